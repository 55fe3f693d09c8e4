// fr_host.hpp -- host-side setup of the FR engine: FR operators (dense, for
// export, plus the tensor-sparse tables the kernels use) and face
// connectivity.  Runs once per fr_create; the hot path is on the device.
#pragma once

#include <cstdint>
#include <array>
#include <string>
#include <vector>

#include "../../include/fr.h"

namespace frhost {

// Dense FR operators of one (element type, order) plus their sparse form.
// Dense shapes follow oper::opp_* (reference include/operators.hpp:133-146).
struct Operators {
  int nDims = 3, order = 3, P1 = 4, nSpts = 0, nFpts = 0, nFptsFace = 0, nFacesEle = 0;
  std::vector<double> spts1D, wts;           // 1D abscissae, per-spt quadrature weights
  std::vector<double> wtsFace;               // quadrature weights of one face's flux points
  std::vector<double> locSpts, locFpts;      // [pt][3]
  std::vector<double> tNorm;                 // [fpt][3]
  std::vector<double> spts_to_fpts;          // nFpts x nSpts
  std::vector<double> grad_spts;             // nDims x nSpts x nSpts
  std::vector<double> extrapolateFn;         // nDims x nFpts x nSpts
  std::vector<double> correction;            // nSpts x nFpts
  std::vector<double> correctU;              // nDims x nSpts x nFpts
  // Shape-function derivatives of the linear element at spts / fpts,
  // [dim][pt][node] (reference solver.cpp:153-168, funcs.cpp dshape_hex).
  std::vector<double> dshapeSpts, dshapeFpts;
  std::vector<double> shapeSpts;  // [pt][node]
  int nNodes = 8;
  // 1D factors consumed by the kernels (layout of frdev::OneD): l_L, l_R,
  // correction coefficients of left faces (-g_L') and right faces (g_R'), and
  // the nodal derivative matrix Dm[i][m].
  std::vector<double> oned;
};

// Builds everything above; returns "" or an error message.
std::string buildOperators(const fr_config &cfg, Operators &ops, int nNodes = 0);

struct FaceRec {
  int eL, fL, eR, fR, rot, bc;  // bc = -1 for interior faces
};

struct Connectivity {
  std::vector<FaceRec> faces;   // interior faces first, then boundary faces
  int nIntFaces = 0;
  std::vector<int> fptR;        // [nFaces][nFptsFace] right-element local fpt per left fpt
  // Element-centric tables:
  std::vector<int32_t> nbr;     // [nEles][nFpts]: paired (ele*nFpts + fpt) or -1 (setup only)
  std::vector<int32_t> fnbr;    // [nEles][nFacesEle]: neighbour element or -1   (kernels)
  std::vector<int8_t> finfo;    // [nEles][nFacesEle]: packed face record, see fr_layout.hpp
};

std::string buildConnectivity(const fr_config &cfg, const fr_mesh &mesh, const Operators &ops,
                              Connectivity &conn);

// Device element order (device index -> mesh element): consecutive blocks of
// `epb` elements are face-connected clusters, grown greedily from seeds taken
// in Z-order of the element centroids ([nE][3]), so that the faces inside a
// block are exchanged through LDS and the blocks follow a space-filling curve.
std::vector<int> blockOrder(const Connectivity &conn, const std::vector<double> &centroid, int nE, int nFaceEle,
                            int epb);

// 1D solution-point abscissae (getPts1D, points.cpp:162-330): sptsType 0
// Gauss-Legendre (order 0..10), 1 Gauss-Lobatto (order 0..6).  Empty for
// other orders.
std::vector<double> pts1D(int sptsType, int order);

// 1D Gauss-Legendre weights (getQptWeights1D, points.cpp:376-469), order
// 0..10; empty for other orders.
std::vector<double> gaussWeights1D(int order);

// Lagrange basis l_mode(y) on the nodes x (polynomials.cpp:9-19).
double lagrange(const std::vector<double> &x, double y, int mode);

// Shape functions of the element's nodes at the reference point r[3] (and,
// with dout, their derivatives dout[node][3]): the 4-node quad, the hex of
// 8 (default), 20 (quadratic serendipity), 27, 64, 125 or 216 (Lagrange)
// nodes in the reference's node order (shape_quad / shape_hex and
// dshape_quad / dshape_hex, funcs.cpp:121-1607), or with nNodes 3 the linear
// triangle (-1,-1), (1,-1), (-1,1).
void shapeFunctions(int nDims, const double *r, double *out, int nNodes = 0, double *dout = nullptr);

// Lattice indices (i, j, k) on the nSide^3 equispaced nodes of each node of
// the Lagrange hex, in the reference's node order: nested shells from the
// outside in, each with its 8 corners, its 12 edges and its 6 faces (bottom,
// front, left, right, back, top; each face by rings of corners and edges,
// then its centre), and the volume centre last (shape_hex, funcs.cpp:213-566).
std::vector<std::array<int, 3>> hexLattice(int nSide);

// Node counts of the supported hexes: 8, 20, 27, 64, 125, 216.
bool hexNodeCount(int nNodes);

// FR operators of the linear triangle (fr_tri_ops.cpp); "" or an error message.
std::string buildTriOperators(const fr_config &cfg, Operators &ops);

}  // namespace frhost
