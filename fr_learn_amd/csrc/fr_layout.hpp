// fr_layout.hpp -- compile-time sizes of a tensor-product element and the
// index maps that the kernels use instead of dense operator rows.
//
// Point numbering is the reference's (src/points.cpp:37-128): hex solution
// point (i,j,k) is i + P1*(j + P1*k); hex faces are 0 z=-1, 1 z=+1, 2 x=-1,
// 3 x=+1, 4 y=-1, 5 y=+1 with flux point face*P1^2 + a + b*P1 and the face-wise
// index reversals of getLocFpts; quad faces are 0 y=-1, 1 x=+1, 2 y=+1,
// 3 x=-1 (counter-clockwise).  Because every flux point shares its tangential
// coordinates with a line of solution points, each row of the extrapolation,
// gradient and correction operators is nonzero only on that line; the maps
// below enumerate those entries in ascending column order.  fr_host.cpp
// checks them against the dense operators at context creation.
#pragma once

#include <hip/hip_runtime.h>

// Build-time switches of the stage kernel's layout (compile with -D to change)
#ifndef FR_LDS_PAD
#define FR_LDS_PAD 1  // pad the solution-point rows in LDS (padIdx)
#endif

namespace frdev {

constexpr int nextPow2(int n) {
  int p = 1;
  while (p < n) p *= 2;
  return p;
}

// Threads per linear-triangle element of order P: one solution point
// ((P+1)(P+2)/2) and one flux point (3(P+1)) per thread, rounded up to a power
// of two (8, 16, 16, 16, 32 for P = 1..5); 256 / triTPE(P) elements per block.
constexpr int triTPE(int P) {
  const int ns = (P + 1) * (P + 2) / 2, nf = 3 * (P + 1);
  return nextPow2(ns > nf ? ns : nf);
}

// Threads per triangle element in tile layout L (as quadTPE).
constexpr int triTPEL(int P, int L) {
  const int ns = (P + 1) * (P + 2) / 2, nf = 3 * (P + 1);
  return L == 1 ? (ns > nf ? ns : nf) : triTPE(P);
}

// Threads per element of a tensor-product element in tile layout L:
// 0 (every per-mesh kernel): the solution-point count rounded up to a power of
// two (at least 8), so that an element's threads never straddle a wavefront;
// 1 (2D, a choice of the batched kernels' tuner): max(solution points, flux
// points), packing e.g. seven P5 quads (36 points) into 252 threads where
// layout 0 gives each 64 threads.
constexpr int tensorTPE(int D, int P, int L) {
  const int P1 = P + 1, ns = (D == 3) ? P1 * P1 * P1 : P1 * P1, nf = (D == 3) ? 6 * P1 * P1 : 4 * P1;
  if (L == 1 && D == 2) return ns > nf ? ns : nf;
  return nextPow2(ns) < 8 ? 8 : nextPow2(ns);
}

template <int D, int P, int L = 0>
struct ElemGeo {
  static constexpr int P1 = P + 1;
  static constexpr int NS = (D == 3) ? P1 * P1 * P1 : P1 * P1;  // solution points
  static constexpr int NFPF = (D == 3) ? P1 * P1 : P1;          // flux points per face
  static constexpr int NFACE = 2 * D;
  static constexpr int NFP = NFACE * NFPF;                      // flux points
  static constexpr int TPE = tensorTPE(D, P, L);                // threads per element
  static constexpr int EPB = TPE >= 256 ? 1 : 256 / TPE;        // elements per block
  static constexpr int NTH = TPE * EPB;                         // threads per block (<= 256)
  // LDS rows of solution-point data can be padded by one entry per x-line
  // (FR_LDS_PAD), so that lanes walking y- or x-face lines (strides P+1, 1
  // with bases P+1 apart) fall in different banks
  static constexpr int NSP = FR_LDS_PAD ? NS + NS / P1 : NS;
};

// LDS index of solution point s (padded: one pad entry per x-line).
template <int P>
__host__ __device__ __forceinline__ int padIdx(int s) {
  return FR_LDS_PAD ? s + s / (P + 1) : s;
}

// LDS stride of a solution-point line of stride `stride` (1, P+1 or
// (P+1)^2): every line starts on an x-line boundary or moves along x, so
// padIdx(base + m*stride) = padIdx(base) + m*padStride(stride).
template <int P>
__host__ __device__ __forceinline__ int padStride(int stride) {
  return FR_LDS_PAD ? stride + stride / (P + 1) : stride;
}

// Solution-point line behind flux point q: spts base + m*stride (m = 0..P)
// with 1D weights l_R (side 1) or l_L (side 0); dn = reference normal axis.
template <int D, int P>
__host__ __device__ __forceinline__ void fptLine(int q, int &base, int &stride, int &side, int &dn) {
  // Written with selects, not a switch over the face: the lanes of a
  // wavefront cover several faces, and a switch would run every case.
  //   hex   f: base                  stride  side dn
  //         0: a + P1 b              P1^2    0    2
  //         1: (P-a) + P1 b          P1^2    1    2
  //         2: P1 (a + P1 b)         1       0    0
  //         3: P1 ((P-a) + P1 b)     1       1    0
  //         4: (P-a) + P1^2 b        P1      0    1
  //         5: a + P1^2 b            P1      1    1
  //   quad  f: 0: a, P1, 0, 1;  1: P1 a, 1, 1, 0;  2: P-a, P1, 1, 1;  3: P1 (P-a), 1, 0, 0
  constexpr int P1 = P + 1;
  if (D == 3) {
    constexpr int F2 = P1 * P1;
    const int f = q / F2, r = q - f * F2, a = r % P1, b = r / P1;
    const int aa = (f == 1 || f == 3 || f == 4) ? P - a : a;
    base = ((f == 2 || f == 3) ? P1 : 1) * aa + ((f < 2) ? P1 : F2) * b;
    stride = (f < 2) ? F2 : (f < 4) ? 1 : P1;
    side = f & 1;
    dn = (f < 2) ? 2 : (f < 4) ? 0 : 1;
  } else {
    const int f = q / P1, a = q - f * P1;
    const int aa = (f >= 2) ? P - a : a;
    base = (f & 1) ? P1 * aa : aa;
    stride = (f & 1) ? 1 : P1;
    side = (f == 1 || f == 2) ? 1 : 0;
    dn = (f & 1) ? 0 : 1;
  }
}

// Correction entries of solution point s, one per face in face order:
// flux point q[f] and coefficient = sgn[f] * g_side[f](x_{axis[f]}), where
// g_L / g_R are the derivatives of the left / right correction functions.
template <int D, int P>
__host__ __device__ __forceinline__ void corrEntries(int s, int *q, int *side, int *axisIdx) {
  constexpr int P1 = P + 1;
  if (D == 3) {
    constexpr int F2 = P1 * P1;
    const int k = s / F2, j = (s - F2 * k) / P1, i = s - F2 * k - P1 * j;
    q[0] = i + j * P1;                side[0] = 0; axisIdx[0] = k;
    q[1] = F2 + (P - i) + j * P1;     side[1] = 1; axisIdx[1] = k;
    q[2] = 2 * F2 + j + k * P1;       side[2] = 0; axisIdx[2] = i;
    q[3] = 3 * F2 + (P - j) + k * P1; side[3] = 1; axisIdx[3] = i;
    q[4] = 4 * F2 + (P - i) + k * P1; side[4] = 0; axisIdx[4] = j;
    q[5] = 5 * F2 + i + k * P1;       side[5] = 1; axisIdx[5] = j;
  } else {
    const int i = s % P1, j = s / P1;
    q[0] = i;                side[0] = 0; axisIdx[0] = j;
    q[1] = P1 + j;           side[1] = 1; axisIdx[1] = i;
    q[2] = 2 * P1 + (P - i); side[2] = 1; axisIdx[2] = j;
    q[3] = 3 * P1 + (P - j); side[3] = 0; axisIdx[3] = i;
  }
}

// Per element-face record of the neighbour: an int32 neighbour element and an
// int8 `finfo`: interior faces pack the neighbour's local face (bits 0-2), the
// relative rotation (bits 3-4) and the side (bit 5: 0 = this element is the
// face's left element, 1 = right); boundary faces are 64 + bc id.
constexpr int FINFO_BOUNDARY = 64;

__host__ __device__ __forceinline__ int packFinfo(int nbrFace, int rot, int side) {
  return nbrFace | (rot << 3) | (side << 5);
}

// Neighbour-local flux point that pairs with flux point q of this element.
// The right-element maps of intFace::setupRightState (intFace.cpp:29-65) are
// involutions -- (a,b) -> (b,a), (P-a,b), (P-b,P-a), (a,P-b) -- so the same
// map takes a left point to its right partner and back.
template <int D, int P>
__host__ __device__ __forceinline__ int neighbourFpt(int q, int finfo) {
  constexpr int P1 = P + 1;
  constexpr int NFPF = (D == 3) ? P1 * P1 : P1;
  const int f = q / NFPF, i = q - f * NFPF;
  const int nf = finfo & 7, rot = (finfo >> 3) & 3;
  int j;
  if (D == 3) {
    // rot 0: (b, a); 1: (P-a, b); 2: (P-b, P-a); 3: (a, P-b), as selects (see fptLine)
    const int a = i % P1, b = i / P1, t = a * P1 + b;
    j = (rot == 0) ? t : (rot == 1) ? (P - a) + b * P1 : (rot == 2) ? NFPF - 1 - t : a + P1 * (P - b);
  } else {
    j = P - i;
  }
  return nf * NFPF + j;
}

// Layout of the per-order 1D factor table: l_L[P1], l_R[P1], g_L[P1], g_R[P1],
// Dm[P1][P1] (Dm[i][m] = derivative of basis m at node i), then the 1D
// solution-point abscissae x[P1] (for the metrics of the nodal geometry).
template <int P>
struct OneD {
  static constexpr int P1 = P + 1;
  static constexpr int LL = 0, LR = P1, GL = 2 * P1, GR = 3 * P1, DM = 4 * P1, X1 = 4 * P1 + P1 * P1;
  // the derivative with the correction of the extrapolated flux folded in,
  // DMF[i][m] = DM[i][m] + GL[i] LL[m] - GR[i] LR[m]: along a line, DM F~ plus
  // the correction of the jump (common flux less the extrapolated outward
  // flux) is DMF F~ plus the correction of the common flux alone
  static constexpr int DMF = 5 * P1 + P1 * P1;
  static constexpr int SIZE = 5 * P1 + 2 * P1 * P1;
};

}  // namespace frdev
