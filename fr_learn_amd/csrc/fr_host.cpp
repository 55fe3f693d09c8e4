// fr_host.cpp -- FR operator construction and face connectivity (host side).
//
// Operators: the dense matrices are produced from 1D Lagrange / correction
// factors exactly as the reference's oper::setupOperators does
// (src/operators.cpp:13-111, 660-1092, 1544-1619; polynomials from
// src/polynomials.cpp:9-118, 392-475), so every entry has the same value.  The
// kernels never apply them densely: each row holds at most P+1 (extrapolation,
// gradient) or 2*nDims (correction) nonzeros.  The kernels enumerate them with
// the index maps of fr_layout.hpp, in ascending column order so that a sparse
// row sum adds the same terms in the same order as the dense product, and
// verifyIndexMaps() checks those maps against the dense matrices here.
//
// Connectivity: faces are matched on sorted canonical vertex lists (the
// reference does the same with quadratic scans, src/geo.cpp:75-267); the left
// element of an interior face is its lowest-index owner (geo.cpp:222-233), the
// relative rotation follows geo::compareOrientation (geo.cpp:789-917) and the
// right flux-point map follows intFace::setupRightState (intFace.cpp:29-65).
#include "fr_host.hpp"

#include "fr_layout.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <unordered_set>

namespace frhost {

namespace {

// 1D abscissae / weights.  The values are the reference's 15-digit tables
// (src/points.cpp:169-253 Gauss-Legendre, 254-330 Gauss-Lobatto, 376-469
// weights), kept digit for digit because operator parity depends on them.
const double kGL[11][11] = {
    {0.0},
    {-0.577350269189626, 0.577350269189626},
    {-0.774596669241483, 0.000000000000000, 0.774596669241483},
    {-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053},
    {-0.906179845938664, -0.538469310105683, 0.000000000000000, 0.538469310105683, 0.906179845938664},
    {-0.932469514203152, -0.661209386466264, -0.238619186083197, 0.238619186083197, 0.661209386466265,
     0.932469514203152},
    {-0.949107912342758, -0.741531185599395, -0.405845151377397, 0.000000000000000, 0.405845151377397,
     0.741531185599394, 0.949107912342758},
    {-0.960289856497536, -0.796666477413627, -0.525532409916329, -0.183434642495650, 0.183434642495650,
     0.525532409916329, 0.796666477413627, 0.960289856497536},
    {-0.968160239507626, -0.836031107326636, -0.613371432700591, -0.324253423403809, 0.000000000000000,
     0.324253423403809, 0.613371432700591, 0.836031107326636, 0.968160239507626},
    {-0.973906528517172, -0.865063366688985, -0.679409568299024, -0.433395394129247, -0.148874338981631,
     0.148874338981632, 0.433395394129247, 0.679409568299024, 0.865063366688984, 0.973906528517171},
    {-0.978228658146057, -0.887062599768095, -0.730152005574049, -0.519096129206812, -0.269543155952345,
     0.000000000000000, 0.269543155952345, 0.519096129206812, 0.730152005574049, 0.887062599768095,
     0.978228658146057}};
const double kGLL[7][7] = {
    {0.0},
    {-1.000000000000000, 1.000000000000000},
    {-1.000000000000000, -0.000000000000000, 1.000000000000000},
    {-1.000000000000000, -0.447213595499958, 0.447213595499958, 1.000000000000000},
    {-1.000000000000000, -0.654653670707977, -0.000000000000000, 0.654653670707977, 1.000000000000000},
    {-1.000000000000000, -0.765055323929465, -0.285231516480645, 0.285231516480645, 0.765055323929465,
     1.000000000000000},
    {-1.000000000000000, -0.830223896278567, -0.468848793470714, 0.000000000000000, 0.468848793470714,
     0.830223896278567, 1.000000000000000}};
const double kGW[11][11] = {
    {2.0},
    {1.0, 1.0},
    {0.5555555555555556, 0.8888888888888888, 0.5555555555555556},
    {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538},
    {0.2369268850561891, 0.4786286704993665, 0.5688888888888889, 0.4786286704993665, 0.2369268850561891},
    {0.1713244923791704, 0.3607615730481386, 0.4679139345726910, 0.4679139345726910, 0.3607615730481386,
     0.1713244923791704},
    {0.1294849661688697, 0.2797053914892766, 0.3818300505051189, 0.4179591836734694, 0.3818300505051189,
     0.2797053914892766, 0.1294849661688697},
    {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620, 0.3626837833783620,
     0.3137066458778873, 0.2223810344533745, 0.1012285362903763},
    {0.0812743883615744, 0.1806481606948574, 0.2606106964029354, 0.3123470770400029, 0.3302393550012598,
     0.3123470770400029, 0.2606106964029354, 0.1806481606948574, 0.0812743883615744},
    {0.0666713443086881, 0.1494513491505806, 0.2190863625159820, 0.2692667193099963, 0.2955242247147529,
     0.2955242247147529, 0.2692667193099963, 0.2190863625159820, 0.1494513491505806, 0.0666713443086881},
    {0.0556685671161737, 0.1255803694649046, 0.1862902109277343, 0.2331937645919905, 0.2628045445102467,
     0.2729250867779006, 0.2628045445102467, 0.2331937645919905, 0.1862902109277343, 0.1255803694649046,
     0.0556685671161737}};

// Lagrange basis l_mode(y) on nodes x (product form of polynomials.cpp:9-19).
double lag(const std::vector<double> &x, double y, int mode) {
  double v = 1.0;
  for (size_t i = 0; i < x.size(); i++)
    if ((int)i != mode) v = v * ((y - x[i]) / (x[mode] - x[i]));
  return v;
}
// d/dy l_mode(y): sum over dropped factor (polynomials.cpp:21-47).
double dlag(const std::vector<double> &x, double y, int mode) {
  double s = 0.0;
  int n = (int)x.size();
  for (int i = 0; i < n; i++) {
    if (i == mode) continue;
    double num = 1.0, den = 1.0;
    for (int j = 0; j < n; j++) {
      if (j != mode && j != i) num = num * (y - x[j]);
      if (j != mode) den = den * (x[mode] - x[j]);
    }
    s = s + num / den;
  }
  return s;
}
// Legendre P_n by the three-term recursion (polynomials.cpp:84-95).
double leg(double r, int n) {
  if (n < 0) return 0.;
  if (n == 0) return 1.0;
  if (n == 1) return r;
  return ((2 * n - 1) * r * leg(r, n - 1) - (n - 1) * leg(r, n - 2)) / n;
}
// P_n'(r) from the identity (r^2-1) P_n' = n (r P_n - P_{n-1}); endpoint
// closed forms at r = +-1 (polynomials.cpp:97-118).
double dleg(double r, int n) {
  if (n == 0) return 0;
  if (r > -1.0 && r < 1.0) return n * ((r * leg(r, n)) - leg(r, n - 1)) / (r * r - 1.0);
  double d = 0.;
  if (r == -1.0) d = std::pow(-1.0, n - 1.0) * 0.5 * n * (n + 1.0);
  if (r == 1.0) d = 0.5 * n * (n + 1.0);
  return d;
}
int fact(int n) { return (n <= 1) ? 1 : fact(n - 1) * n; }
// VCJH parameter eta for DG / SD / HU / C+ (polynomials.cpp:392-428).
double vcjhEta(int scheme, int P) {
  if (scheme == 0) return 0.0;
  if (scheme == 1) return (1.0 * P) / (1.0 * (P + 1));
  if (scheme == 2) return (1.0 * (P + 1)) / (1.0 * P);
  double c = 0;
  if (P == 2) c = 0.206;
  else if (P == 3) c = 3.80e-3;
  else if (P == 4) c = 4.67e-5;
  else if (P == 5) c = 4.28e-7;
  double ap = 1. / std::pow(2.0, P) * fact(2 * P) / (fact(P) * fact(P));
  return c * (2 * P + 1) / 2 * (fact(P) * ap) * (fact(P) * ap);
}
// Derivative of the left (side 0) / right (side 1) VCJH correction function
// (polynomials.cpp:447-475).
double dvcjh(double r, int side, int P, double eta) {
  double sgn = std::pow(-1.0, P);
  if (side == 0) {
    if (P == 0) return 0.5 * sgn * (dleg(r, P) - (dleg(r, P + 1) / (1.0 + eta)));
    return 0.5 * sgn * (dleg(r, P) - ((eta * dleg(r, P - 1) + dleg(r, P + 1)) / (1.0 + eta)));
  }
  if (P == 0) return 0.5 * (dleg(r, P) + (dleg(r, P + 1) / (1.0 + eta)));
  return 0.5 * (dleg(r, P) + ((eta * dleg(r, P - 1) + dleg(r, P + 1)) / (1.0 + eta)));
}

void faceNormal(int nDims, int face, double *t) {
  static const double H[6][3] = {{0, 0, -1}, {0, 0, 1}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}};
  static const double Q[4][3] = {{0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}};
  for (int d = 0; d < 3; d++) t[d] = (nDims == 3) ? H[face][d] : Q[face][d];
}


// Compare the kernels' sparse index maps against the dense operators.
template <int D, int P>
std::string verifyMapsT(const Operators &o) {
  using frdev::OneD;
  const int P1 = P + 1, nS = o.nSpts, nF = o.nFpts;
  // extrapolation and reference-normal flux rows
  for (int q = 0; q < nF; q++) {
    int base, stride, side, dn;
    frdev::fptLine<D, P>(q, base, stride, side, dn);
    std::vector<double> row(nS, 0.0);
    if (base < 0 || base + P * stride >= nS || dn < 0 || dn >= D) return "extrapolation index map out of range";
    for (int m = 0; m < P1; m++) row[base + m * stride] = o.oned[(side ? OneD<P>::LR : OneD<P>::LL) + m];
    for (int s = 0; s < nS; s++) {
      if (row[s] != o.spts_to_fpts[(size_t)q * nS + s]) return "extrapolation index map mismatch";
      for (int d = 0; d < D; d++) {
        const double want = (d == dn) ? row[s] * (side ? 1.0 : -1.0) : 0.0;
        if (want != o.extrapolateFn[((size_t)d * nF + q) * nS + s]) return "normal-flux index map mismatch";
      }
    }
  }
  // gradient rows
  for (int s = 0; s < nS; s++) {
    int ijk[3] = {s % P1, (s / P1) % P1, (D == 3) ? s / (P1 * P1) : 0};
    for (int d = 0; d < D; d++) {
      std::vector<double> row(nS, 0.0);
      for (int m = 0; m < P1; m++) {
        int c[3] = {ijk[0], ijk[1], ijk[2]};
        c[d] = m;
        const int col = c[0] + P1 * (c[1] + P1 * c[2]);
        row[col] = o.oned[OneD<P>::DM + ijk[d] * P1 + m];
      }
      for (int b = 0; b < nS; b++)
        if (row[b] != o.grad_spts[((size_t)d * nS + s) * nS + b]) return "gradient index map mismatch";
    }
  }
  // correction rows
  for (int s = 0; s < nS; s++) {
    int q[6], side[6], ax[6];
    frdev::corrEntries<D, P>(s, q, side, ax);
    std::vector<double> row(nF, 0.0);
    for (int f = 0; f < 2 * D; f++) {
      if (q[f] < 0 || q[f] >= nF || ax[f] < 0 || ax[f] > P) return "correction index map out of range";
      row[q[f]] = o.oned[(side[f] ? OneD<P>::GR : OneD<P>::GL) + ax[f]];
    }
    for (int c = 0; c < nF; c++)
      if (row[c] != o.correction[(size_t)s * nF + c]) return "correction index map mismatch";
  }
  return "";
}

// The kernels derive each flux point's partner from the per-face record
// (neighbourFpt); check that against the explicit pairing built from
// intFace::setupRightState's maps.
template <int D, int P>
std::string verifyNeighbourT(const Operators &o, const Connectivity &c, int nE) {
  const int nF = o.nFpts, nfpf = o.nFptsFace, nfe = o.nFacesEle;
  for (int e = 0; e < nE; e++)
    for (int q = 0; q < nF; q++) {
      const int info = c.finfo[(size_t)e * nfe + q / nfpf];
      const int32_t want = c.nbr[(size_t)e * nF + q];
      if (info >= frdev::FINFO_BOUNDARY) {
        if (want != -1) return "boundary face with a paired flux point";
        continue;
      }
      const int32_t got = c.fnbr[(size_t)e * nfe + q / nfpf] * nF + frdev::neighbourFpt<D, P>(q, info);
      if (got != want) return "neighbour flux-point map mismatch";
    }
  return "";
}

std::string verifyNeighbourMaps(const Operators &o, const Connectivity &c, int nE) {
  switch (o.nDims * 10 + o.order) {
    case 31: return verifyNeighbourT<3, 1>(o, c, nE);
    case 32: return verifyNeighbourT<3, 2>(o, c, nE);
    case 33: return verifyNeighbourT<3, 3>(o, c, nE);
    case 34: return verifyNeighbourT<3, 4>(o, c, nE);
    case 35: return verifyNeighbourT<3, 5>(o, c, nE);
    case 21: return verifyNeighbourT<2, 1>(o, c, nE);
    case 22: return verifyNeighbourT<2, 2>(o, c, nE);
    case 23: return verifyNeighbourT<2, 3>(o, c, nE);
    case 24: return verifyNeighbourT<2, 4>(o, c, nE);
    case 25: return verifyNeighbourT<2, 5>(o, c, nE);
  }
  return "unsupported element order";
}

std::string verifyIndexMaps(const Operators &o) {
  const int key = o.nDims * 10 + o.order;
  switch (key) {
    case 31: return verifyMapsT<3, 1>(o);
    case 32: return verifyMapsT<3, 2>(o);
    case 33: return verifyMapsT<3, 3>(o);
    case 34: return verifyMapsT<3, 4>(o);
    case 35: return verifyMapsT<3, 5>(o);
    case 21: return verifyMapsT<2, 1>(o);
    case 22: return verifyMapsT<2, 2>(o);
    case 23: return verifyMapsT<2, 3>(o);
    case 24: return verifyMapsT<2, 4>(o);
    case 25: return verifyMapsT<2, 5>(o);
  }
  return "unsupported element order";
}

}  // namespace

std::string buildOperators(const fr_config &cfg, Operators &o, int nNodes) {
  const int D = cfg.nDims, P = cfg.order;
  if (D != 2 && D != 3) return "nDims must be 2 or 3";
  if (P < 1 || P > 5) return "order must be in 1..5";
  if (cfg.sptsType != 0 && cfg.sptsType != 1) return "sptsType must be 0 (Legendre) or 1 (Lobatto)";
  if (cfg.vcjhScheme < 0 || cfg.vcjhScheme > 3) return "Invalid VCJH scheme.";
  if (cfg.vcjhScheme == 3 && P < 2) return "C_plus scheme not implemented for this order";
  o.nDims = D;
  o.order = P;
  const int P1 = o.P1 = P + 1;
  o.nSpts = (D == 3) ? P1 * P1 * P1 : P1 * P1;
  o.nFptsFace = (D == 3) ? P1 * P1 : P1;
  o.nFacesEle = 2 * D;
  o.nFpts = o.nFacesEle * o.nFptsFace;
  o.nNodes = (D == 3) ? (nNodes ? nNodes : 8) : 4;
  if (D == 3 && !hexNodeCount(o.nNodes)) return "hexes have 8, 20, 27, 64, 125 or 216 nodes";
  const int nS = o.nSpts, nF = o.nFpts;
  std::vector<double> x(P1);
  for (int i = 0; i < P1; i++) x[i] = (cfg.sptsType == 0) ? kGL[P][i] : kGLL[P][i];
  o.spts1D = x;

  // point sets (points.cpp:37-128) and spt quadrature weights (points.cpp:347-374)
  o.locSpts.assign(3 * nS, 0.0);
  o.locFpts.assign(3 * nF, 0.0);
  o.wts.assign(nS, 0.0);
  for (int s = 0; s < nS; s++) {
    int i = s % P1, j = (s / P1) % P1, k = s / (P1 * P1);
    o.locSpts[3 * s + 0] = x[i];
    o.locSpts[3 * s + 1] = x[j];
    o.locSpts[3 * s + 2] = (D == 3) ? x[k] : 0.0;
    o.wts[s] = (D == 3) ? kGW[P][k] * kGW[P][j] * kGW[P][i] : kGW[P][j] * kGW[P][i];
  }
  // face weights indexed by face-local flux point, getQptWeights(order,
  // nDims - 1) as ele::calcDt uses them (ele.cpp:1087, 1100)
  const int nFF = (D == 3) ? P1 * P1 : P1;
  o.wtsFace.assign(nFF, 0.0);
  for (int m = 0; m < nFF; m++) o.wtsFace[m] = (D == 3) ? kGW[P][m / P1] * kGW[P][m % P1] : kGW[P][m];
  if (D == 3) {
    const int F2 = P1 * P1;
    for (int i = 0; i < P1; i++)
      for (int j = 0; j < P1; j++) {
        const int q = i + j * P1;
        const double pts[6][3] = {{x[i], x[j], -1.},     {x[P - i], x[j], 1.},     {-1., x[i], x[j]},
                                  {1., x[P - i], x[j]}, {x[P - i], -1., x[j]}, {x[i], 1., x[j]}};
        for (int f = 0; f < 6; f++)
          for (int d = 0; d < 3; d++) o.locFpts[3 * (q + f * F2) + d] = pts[f][d];
      }
  } else {
    for (int i = 0; i < P1; i++) {
      const double pts[4][2] = {{x[i], -1.}, {1., x[i]}, {x[P - i], 1.}, {-1., x[P - i]}};
      for (int f = 0; f < 4; f++)
        for (int d = 0; d < 2; d++) o.locFpts[3 * (i + f * P1) + d] = pts[f][d];
    }
  }
  o.tNorm.assign(3 * nF, 0.0);
  for (int q = 0; q < nF; q++) faceNormal(D, q / o.nFptsFace, &o.tNorm[3 * q]);

  auto tripleIdx = [&](int s, int &i, int &j, int &k) {
    if (D == 3) {
      k = s / (P1 * P1);
      j = (s - P1 * P1 * k) / P1;
      i = s - P1 * j - P1 * P1 * k;
    } else {
      i = s % P1;
      j = s / P1;
      k = 0;
    }
  };
  // extrapolation to flux points (operators.cpp:78-111)
  o.spts_to_fpts.assign((size_t)nF * nS, 0.0);
  for (int q = 0; q < nF; q++) {
    const double *f = &o.locFpts[3 * q];
    for (int s = 0; s < nS; s++) {
      int i, j, k;
      tripleIdx(s, i, j, k);
      double v = lag(x, f[0], i) * lag(x, f[1], j);
      if (D == 3) v = v * lag(x, f[2], k);
      o.spts_to_fpts[(size_t)q * nS + s] = v;
    }
  }
  // gradient at spts (operators.cpp:660-740)
  o.grad_spts.assign((size_t)D * nS * nS, 0.0);
  for (int d = 0; d < D; d++)
    for (int g = 0; g < nS; g++) {
      int ig, jg, kg;
      tripleIdx(g, ig, jg, kg);
      for (int b = 0; b < nS; b++) {
        int ib, jb, kb;
        tripleIdx(b, ib, jb, kb);
        double v;
        if (D == 3) {
          if (d == 0) v = dlag(x, x[ig], ib) * lag(x, x[jg], jb) * lag(x, x[kg], kb);
          else if (d == 1) v = dlag(x, x[jg], jb) * lag(x, x[ig], ib) * lag(x, x[kg], kb);
          else v = dlag(x, x[kg], kb) * lag(x, x[ig], ib) * lag(x, x[jg], jb);
        } else {
          if (d == 0) v = dlag(x, x[ig], ib) * lag(x, x[jg], jb);
          else v = dlag(x, x[jg], jb) * lag(x, x[ig], ib);
        }
        o.grad_spts[((size_t)d * nS + g) * nS + b] = v;
      }
    }
  // reference-normal flux extrapolation (operators.cpp:807-889)
  o.extrapolateFn.assign((size_t)D * nF * nS, 0.0);
  for (int d = 0; d < D; d++)
    for (int q = 0; q < nF; q++)
      for (int s = 0; s < nS; s++)
        o.extrapolateFn[((size_t)d * nF + q) * nS + s] = o.spts_to_fpts[(size_t)q * nS + s] * o.tNorm[3 * q + d];
  // correction = divergence of the VCJH correction field (operators.cpp:891-914,
  // divVCJH_hex 1575-1619, divVCJH_quad 1544-1573)
  const double eta = (cfg.vcjhScheme == 0) ? 0. : vcjhEta(cfg.vcjhScheme, P);
  o.correction.assign((size_t)nS * nF, 0.0);
  for (int s = 0; s < nS; s++) {
    const double *l = &o.locSpts[3 * s];
    for (int q = 0; q < nF; q++) {
      double v;
      if (D == 3) {
        const int F2 = P1 * P1, f = q / F2, j = (q - F2 * f) / P1, i = q - F2 * f - P1 * j;
        switch (f) {
          case 0: v = lag(x, l[0], i) * lag(x, l[1], j) * -dvcjh(l[2], 0, P, eta); break;
          case 1: v = lag(x, l[0], P - i) * lag(x, l[1], j) * dvcjh(l[2], 1, P, eta); break;
          case 2: v = lag(x, l[1], i) * lag(x, l[2], j) * -dvcjh(l[0], 0, P, eta); break;
          case 3: v = lag(x, l[1], P - i) * lag(x, l[2], j) * dvcjh(l[0], 1, P, eta); break;
          case 4: v = lag(x, l[0], P - i) * lag(x, l[2], j) * -dvcjh(l[1], 0, P, eta); break;
          default: v = lag(x, l[0], i) * lag(x, l[2], j) * dvcjh(l[1], 1, P, eta); break;
        }
      } else {
        const int f = q / P1, j = q - P1 * f;
        switch (f) {
          case 0: v = -lag(x, l[0], j) * dvcjh(l[1], 0, P, eta); break;
          case 1: v = lag(x, l[1], j) * dvcjh(l[0], 1, P, eta); break;
          case 2: v = lag(x, l[0], P - j) * dvcjh(l[1], 1, P, eta); break;
          default: v = -lag(x, l[1], P - j) * dvcjh(l[0], 0, P, eta); break;
        }
      }
      o.correction[(size_t)s * nF + q] = v;
    }
  }
  // gradient correction (operators.cpp:1008-1092)
  o.correctU.assign((size_t)D * nS * nF, 0.0);
  for (int d = 0; d < D; d++)
    for (int q = 0; q < nF; q++)
      for (int s = 0; s < nS; s++)
        o.correctU[((size_t)d * nS + s) * nF + q] = o.correction[(size_t)s * nF + q] * o.tNorm[3 * q + d];

  // 1D factors for the kernels and a check that the kernels' index maps
  // (fr_layout.hpp) enumerate exactly the nonzeros of the dense operators
  {
    const int L = 5 * P1 + 2 * P1 * P1;
    o.oned.assign(L, 0.0);
    for (int m = 0; m < P1; m++) {
      o.oned[4 * P1 + P1 * P1 + m] = x[m];
      o.oned[m] = lag(x, -1.0, m);
      o.oned[P1 + m] = lag(x, 1.0, m);
      o.oned[2 * P1 + m] = -dvcjh(x[m], 0, P, eta);
      o.oned[3 * P1 + m] = dvcjh(x[m], 1, P, eta);
      for (int i = 0; i < P1; i++) o.oned[4 * P1 + i * P1 + m] = dlag(x, x[i], m);
    }
    // folded derivative (OneD::DMF)
    for (int i = 0; i < P1; i++)
      for (int m = 0; m < P1; m++)
        o.oned[5 * P1 + P1 * P1 + i * P1 + m] = o.oned[4 * P1 + i * P1 + m] + o.oned[2 * P1 + i] * o.oned[m] -
                                                o.oned[3 * P1 + i] * o.oned[P1 + m];
    std::string err = verifyIndexMaps(o);
    if (!err.empty()) return err;
  }

  // shape-function derivatives of the element's nodes at the solution and
  // flux points (dshape_quad / dshape_hex, funcs.cpp:570-1607)
  const int nN = o.nNodes;
  auto fillDshape = [&](const std::vector<double> &loc, int nPts, std::vector<double> &out) {
    out.assign((size_t)D * nPts * nN, 0.0);
    std::vector<double> sh(nN), dsh((size_t)3 * nN);
    for (int p = 0; p < nPts; p++) {
      shapeFunctions(D, &loc[3 * p], sh.data(), nN, dsh.data());
      for (int n = 0; n < nN; n++)
        for (int d = 0; d < D; d++) out[((size_t)d * nPts + p) * nN + n] = dsh[3 * n + d];
    }
  };
  fillDshape(o.locSpts, nS, o.dshapeSpts);
  o.shapeSpts.assign((size_t)nS * nN, 0.0);
  for (int p = 0; p < nS; p++) shapeFunctions(D, &o.locSpts[3 * p], &o.shapeSpts[(size_t)p * nN], nN);
  fillDshape(o.locFpts, nF, o.dshapeFpts);
  return "";
}

std::vector<double> pts1D(int sptsType, int order) {
  if (order < 0 || (sptsType != 0 && sptsType != 1)) return {};
  if ((sptsType == 0 && order > 10) || (sptsType == 1 && order > 6)) return {};
  std::vector<double> x(order + 1);
  for (int i = 0; i <= order; i++) x[i] = (sptsType == 0) ? kGL[order][i] : kGLL[order][i];
  return x;
}

std::vector<double> gaussWeights1D(int order) {
  if (order < 0 || order > 10) return {};
  return std::vector<double>(kGW[order], kGW[order] + order + 1);
}

double lagrange(const std::vector<double> &x, double y, int mode) { return lag(x, y, mode); }

std::vector<std::array<int, 3>> hexLattice(int nSide) {
  std::vector<std::array<int, 3>> nd;
  const int nLevels = nSide / 2, c = nSide / 2;
  for (int i = 0; i < nLevels; i++) {
    const int i2 = (nSide - 1) - i;
    // the shell's corners
    const int cs[8][3] = {{i, i, i}, {i2, i, i}, {i2, i2, i}, {i, i2, i}, {i, i, i2}, {i2, i, i2}, {i2, i2, i2}, {i, i2, i2}};
    for (auto &q : cs) nd.push_back({q[0], q[1], q[2]});
    // its edges: slot s of 12 holds the nSide2 points of one edge
    const int nSide2 = nSide - 2 * (i + 1);
    std::vector<std::array<int, 3>> ed((size_t)12 * nSide2);
    for (int j = 0; j < nSide2; j++) {
      const int a = i + 1 + j, b = i2 - 1 - j;
      ed[0 * nSide2 + j] = {a, i, i};
      ed[1 * nSide2 + j] = {i, a, i};
      ed[2 * nSide2 + j] = {i, i, a};
      ed[3 * nSide2 + j] = {i2, a, i};
      ed[4 * nSide2 + j] = {i2, i, a};
      ed[5 * nSide2 + j] = {b, i2, i};
      ed[6 * nSide2 + j] = {i2, i2, a};
      ed[7 * nSide2 + j] = {i, i2, a};
      ed[8 * nSide2 + j] = {a, i, i2};
      ed[9 * nSide2 + j] = {i, a, i2};
      ed[10 * nSide2 + j] = {i2, a, i2};
      ed[11 * nSide2 + j] = {b, i2, i2};
    }
    nd.insert(nd.end(), ed.begin(), ed.end());
    // its faces (bottom, front, left, right, back, top), each by rings of 4
    // corners and 4 edges, then the face centre when nSide2 is odd.  f(u, v)
    // places the face's in-plane indices (u, v) at the face's fixed index.
    const int nLevels2 = nSide2 / 2;
    auto face = [&](auto f, const int (*crn)[2], const int (*edge)[4]) {
      for (int j0 = 0; j0 < nLevels2; j0++) {
        const int j = j0 + i + 1, j2 = i + 1 + (nSide2 - 1) - j0;
        const int val[2] = {j, j2};
        for (int q = 0; q < 4; q++) nd.push_back(f(val[crn[q][0]], val[crn[q][1]]));
        const int nSide3 = nSide2 - 2 * (j0 + 1);
        std::vector<std::array<int, 3>> fe((size_t)4 * nSide3);
        for (int k = 0; k < nSide3; k++) {
          // per edge: (u, v) = (A + sa*k, B + sb*k) with A, B from {j, j2, j+1, j2-1}
          const int opts[4] = {j, j2, j + 1 + k, j2 - 1 - k};
          for (int q = 0; q < 4; q++) fe[(size_t)q * nSide3 + k] = f(opts[edge[q][0]], opts[edge[q][1]]);
        }
        nd.insert(nd.end(), fe.begin(), fe.end());
      }
      if (nSide2 % 2) nd.push_back(f(c, c));
    };
    // corner orders as indices into (j, j2); edge orders as indices into
    // (j, j2, j+1+k, j2-1-k), both as (u, v) -- funcs.cpp:292-556
    static const int cBot[4][2] = {{0, 0}, {0, 1}, {1, 1}, {1, 0}}, eBot[4][4] = {{0, 2}, {2, 1}, {1, 3}, {3, 0}};
    static const int cFro[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}}, eFro[4][4] = {{2, 0}, {1, 2}, {3, 1}, {0, 3}};
    static const int cLef[4][2] = {{0, 0}, {0, 1}, {1, 1}, {1, 0}}, eLef[4][4] = {{0, 2}, {2, 1}, {1, 3}, {3, 0}};
    static const int cRig[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}}, eRig[4][4] = {{2, 0}, {1, 2}, {3, 1}, {0, 3}};
    static const int cBac[4][2] = {{1, 0}, {0, 0}, {0, 1}, {1, 1}}, eBac[4][4] = {{3, 0}, {0, 2}, {2, 1}, {1, 3}};
    static const int cTop[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}}, eTop[4][4] = {{2, 0}, {1, 2}, {3, 1}, {0, 3}};
    face([&](int u, int v) { return std::array<int, 3>{u, v, i}; }, cBot, eBot);    // (xi, eta) at mu = i
    face([&](int u, int v) { return std::array<int, 3>{u, i, v}; }, cFro, eFro);    // (xi, mu) at eta = i
    face([&](int u, int v) { return std::array<int, 3>{i, u, v}; }, cLef, eLef);    // (eta, mu) at xi = i
    face([&](int u, int v) { return std::array<int, 3>{i2, u, v}; }, cRig, eRig);   // (eta, mu) at xi = i2
    face([&](int u, int v) { return std::array<int, 3>{u, i2, v}; }, cBac, eBac);   // (xi, mu) at eta = i2
    face([&](int u, int v) { return std::array<int, 3>{u, v, i2}; }, cTop, eTop);   // (xi, eta) at mu = i2
  }
  if (nSide % 2) nd.push_back({c, c, c});
  return nd;
}

void shapeFunctions(int nDims, const double *r, double *out, int nNodes, double *dout) {
  if (nNodes == 3) {
    out[0] = -(r[0] + r[1]) / 2;
    out[1] = (1 + r[0]) / 2;
    out[2] = (1 + r[1]) / 2;
    if (dout) {
      const double d[3][3] = {{-0.5, -0.5, 0}, {0.5, 0, 0}, {0, 0.5, 0}};
      for (int n = 0; n < 3; n++)
        for (int k = 0; k < 3; k++) dout[3 * n + k] = d[n][k];
    }
    return;
  }
  if (nDims == 2) nNodes = 4;
  if (nDims == 3 && nNodes == 0) nNodes = 8;
  if (nDims == 3 && nNodes == 20) {
    // quadratic serendipity hex (shape_hex / dshape_hex, funcs.cpp:188-212, 685-750)
    static const double XI[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, ETA[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                        MU[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    const double xi = r[0], eta = r[1], mu = r[2];
    for (int i = 0; i < 8; i++) {
      const double a = 1 + xi * XI[i], b = 1 + eta * ETA[i], cc = 1 + mu * MU[i];
      out[i] = .125 * a * b * cc * (xi * XI[i] + eta * ETA[i] + mu * MU[i] - 2);
      if (dout) {
        dout[3 * i + 0] = .125 * XI[i] * b * cc * (2 * xi * XI[i] + eta * ETA[i] + mu * MU[i] - 1);
        dout[3 * i + 1] = .125 * ETA[i] * a * cc * (xi * XI[i] + 2 * eta * ETA[i] + mu * MU[i] - 1);
        dout[3 * i + 2] = .125 * MU[i] * a * b * (xi * XI[i] + eta * ETA[i] + 2 * mu * MU[i] - 1);
      }
    }
    // mid-edge nodes: (node, axis along the edge, signs of the two others)
    static const int EDG[12][4] = {{8, 0, -1, -1}, {10, 0, 1, -1}, {16, 0, -1, 1}, {18, 0, 1, 1},
                                   {9, 1, 1, -1},  {11, 1, -1, -1}, {17, 1, 1, 1}, {19, 1, -1, 1},
                                   {12, 2, -1, -1}, {13, 2, 1, -1}, {14, 2, 1, 1}, {15, 2, -1, 1}};
    for (auto &e : EDG) {
      const int n = e[0], ax = e[1];
      const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;  // the other two axes, ascending
      const double s = r[ax], p = 1 + e[2] * r[o1], q = 1 + e[3] * r[o2];
      out[n] = .25 * (1 - s * s) * p * q;
      if (dout) {
        dout[3 * n + ax] = -.5 * s * p * q;
        dout[3 * n + o1] = .25 * (1 - s * s) * e[2] * q;
        dout[3 * n + o2] = .25 * (1 - s * s) * p * e[3];
      }
    }
    return;
  }
  // Lagrange elements on equispaced nodes: the node order of the quad
  // (Gmsh corners) or of the hex recursion (hexLattice)
  const int nSide = (nDims == 3) ? (int)std::lround(std::cbrt((double)nNodes)) : 2;
  std::vector<double> xl(nSide);
  for (int i = 0; i < nSide; i++) xl[i] = -1. + i * (2. / (nSide - 1));
  if (nDims == 2) {
    static const int QI[4] = {0, 1, 1, 0}, QJ[4] = {0, 0, 1, 1};
    for (int n = 0; n < 4; n++) {
      out[n] = lag(xl, r[0], QI[n]) * lag(xl, r[1], QJ[n]);
      if (dout) {
        dout[3 * n + 0] = dlag(xl, r[0], QI[n]) * lag(xl, r[1], QJ[n]);
        dout[3 * n + 1] = lag(xl, r[0], QI[n]) * dlag(xl, r[1], QJ[n]);
        dout[3 * n + 2] = 0;
      }
    }
    return;
  }
  const std::vector<std::array<int, 3>> nd = hexLattice(nSide);
  for (int n = 0; n < nNodes; n++) {
    const int a = nd[n][0], b = nd[n][1], cc = nd[n][2];
    const double la = lag(xl, r[0], a), lb = lag(xl, r[1], b), lc = lag(xl, r[2], cc);
    out[n] = la * lb * lc;
    if (dout) {
      dout[3 * n + 0] = dlag(xl, r[0], a) * lb * lc;
      dout[3 * n + 1] = la * dlag(xl, r[1], b) * lc;
      dout[3 * n + 2] = la * lb * dlag(xl, r[2], cc);
    }
  }
}

bool hexNodeCount(int nNodes) {
  if (nNodes == 8 || nNodes == 20) return true;
  for (int n = 3; n <= 6; n++)
    if (nNodes == n * n * n) return true;
  return false;
}

std::string buildConnectivity(const fr_config &cfg, const fr_mesh &m, const Operators &o, Connectivity &c) {
  const int D = cfg.nDims, nE = m.nEles, nN = m.nNodes, nfe = o.nFacesEle, nfv = (D == 3) ? 4 : 2;
  const int nfp = o.nFptsFace, P = o.order, nF = o.nFpts;
  // face vertex lists: matching order (geo.cpp:86-91) and orientation order
  // (geo.cpp:804-900); quad edges follow the quad flux-point faces
  static const int HFV[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {3, 0, 4, 7}, {2, 1, 5, 6}, {1, 0, 4, 5}, {3, 2, 6, 7}};
  static const int HOR[6][4] = {{0, 1, 2, 3}, {5, 4, 7, 6}, {0, 3, 7, 4}, {2, 1, 5, 6}, {1, 0, 4, 5}, {3, 2, 6, 7}};
  static const int QFV[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const int TFV[3][2] = {{0, 1}, {1, 2}, {2, 0}};  // triangle edges (fr_tri_ops.cpp)
  const bool triEle = (nN == 3);
  auto cv = [&](int e, int n) -> int {
    int v = m.c2v[(size_t)e * nN + n];
    return m.vmap ? m.vmap[v] : v;
  };
  struct Key {
    std::array<int, 4> v;
    int e, f;
  };
  std::vector<Key> keys((size_t)nE * nfe);
  for (int e = 0; e < nE; e++)
    for (int f = 0; f < nfe; f++) {
      Key &k = keys[(size_t)e * nfe + f];
      k.v = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
      for (int i = 0; i < nfv; i++) k.v[i] = (D == 3) ? cv(e, HFV[f][i]) : triEle ? cv(e, TFV[f][i]) : cv(e, QFV[f][i]);
      std::sort(k.v.begin(), k.v.begin() + nfv);
      k.e = e;
      k.f = f;
    }
  std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
    if (a.v != b.v) return a.v < b.v;
    if (a.e != b.e) return a.e < b.e;
    return a.f < b.f;
  });
  std::vector<std::unordered_set<int>> bnd(m.nBounds);
  for (int b = 0; b < m.nBounds; b++)
    for (int q = m.bndOff[b]; q < m.bndOff[b + 1]; q++) {
      int v = m.bndPts[q];
      if (v < 0 || v >= m.nVerts) return "boundary vertex id out of range";
      bnd[b].insert(m.vmap ? m.vmap[v] : v);
    }
  c.faces.clear();
  std::vector<FaceRec> bfaces;
  for (size_t a = 0; a < keys.size();) {
    size_t b = a + 1;
    while (b < keys.size() && keys[b].v == keys[a].v) b++;
    size_t cnt = b - a;
    if (cnt > 2) return "More than 2 cells for face";
    if (cnt == 2) {
      c.faces.push_back({keys[a].e, keys[a].f, keys[a + 1].e, keys[a + 1].f, 0, -1});
    } else {
      int bc = -1;
      for (int bb = 0; bb < m.nBounds && bc < 0; bb++) {
        bool ok = true;
        for (int i = 0; i < nfv && ok; i++) ok = bnd[bb].count(keys[a].v[i]) > 0;
        if (ok) bc = m.bcType[bb];
      }
      if (bc < 0) return "boundary face matches no boundary";
      static const int known[] = {1, 2, 3, 4, 5, 8, 9, 10, 12, 13};
      if (std::find(std::begin(known), std::end(known), bc) == std::end(known))
        return "Boundary condition not recognized.";
      bfaces.push_back({keys[a].e, keys[a].f, -1, -1, 0, bc});
    }
    a = b;
  }
  c.nIntFaces = (int)c.faces.size();
  c.faces.insert(c.faces.end(), bfaces.begin(), bfaces.end());
  const int nFaces = (int)c.faces.size();
  c.fptR.assign((size_t)nFaces * nfp, -1);
  c.nbr.assign((size_t)nE * nF, -1);
  c.fnbr.assign((size_t)nE * nfe, -1);
  c.finfo.assign((size_t)nE * nfe, -1);
  for (int fi = 0; fi < nFaces; fi++) {
    FaceRec &fr = c.faces[fi];
    if (fr.bc >= 0) {
      c.finfo[(size_t)fr.eL * nfe + fr.fL] = (int8_t)(frdev::FINFO_BOUNDARY + fr.bc);
      continue;
    }
    int *map = &c.fptR[(size_t)fi * nfp];
    if (D == 3) {
      const int v0 = cv(fr.eR, HOR[fr.fR][0]);
      int rot = -1;
      for (int i = 0; i < 4; i++)
        if (cv(fr.eL, HOR[fr.fL][i]) == v0) {
          rot = i;
          break;
        }
      if (rot < 0) return "Internal faces improperly matched.";
      fr.rot = rot;
      const int start = nfp * fr.fR, end = nfp * (fr.fR + 1);
      for (int i = 0; i < nfp; i++) {
        const int a = i % (P + 1), b = i / (P + 1);
        switch (rot) {
          case 0: map[i] = start + a * (P + 1) + b; break;
          case 1: map[i] = start + P - a + b * (P + 1); break;
          case 2: map[i] = end - 1 - (a * (P + 1) + b); break;
          default: map[i] = end - (P + 1) * (b + 1) + a; break;
        }
      }
    } else {
      fr.rot = 1;
      for (int i = 0; i < nfp; i++) map[i] = fr.fR * nfp + nfp - 1 - i;
    }
    c.fnbr[(size_t)fr.eL * nfe + fr.fL] = fr.eR;
    c.fnbr[(size_t)fr.eR * nfe + fr.fR] = fr.eL;
    c.finfo[(size_t)fr.eL * nfe + fr.fL] = (int8_t)frdev::packFinfo(fr.fR, fr.rot, 0);
    c.finfo[(size_t)fr.eR * nfe + fr.fR] = (int8_t)frdev::packFinfo(fr.fL, fr.rot, 1);
    for (int i = 0; i < nfp; i++) {
      const int qL = fr.fL * nfp + i, qR = map[i];
      c.nbr[(size_t)fr.eL * nF + qL] = fr.eR * nF + qR;
      c.nbr[(size_t)fr.eR * nF + qR] = fr.eL * nF + qL;
    }
  }
  for (size_t i = 0; i < c.finfo.size(); i++)
    if (c.finfo[i] < 0) return "element face without a face record";
  return verifyNeighbourMaps(o, c, nE);
}

// Morton (Z-order) key of a point quantised to 21 bits per axis.
static uint64_t mortonKey(const double *x, const double *lo, const double *span) {
  uint64_t key = 0;
  uint32_t q[3];
  for (int d = 0; d < 3; d++) {
    double t = span[d] > 0 ? (x[d] - lo[d]) / span[d] : 0.0;
    t = std::min(std::max(t, 0.0), 1.0);
    q[d] = (uint32_t)(t * 2097151.0);
  }
  for (int bit = 20; bit >= 0; bit--)
    for (int d = 2; d >= 0; d--) key = (key << 1) | ((q[d] >> bit) & 1u);
  return key;
}

std::vector<int> blockOrder(const Connectivity &c, const std::vector<double> &centroid, int nE, int nFaceEle,
                            int epb) {
  std::vector<int> order;
  order.reserve(nE);
  if (epb <= 1) {
    for (int e = 0; e < nE; e++) order.push_back(e);
    return order;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, span[3];
  for (int e = 0; e < nE; e++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], centroid[3 * (size_t)e + d]);
      hi[d] = std::max(hi[d], centroid[3 * (size_t)e + d]);
    }
  for (int d = 0; d < 3; d++) span[d] = hi[d] - lo[d];
  std::vector<std::pair<uint64_t, int>> keyed(nE);
  for (int e = 0; e < nE; e++) keyed[e] = {mortonKey(&centroid[3 * (size_t)e], lo, span), e};
  std::sort(keyed.begin(), keyed.end());
  std::vector<int> rank(nE), byRank(nE), cluster(nE, -1);
  for (int i = 0; i < nE; i++) {
    rank[keyed[i].second] = i;
    byRank[i] = keyed[i].second;
  }
  std::vector<int> members;
  size_t next = 0;  // first rank that may still be unassigned
  int id = 0;
  for (int i = 0; i < nE; i++) {
    const int seed = byRank[i];
    if (cluster[seed] >= 0) continue;
    members.assign(1, seed);
    cluster[seed] = id;
    // grow by the unassigned face neighbour with the most faces into the
    // cluster; ties: the one that leaves more unassigned elements touching the
    // cluster twice (on a grid: turn a corner, so that the fourth element
    // closes a 2x2 square instead of extending a line), then lowest Z-order rank
    while ((int)members.size() < epb) {
      int best = -1, bestCnt = 0, bestLook = -1;
      for (int m : members)
        for (int f = 0; f < nFaceEle; f++) {
          const int n = c.fnbr[(size_t)m * nFaceEle + f];
          if (n < 0 || cluster[n] >= 0) continue;
          int cnt = 0;
          for (int g = 0; g < nFaceEle; g++) {
            const int o = c.fnbr[(size_t)n * nFaceEle + g];
            if (o >= 0 && cluster[o] == id) cnt++;
          }
          int look = 0;
          if ((int)members.size() + 1 < epb)
            for (int g = 0; g < nFaceEle; g++) {
              const int o = c.fnbr[(size_t)n * nFaceEle + g];
              if (o < 0 || o == n || cluster[o] >= 0) continue;
              for (int h = 0; h < nFaceEle; h++) {
                const int w = c.fnbr[(size_t)o * nFaceEle + h];
                if (w >= 0 && cluster[w] == id) {
                  look++;
                  break;
                }
              }
            }
          if (cnt > bestCnt ||
              (cnt == bestCnt && best >= 0 && (look > bestLook || (look == bestLook && rank[n] < rank[best])))) {
            best = n;
            bestCnt = cnt;
            bestLook = look;
          }
        }
      if (best < 0) break;
      cluster[best] = id;
      members.push_back(best);
    }
    // a cluster that cannot grow is filled up in Z-order, so that every block
    // of epb device elements is one cluster
    while ((int)members.size() < epb) {
      while (next < (size_t)nE && cluster[byRank[next]] >= 0) next++;
      if (next == (size_t)nE) break;
      cluster[byRank[next]] = id;
      members.push_back(byRank[next]);
    }
    for (int m : members) order.push_back(m);
    id++;
  }
  return order;
}

}  // namespace frhost
