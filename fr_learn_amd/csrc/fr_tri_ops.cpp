// fr_tri_ops.cpp -- FR operators of the linear triangle (2D), orders 1..5.
//
// FlurryPP builds operators for QUAD and HEX only (operators.cpp:24-26); the
// triangle here is this library's own definition:
//  * solution points: the warp-and-blend nodes of Hesthaven & Warburton
//    (Nodal Discontinuous Galerkin Methods, 2008, sec. 6.1: Nodes2D,
//    Warpfactor) on the reference triangle (-1,-1), (1,-1), (-1,1);
//  * flux points: the P+1 Gauss-Legendre abscissae of the quad operators on
//    each edge, edge f running from vertex f to vertex f+1 (counter-
//    clockwise, the Gmsh node order), so that two elements sharing an edge
//    meet at the same points in opposite order (neighbourFpt<2, P>);
//  * the orthonormal Dubiner basis (Simplex2DP / GradSimplex2DP, sec. 6.1):
//    Vandermonde V, differentiation Dr = V_r V^-1, Ds = V_s V^-1,
//    extrapolation E = Psi(fpts) V^-1;
//  * correction: the DG-equivalent VCJH lifting (c = 0; Castonguay, Vincent
//    & Jameson 2012), L = V Psi(fpts)^T W with W the edge quadrature weights:
//    for every v of degree <= P, int v (L g) = sum_j w_j v(x_j) g_j.
// Operators exported as the quad ones: spts_to_fpts = E, grad_spts = {Dr, Ds},
// extrapolateFn[d] = tNorm_d(fpt) E, correction = L.  The kernel table
// (oned) is Dr | Ds | E | L, row-major.
#include <cmath>
#include <string>
#include <vector>

#include "fr_host.hpp"

namespace frhost {

namespace {

// Orthonormal Jacobi polynomial P_n^(a,b)(x) (JacobiP).
double jacobiP(double x, double a, double b, int n) {
  const double g0 = std::pow(2.0, a + b + 1) / (a + b + 1) * std::tgamma(a + 1) * std::tgamma(b + 1) /
                    std::tgamma(a + b + 1);
  double pm = 1.0 / std::sqrt(g0);
  if (n == 0) return pm;
  const double g1 = (a + 1) * (b + 1) / (a + b + 3) * g0;
  double pc = ((a + b + 2) * x / 2 + (a - b) / 2) / std::sqrt(g1);
  if (n == 1) return pc;
  double aold = 2 / (2 + a + b) * std::sqrt((a + 1) * (b + 1) / (a + b + 3));
  for (int i = 1; i < n; i++) {
    const double h1 = 2 * i + a + b;
    const double anew =
        2 / (h1 + 2) * std::sqrt((i + 1) * (i + 1 + a + b) * (i + 1 + a) * (i + 1 + b) / (h1 + 1) / (h1 + 3));
    const double bnew = -(a * a - b * b) / h1 / (h1 + 2);
    const double pn = 1 / anew * (-aold * pm + (x - bnew) * pc);
    aold = anew;
    pm = pc;
    pc = pn;
  }
  return pc;
}

double gradJacobiP(double x, double a, double b, int n) {
  return n == 0 ? 0.0 : std::sqrt(n * (n + a + b + 1)) * jacobiP(x, a + 1, b + 1, n - 1);
}

void rsToAb(double r, double s, double &a, double &b) {
  a = (s != 1.0) ? 2 * (1 + r) / (1 - s) - 1 : -1.0;
  b = s;
}

// Dubiner mode (i, j) and its gradient at (r, s).
double simplexP(double r, double s, int i, int j) {
  double a, b;
  rsToAb(r, s, a, b);
  return std::sqrt(2.0) * jacobiP(a, 0, 0, i) * jacobiP(b, 2 * i + 1, 0, j) * std::pow(1 - b, i);
}

void gradSimplexP(double r, double s, int i, int j, double &dr, double &ds) {
  double a, b;
  rsToAb(r, s, a, b);
  const double fa = jacobiP(a, 0, 0, i), dfa = gradJacobiP(a, 0, 0, i);
  const double gb = jacobiP(b, 2 * i + 1, 0, j), dgb = gradJacobiP(b, 2 * i + 1, 0, j);
  dr = dfa * gb;
  if (i > 0) dr *= std::pow(0.5 * (1 - b), i - 1);
  ds = dfa * (gb * (0.5 * (1 + a)));
  if (i > 0) ds *= std::pow(0.5 * (1 - b), i - 1);
  double tmp = dgb * std::pow(0.5 * (1 - b), i);
  if (i > 0) tmp -= 0.5 * i * gb * std::pow(0.5 * (1 - b), i - 1);
  ds += fa * tmp;
  const double sc = std::pow(2.0, i + 0.5);
  dr *= sc;
  ds *= sc;
}

// Gauss-Jordan inverse with partial pivoting (n <= 21); "" or an error.
std::string invert(std::vector<double> A, int n, std::vector<double> &inv) {
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)p * n + c])) p = r;
    if (std::fabs(A[(size_t)p * n + c]) < 1e-14) return "singular triangle Vandermonde matrix";
    for (int k = 0; k < n; k++) {
      std::swap(A[(size_t)c * n + k], A[(size_t)p * n + k]);
      std::swap(inv[(size_t)c * n + k], inv[(size_t)p * n + k]);
    }
    const double d = A[(size_t)c * n + c];
    for (int k = 0; k < n; k++) {
      A[(size_t)c * n + k] /= d;
      inv[(size_t)c * n + k] /= d;
    }
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      const double f = A[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; k++) {
        A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  return "";
}

// Gauss-Lobatto-Legendre nodes (P+1) by Newton on (1 - x^2) P_P'(x).
std::vector<double> gll(int P) {
  std::vector<double> x(P + 1);
  for (int i = 0; i <= P; i++) {
    double t = -std::cos(M_PI * i / P);  // Chebyshev-Gauss-Lobatto start
    if (i > 0 && i < P)
      for (int it = 0; it < 100; it++) {
        // Legendre P_P, P_{P-1} at t by recurrence
        double p0 = 1.0, p1 = t;
        for (int k = 2; k <= P; k++) {
          const double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        // interior nodes: roots of q = P_{P-1} - t P_P = (1 - t^2) P_P' / P,
        // whose derivative is -(P+1) P_P (Legendre's equation)
        const double q = p0 - t * p1;
        const double dt = q / ((P + 1) * p1);
        t += dt;
        if (std::fabs(dt) < 1e-16) break;
      }
    x[i] = t;
  }
  return x;
}

}  // namespace

std::string buildTriOperators(const fr_config &cfg, Operators &o) {
  const int P = cfg.order;
  if (P < 1 || P > 5) return "order must be in 1..5";
  if (cfg.viscous) return "triangles: inviscid equations only (no LDG operators for triangles)";
  if (cfg.vcjhScheme != 0) return "triangles: only the DG correction (vcjhScheme 0)";
  o.nDims = 2;
  o.order = P;
  const int P1 = o.P1 = P + 1;
  const int nS = o.nSpts = P1 * (P + 2) / 2;
  o.nFptsFace = P1;
  o.nFacesEle = 3;
  const int nF = o.nFpts = 3 * P1;
  o.nNodes = 3;
  o.spts1D = pts1D(0, P);  // Gauss-Legendre edge points, the quads' tables
  std::vector<double> w1d(P1);
  // weights of the P+1-point Gauss rule: w_i = 2 / ((1 - x_i^2) P'_{P+1}(x_i)^2)
  for (int i = 0; i < P1; i++) {
    const double t = o.spts1D[i];
    double p0 = 1.0, p1 = t;
    for (int k = 2; k <= P1; k++) {
      const double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    const double dp = P1 * (t * p1 - p0) / (t * t - 1.0);
    w1d[i] = 2.0 / ((1.0 - t * t) * dp * dp);
  }

  // solution points: warp & blend
  static const double alpopt[6] = {0.0, 0.0, 0.0, 1.4152, 0.1001, 0.2751};
  const double alpha = alpopt[P];
  const std::vector<double> lgl = gll(P);
  std::vector<double> req(P1), veqT((size_t)P1 * P1), veqTi;
  for (int i = 0; i < P1; i++) req[i] = -1.0 + 2.0 * i / P;
  for (int i = 0; i < P1; i++)
    for (int n = 0; n < P1; n++) veqT[(size_t)n * P1 + i] = jacobiP(req[i], 0, 0, n);  // Veq^T
  std::string err = invert(veqT, P1, veqTi);
  if (!err.empty()) return err;
  auto warpf = [&](double rout) {
    // warp = Lmat^T (lgl - req), Lmat = Veq^-T Pmat(rout)
    double w = 0.0;
    for (int i = 0; i < P1; i++) {
      double l = 0.0;
      for (int n = 0; n < P1; n++) l += veqTi[(size_t)i * P1 + n] * jacobiP(rout, 0, 0, n);
      w += l * (lgl[i] - req[i]);
    }
    const bool zf = std::fabs(rout) < 1.0 - 1e-10;
    const double sf = 1.0 - (zf ? rout * rout : 0.0);
    return w / sf + w * ((zf ? 1.0 : 0.0) - 1.0);
  };
  o.locSpts.assign((size_t)nS * 3, 0.0);
  {
    int k = 0;
    for (int n = 0; n <= P; n++)
      for (int m = 0; m <= P - n; m++, k++) {
        const double L1 = (double)n / P, L3 = (double)m / P, L2 = 1.0 - L1 - L3;
        double x = -L2 + L3, y = (-L2 - L3 + 2 * L1) / std::sqrt(3.0);
        const double w1 = 4 * L2 * L3 * warpf(L3 - L2) * (1 + (alpha * L1) * (alpha * L1));
        const double w2 = 4 * L1 * L3 * warpf(L1 - L3) * (1 + (alpha * L2) * (alpha * L2));
        const double w3 = 4 * L1 * L2 * warpf(L2 - L1) * (1 + (alpha * L3) * (alpha * L3));
        x += w1 + std::cos(2 * M_PI / 3) * w2 + std::cos(4 * M_PI / 3) * w3;
        y += std::sin(2 * M_PI / 3) * w2 + std::sin(4 * M_PI / 3) * w3;
        const double l1 = (std::sqrt(3.0) * y + 1.0) / 3.0;
        const double l2 = (-3.0 * x - std::sqrt(3.0) * y + 2.0) / 6.0;
        const double l3 = (3.0 * x - std::sqrt(3.0) * y + 2.0) / 6.0;
        o.locSpts[3 * k] = -l2 + l3 - l1;
        o.locSpts[3 * k + 1] = -l2 - l3 + l1;
      }
  }
  // flux points and reference normals
  static const double VX[3] = {-1.0, 1.0, -1.0}, VY[3] = {-1.0, -1.0, 1.0};
  const double is2 = 1.0 / std::sqrt(2.0);
  const double TN[3][2] = {{0.0, -1.0}, {is2, is2}, {-1.0, 0.0}};
  const double LEN[3] = {2.0, 2.0 * std::sqrt(2.0), 2.0};
  o.locFpts.assign((size_t)nF * 3, 0.0);
  o.tNorm.assign((size_t)nF * 3, 0.0);
  o.wtsFace = w1d;
  std::vector<double> W(nF);
  for (int f = 0; f < 3; f++)
    for (int j = 0; j < P1; j++) {
      const int q = f * P1 + j;
      const double t = o.spts1D[j];
      const int g = (f + 1) % 3;
      o.locFpts[3 * q] = VX[f] * (1 - t) / 2 + VX[g] * (1 + t) / 2;
      o.locFpts[3 * q + 1] = VY[f] * (1 - t) / 2 + VY[g] * (1 + t) / 2;
      o.tNorm[3 * q] = TN[f][0];
      o.tNorm[3 * q + 1] = TN[f][1];
      W[q] = w1d[j] * LEN[f] / 2;
    }
  // Vandermonde matrices
  std::vector<std::pair<int, int>> md;
  for (int i = 0; i <= P; i++)
    for (int j = 0; j <= P - i; j++) md.push_back({i, j});
  std::vector<double> V((size_t)nS * nS), Vr((size_t)nS * nS), Vs((size_t)nS * nS), Vi, Psi((size_t)nF * nS);
  for (int p = 0; p < nS; p++)
    for (int n = 0; n < nS; n++) {
      const double r = o.locSpts[3 * p], s = o.locSpts[3 * p + 1];
      V[(size_t)p * nS + n] = simplexP(r, s, md[n].first, md[n].second);
      gradSimplexP(r, s, md[n].first, md[n].second, Vr[(size_t)p * nS + n], Vs[(size_t)p * nS + n]);
    }
  for (int q = 0; q < nF; q++)
    for (int n = 0; n < nS; n++)
      Psi[(size_t)q * nS + n] = simplexP(o.locFpts[3 * q], o.locFpts[3 * q + 1], md[n].first, md[n].second);
  if (!(err = invert(V, nS, Vi)).empty()) return err;
  auto mul = [&](const std::vector<double> &A, int m, const std::vector<double> &B, int n) {
    std::vector<double> C((size_t)m * n, 0.0);  // A (m x nS) * B (nS x n)
    for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) {
        double acc = 0.0;
        for (int k = 0; k < nS; k++) acc += A[(size_t)i * nS + k] * B[(size_t)k * n + j];
        C[(size_t)i * n + j] = acc;
      }
    return C;
  };
  const std::vector<double> Dr = mul(Vr, nS, Vi, nS), Ds = mul(Vs, nS, Vi, nS), E = mul(Psi, nF, Vi, nS);
  std::vector<double> L((size_t)nS * nF, 0.0);
  for (int i = 0; i < nS; i++)
    for (int q = 0; q < nF; q++) {
      double acc = 0.0;
      for (int n = 0; n < nS; n++) acc += V[(size_t)i * nS + n] * Psi[(size_t)q * nS + n];
      L[(size_t)i * nF + q] = acc * W[q];
    }
  o.spts_to_fpts = E;
  o.grad_spts.assign((size_t)2 * nS * nS, 0.0);
  for (size_t i = 0; i < Dr.size(); i++) {
    o.grad_spts[i] = Dr[i];
    o.grad_spts[Dr.size() + i] = Ds[i];
  }
  o.extrapolateFn.assign((size_t)2 * nF * nS, 0.0);
  for (int d = 0; d < 2; d++)
    for (int q = 0; q < nF; q++)
      for (int s = 0; s < nS; s++)
        o.extrapolateFn[((size_t)d * nF + q) * nS + s] = o.tNorm[3 * q + d] * E[(size_t)q * nS + s];
  o.correction = L;
  o.correctU.clear();
  // quadrature weights of the nodal basis: int l_i = sqrt(2) (V^-1)_{0 i}
  o.wts.assign(nS, 0.0);
  for (int i = 0; i < nS; i++) o.wts[i] = std::sqrt(2.0) * Vi[i];
  // linear-triangle shape functions and derivatives (Gmsh order)
  auto fill = [&](const std::vector<double> &loc, int nPts, std::vector<double> &out) {
    out.assign((size_t)2 * nPts * 3, 0.0);
    for (int p = 0; p < nPts; p++) {
      out[((size_t)0 * nPts + p) * 3 + 0] = -0.5;
      out[((size_t)0 * nPts + p) * 3 + 1] = 0.5;
      out[((size_t)1 * nPts + p) * 3 + 0] = -0.5;
      out[((size_t)1 * nPts + p) * 3 + 2] = 0.5;
    }
    (void)loc;
  };
  fill(o.locSpts, nS, o.dshapeSpts);
  fill(o.locFpts, nF, o.dshapeFpts);
  o.shapeSpts.assign((size_t)nS * 3, 0.0);
  for (int p = 0; p < nS; p++) shapeFunctions(2, &o.locSpts[3 * p], &o.shapeSpts[(size_t)p * 3], 3);
  // kernel table: Dr | Ds | E | L | Dr - L tNorm_r E | Ds - L tNorm_s E.  The
  // last two fold the jump into the divergence: with J = Fn dA - (tNorm_r E
  // F~_r + tNorm_s E F~_s), Dr F~_r + Ds F~_s + L J equals
  // (Dr - L tNorm_r E) F~_r + (Ds - L tNorm_s E) F~_s + L Fn dA, tNorm_d the
  // diagonal of the flux points' reference normals (the FP64 kernel's form).
  o.oned.clear();
  o.oned.insert(o.oned.end(), Dr.begin(), Dr.end());
  o.oned.insert(o.oned.end(), Ds.begin(), Ds.end());
  o.oned.insert(o.oned.end(), E.begin(), E.end());
  o.oned.insert(o.oned.end(), L.begin(), L.end());
  for (int d = 0; d < 2; d++) {
    const std::vector<double> &D = d == 0 ? Dr : Ds;
    for (int r = 0; r < nS; r++)
      for (int k = 0; k < nS; k++) {
        double acc = 0.0;
        for (int q = 0; q < nF; q++) acc += L[(size_t)r * nF + q] * o.tNorm[3 * q + d] * E[(size_t)q * nS + k];
        o.oned.push_back(D[(size_t)r * nS + k] - acc);
      }
  }
  return "";
}

}  // namespace frhost
