/*
 * fr_oracle.c -- CPU FP64 oracle (TEST INFRASTRUCTURE ONLY; see fr_oracle.h).
 *
 * Plain-C restatement of the FlurryPP residual + RK path.  Each function cites
 * the reference file:line it follows (paths relative to the reference root).
 * Design notes:
 *  - arrays use the reference layout and every operator application is a dense
 *    row-major dgemm (oracle_dgemm) like the reference's cblas_dgemm calls;
 *  - face connectivity is found by sorting canonical face-vertex lists instead
 *    of the reference's quadratic scans (src/geo.cpp:75-267); the left element
 *    of a face is still the first (lowest-index) element that owns it;
 *  - periodic faces (absent from the reference) come from vertex
 *    identification: vmap[] maps each vertex to a canonical id used only for
 *    matching, so a periodic box is matched like a torus.
 * Parity pinning of each part is done by tests/ (closed forms, invariants,
 * an independent weak-form DG evaluation, geometric face matching).
 */
#include "fr_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static char g_err[512];
static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
const char *or_last_error(void) { return g_err; }

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}


/* ------------------------------------------------------------------------ */
/* 1D point tables: 15-decimal Gauss-Legendre / Gauss-Lobatto abscissae and  */
/* 16-decimal Gauss weights, as tabulated in src/points.cpp:162-345 and      */
/* 376-469.  Kept as data (including their last-digit asymmetries) because   */
/* operator parity depends on the exact abscissae.                           */
/* ------------------------------------------------------------------------ */
static const double GL_PTS[11][11] = {
    {0.0},
    {-0.577350269189626, 0.577350269189626},
    {-0.774596669241483, 0.000000000000000, 0.774596669241483},
    {-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053},
    {-0.906179845938664, -0.538469310105683, 0.000000000000000, 0.538469310105683, 0.906179845938664},
    {-0.932469514203152, -0.661209386466264, -0.238619186083197, 0.238619186083197, 0.661209386466265, 0.932469514203152},
    {-0.949107912342758, -0.741531185599395, -0.405845151377397, 0.000000000000000, 0.405845151377397, 0.741531185599394, 0.949107912342758},
    {-0.960289856497536, -0.796666477413627, -0.525532409916329, -0.183434642495650, 0.183434642495650, 0.525532409916329, 0.796666477413627, 0.960289856497536},
    {-0.968160239507626, -0.836031107326636, -0.613371432700591, -0.324253423403809, 0.000000000000000, 0.324253423403809, 0.613371432700591, 0.836031107326636, 0.968160239507626},
    {-0.973906528517172, -0.865063366688985, -0.679409568299024, -0.433395394129247, -0.148874338981631, 0.148874338981632, 0.433395394129247, 0.679409568299024, 0.865063366688984, 0.973906528517171},
    {-0.978228658146057, -0.887062599768095, -0.730152005574049, -0.519096129206812, -0.269543155952345, 0.000000000000000, 0.269543155952345, 0.519096129206812, 0.730152005574049, 0.887062599768095, 0.978228658146057}};

static const double GLL_PTS[11][11] = {
    {0.0},
    {-1.000000000000000, 1.000000000000000},
    {-1.000000000000000, -0.000000000000000, 1.000000000000000},
    {-1.000000000000000, -0.447213595499958, 0.447213595499958, 1.000000000000000},
    {-1.000000000000000, -0.654653670707977, -0.000000000000000, 0.654653670707977, 1.000000000000000},
    {-1.000000000000000, -0.765055323929465, -0.285231516480645, 0.285231516480645, 0.765055323929465, 1.000000000000000},
    {-1.000000000000000, -0.830223896278567, -0.468848793470714, 0.000000000000000, 0.468848793470714, 0.830223896278567, 1.000000000000000},
    {-1.000000000000000, -0.871740148509607, -0.591700181433142, -0.209299217902479, 0.209299217902479, 0.591700181433142, 0.871740148509607, 1.000000000000000},
    {-1.000000000000000, -0.899757995411460, -0.677186279510738, -0.363117463826178, -0.000000000000000, 0.363117463826178, 0.677186279510738, 0.899757995411460, 1.000000000000000},
    {-1.000000000000000, -0.919533908166459, -0.738773865105505, -0.477924949810445, -0.165278957666387, 0.165278957666387, 0.477924949810444, 0.738773865105505, 0.919533908166459, 1.000000000000000},
    {-1.000000000000000, -0.934001430408059, -0.784483473663144, -0.565235326996205, -0.295758135586939, 0.000000000000000, 0.295758135586939, 0.565235326996205, 0.784483473663144, 0.934001430408059, 1.000000000000000}};

static const double GL_WTS[11][11] = {
    {2.0},
    {1.0, 1.0},
    {0.5555555555555556, 0.8888888888888888, 0.5555555555555556},
    {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538},
    {0.2369268850561891, 0.4786286704993665, 0.5688888888888889, 0.4786286704993665, 0.2369268850561891},
    {0.1713244923791704, 0.3607615730481386, 0.4679139345726910, 0.4679139345726910, 0.3607615730481386, 0.1713244923791704},
    {0.1294849661688697, 0.2797053914892766, 0.3818300505051189, 0.4179591836734694, 0.3818300505051189, 0.2797053914892766, 0.1294849661688697},
    {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620, 0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763},
    {0.0812743883615744, 0.1806481606948574, 0.2606106964029354, 0.3123470770400029, 0.3302393550012598, 0.3123470770400029, 0.2606106964029354, 0.1806481606948574, 0.0812743883615744},
    {0.0666713443086881, 0.1494513491505806, 0.2190863625159820, 0.2692667193099963, 0.2955242247147529, 0.2955242247147529, 0.2692667193099963, 0.2190863625159820, 0.1494513491505806, 0.0666713443086881},
    {0.0556685671161737, 0.1255803694649046, 0.1862902109277343, 0.2331937645919905, 0.2628045445102467, 0.2729250867779006, 0.2628045445102467, 0.2331937645919905, 0.1862902109277343, 0.1255803694649046, 0.0556685671161737}};

/* getPts1D, src/points.cpp:162-345 (Equidistant branch 165-168) */
static int pts1d(int type, int order, double *out) {
  if (order < 0 || order > 10) { set_err("points for order %d not tabulated", order); return -1; }
  for (int i = 0; i <= order; i++) {
    if (type == 0) out[i] = GL_PTS[order][i];
    else if (type == 1) out[i] = GLL_PTS[order][i];
    else out[i] = -1 + (2. / order) * i; /* Equidistant */
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* polynomials, src/polynomials.cpp                                          */
/* ------------------------------------------------------------------------ */
/* Lagrange, polynomials.cpp:9-19 */
static double lagrange(const double *x, int n, double y, int mode) {
  double lag = 1.0;
  for (int i = 0; i < n; i++)
    if (i != mode) lag = lag * ((y - x[i]) / (x[mode] - x[i]));
  return lag;
}
/* dLagrange, polynomials.cpp:21-47 */
static double dlagrange(const double *x, int n, double y, int mode) {
  double dLag = 0.0;
  for (int i = 0; i < n; i++) {
    if (i == mode) continue;
    double num = 1.0, den = 1.0;
    for (int j = 0; j < n; j++) {
      if (j != mode && j != i) num = num * (y - x[j]);
      if (j != mode) den = den * (x[mode] - x[j]);
    }
    dLag = dLag + (num / den);
  }
  return dLag;
}
/* Legendre, polynomials.cpp:84-95 (same recursion) */
static double legendre(double r, int mode) {
  if (mode < 0) return 0.;
  if (mode == 0) return 1.0;
  if (mode == 1) return r;
  return ((2 * mode - 1) * r * legendre(r, mode - 1) - (mode - 1) * legendre(r, mode - 2)) / mode;
}
/* dLegendre, polynomials.cpp:97-118 */
static double dlegendre(double r, int mode) {
  double d = 0.;
  if (mode == 0) return 0;
  if (r > -1.0 && r < 1.0) {
    d = mode * ((r * legendre(r, mode)) - legendre(r, mode - 1)) / (r * r - 1.0);
  } else {
    if (r == -1.0) d = pow(-1.0, mode - 1.0) * 0.5 * mode * (mode + 1.0);
    if (r == 1.0) d = 0.5 * mode * (mode + 1.0);
  }
  return d;
}
static int factorial_i(int n) { return (n == 1 || n == 0) ? 1 : factorial_i(n - 1) * n; } /* global.cpp:27 */

/* compute_eta, polynomials.cpp:392-428 */
static double compute_eta(int scheme, int order) {
  if (scheme == OR_DG) return 0.0;
  if (scheme == OR_SD) return (1.0 * order) / (1.0 * (order + 1));
  if (scheme == OR_HU) return (1.0 * (order + 1)) / (1.0 * order);
  double c1d = 0;
  if (order == 2) c1d = 0.206;
  else if (order == 3) c1d = 3.80e-3;
  else if (order == 4) c1d = 4.67e-5;
  else if (order == 5) c1d = 4.28e-7;
  double ap = 1. / pow(2.0, order) * factorial_i(2 * order) / (factorial_i(order) * factorial_i(order));
  return c1d * (2 * order + 1) / 2 * (factorial_i(order) * ap) * (factorial_i(order) * ap);
}
/* dVCJH_1d, polynomials.cpp:447-475 */
static double dvcjh_1d(double r, int mode, int order, double eta) {
  double d = 0.;
  if (mode == 0) {
    if (order == 0)
      d = 0.5 * pow(-1.0, order) * (dlegendre(r, order) - ((dlegendre(r, order + 1)) / (1.0 + eta)));
    else
      d = 0.5 * pow(-1.0, order) *
          (dlegendre(r, order) - (((eta * dlegendre(r, order - 1)) + dlegendre(r, order + 1)) / (1.0 + eta)));
  } else if (mode == 1) {
    if (order == 0)
      d = 0.5 * (dlegendre(r, order) + ((dlegendre(r, order + 1)) / (1.0 + eta)));
    else
      d = 0.5 * (dlegendre(r, order) + (((eta * dlegendre(r, order - 1)) + dlegendre(r, order + 1)) / (1.0 + eta)));
  }
  return d;
}

/* ------------------------------------------------------------------------ */
/* solver state                                                               */
/* ------------------------------------------------------------------------ */
struct or_solver {
  or_params p;
  int nDims, order, P1, nSpts, nFpts, nFptsFace, nFacesEle, nFields, nEles, nNodes, nRK;
  double RKa[5], RKb[4];
  double *spts1D, *wts;                 /* spts1D[P+1], quadrature weights per spt */
  double *loc_spts, *loc_fpts, *tNorm;  /* [pt][3] */
  double *opp_spts_to_fpts, *opp_grad_spts, *opp_extrapolateFn, *opp_correction, *opp_correctU;
  /* geometry */
  double *nodes, *pos_spts, *Jac_spts, *JGinv_spts, *detJac_spts;
  double *Jac_fpts, *JGinv_fpts, *detJac_fpts, *norm_fpts, *dA_fpts;
  /* faces */
  int nFaces, nIntFaces;
  int *fL, *flocL, *fR, *flocR, *frot, *fbc; /* per face */
  int *fptR;                                  /* [nFaces][nFptsFace], element-local fpt id in R */
  /* solution */
  double *U_spts, *U_fpts, *F_spts, *Fn_fpts, *disFn_fpts, *U0, *divF[4];
  double *dU_spts, *dU_fpts, *dUc_fpts;
  double *faceFn; /* [nFaces][nFptsFace][nFields]: each face's common flux kept between
                     the inviscid and viscous passes (face.hpp Fn member) */
  double *faceUC; /* [nFaces][nFptsFace][nFields] LDG common solution */
  int dtType;      /* 0 fixed, 1 global CFL dt, 2 local (per-element) CFL dt */
  double *dtEle;   /* [nEles] ele::dt for dtType 2 */
  /* CFL time stepping: ele::waveSp_fpts as [fpt][ele] (written by
   * calcWaveSpFpts and by the faces' Rusanov / Lax-Friedrichs fluxes at the
   * left element's points), the CFL number, and whether the next update is
   * the first (params->iter == initIter + 1, solver.cpp:294) */
  double *waveSp;
  double CFL;
  int dtFirst;
};

#define IDX3(a, b, c, B, C) ((c) + (C) * ((b) + (B) * (a)))

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* Row-major C = alpha*A*B + beta*C, A m x k, B k x n (all NoTrans) -- the
 * textbook product that cblas_dgemm computes at every call site in
 * src/solver.cpp.  Summation over k runs in ascending order. */
static void oracle_dgemm(int m, int n, int k, double alpha, const double *A, const double *B,
                         double beta, double *C) {
#pragma omp parallel
  {
#ifdef _OPENMP
    int nt = omp_get_num_threads(), t = omp_get_thread_num();
#else
    int nt = 1, t = 0;
#endif
    const int NB = 512;
    int nblk = (n + NB - 1) / NB;
    double acc[512];
    for (int blk = t; blk < nblk; blk += nt) {
      int j0 = blk * NB, j1 = j0 + NB < n ? j0 + NB : n, w = j1 - j0;
      for (int i = 0; i < m; i++) {
        for (int j = 0; j < w; j++) acc[j] = 0.0;
        const double *Ai = A + (size_t)i * k;
        for (int kk = 0; kk < k; kk++) {
          double a = Ai[kk];
          const double *Bk = B + (size_t)kk * n + j0;
          for (int j = 0; j < w; j++) acc[j] += a * Bk[j];
        }
        double *Ci = C + (size_t)i * n + j0;
        if (beta == 0.0)
          for (int j = 0; j < w; j++) Ci[j] = alpha * acc[j];
        else
          for (int j = 0; j < w; j++) Ci[j] = alpha * acc[j] + beta * Ci[j];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* operators, src/operators.cpp                                               */
/* ------------------------------------------------------------------------ */
/* reference-face unit normals: hex tNorm (solver.cpp:171-203, operators.cpp
 * 847-881), quad tNorm (operators.cpp:814-843) */
static void tnorm_face(int nDims, int face, double *t) {
  t[0] = t[1] = t[2] = 0;
  if (nDims == 3) {
    static const double T[6][3] = {{0, 0, -1}, {0, 0, 1}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}};
    t[0] = T[face][0]; t[1] = T[face][1]; t[2] = T[face][2];
  } else {
    static const double T[4][2] = {{0, -1}, {1, 0}, {0, 1}, {-1, 0}};
    t[0] = T[face][0]; t[1] = T[face][1];
  }
}

/* getLocSpts / getLocFpts, points.cpp:37-128 */
static void setup_points(or_solver *s) {
  int P = s->order, P1 = P + 1;
  const double *x = s->spts1D;
  if (s->nDims == 3) {
    for (int k = 0; k < P1; k++)
      for (int j = 0; j < P1; j++)
        for (int i = 0; i < P1; i++) {
          double *q = s->loc_spts + 3 * (i + P1 * (j + P1 * k));
          q[0] = x[i]; q[1] = x[j]; q[2] = x[k];
        }
    int P12 = P1 * P1;
    for (int i = 0; i < P1; i++)
      for (int j = 0; j < P1; j++) {
        int ind = i + j * P1;
        double *f;
        f = s->loc_fpts + 3 * (ind);           f[0] = x[i];     f[1] = x[j];     f[2] = -1.;
        f = s->loc_fpts + 3 * (ind + P12);     f[0] = x[P - i]; f[1] = x[j];     f[2] = 1.;
        f = s->loc_fpts + 3 * (ind + 2 * P12); f[0] = -1;       f[1] = x[i];     f[2] = x[j];
        f = s->loc_fpts + 3 * (ind + 3 * P12); f[0] = 1;        f[1] = x[P - i]; f[2] = x[j];
        f = s->loc_fpts + 3 * (ind + 4 * P12); f[0] = x[P - i]; f[1] = -1;       f[2] = x[j];
        f = s->loc_fpts + 3 * (ind + 5 * P12); f[0] = x[i];     f[1] = 1;        f[2] = x[j];
      }
  } else {
    for (int i = 0; i < P1; i++)
      for (int j = 0; j < P1; j++) {
        double *q = s->loc_spts + 3 * (j + i * P1);
        q[0] = x[j]; q[1] = x[i]; q[2] = 0;
      }
    for (int i = 0; i < P1; i++) {
      double *f;
      f = s->loc_fpts + 3 * i;            f[0] = x[i];     f[1] = -1.;      f[2] = 0;
      f = s->loc_fpts + 3 * (i + P1);     f[0] = 1.;       f[1] = x[i];     f[2] = 0;
      f = s->loc_fpts + 3 * (i + 2 * P1); f[0] = x[P - i]; f[1] = 1.;       f[2] = 0;
      f = s->loc_fpts + 3 * (i + 3 * P1); f[0] = -1.;      f[1] = x[P - i]; f[2] = 0;
    }
  }
  for (int fpt = 0; fpt < s->nFpts; fpt++) tnorm_face(s->nDims, fpt / s->nFptsFace, s->tNorm + 3 * fpt);
  /* getQptWeights, points.cpp:347-374 */
  const double *w1 = GL_WTS[P];
  for (int spt = 0; spt < s->nSpts; spt++) {
    int i = spt % P1, j = (spt / P1) % P1, k = spt / (P1 * P1);
    s->wts[spt] = (s->nDims == 3) ? w1[k] * w1[j] * w1[i] : w1[j] * w1[i];
  }
}

/* divVCJH_hex operators.cpp:1575-1619, divVCJH_quad 1544-1573 */
static double div_vcjh(const or_solver *s, int fpt, const double *loc) {
  int P = s->order, P1 = P + 1, n = P1;
  const double *x = s->spts1D;
  double eta = (s->p.vcjhScheme == OR_DG) ? 0. : compute_eta(s->p.vcjhScheme, P);
  if (s->nDims == 3) {
    int P12 = P1 * P1, f = fpt / P12, j = (fpt - P12 * f) / P1, i = fpt - P12 * f - P1 * j;
    switch (f) {
    case 0: return lagrange(x, n, loc[0], i) * lagrange(x, n, loc[1], j) * -dvcjh_1d(loc[2], 0, P, eta);
    case 1: return lagrange(x, n, loc[0], P - i) * lagrange(x, n, loc[1], j) * dvcjh_1d(loc[2], 1, P, eta);
    case 2: return lagrange(x, n, loc[1], i) * lagrange(x, n, loc[2], j) * -dvcjh_1d(loc[0], 0, P, eta);
    case 3: return lagrange(x, n, loc[1], P - i) * lagrange(x, n, loc[2], j) * dvcjh_1d(loc[0], 1, P, eta);
    case 4: return lagrange(x, n, loc[0], P - i) * lagrange(x, n, loc[2], j) * -dvcjh_1d(loc[1], 0, P, eta);
    default: return lagrange(x, n, loc[0], i) * lagrange(x, n, loc[2], j) * dvcjh_1d(loc[1], 1, P, eta);
    }
  } else {
    int i = fpt / P1, j = fpt - P1 * i;
    switch (i) {
    case 0: return -lagrange(x, n, loc[0], j) * dvcjh_1d(loc[1], 0, P, eta);
    case 1: return lagrange(x, n, loc[1], j) * dvcjh_1d(loc[0], 1, P, eta);
    case 2: return lagrange(x, n, loc[0], P - j) * dvcjh_1d(loc[1], 1, P, eta);
    default: return -lagrange(x, n, loc[1], P - j) * dvcjh_1d(loc[0], 0, P, eta);
    }
  }
}

static void setup_operators(or_solver *s) {
  int P1 = s->P1, nS = s->nSpts, nF = s->nFpts, D = s->nDims;
  const double *x = s->spts1D;
  /* setupExtrapolateSptsFpts, operators.cpp:78-111 */
  for (int fpt = 0; fpt < nF; fpt++) {
    const double *f = s->loc_fpts + 3 * fpt;
    for (int spt = 0; spt < nS; spt++) {
      double v;
      if (D == 3) {
        int k = spt / (P1 * P1), j = (spt - P1 * P1 * k) / P1, i = spt - P1 * j - P1 * P1 * k;
        v = lagrange(x, P1, f[0], i) * lagrange(x, P1, f[1], j) * lagrange(x, P1, f[2], k);
      } else {
        int i = spt % P1, j = spt / P1;
        v = lagrange(x, P1, f[0], i) * lagrange(x, P1, f[1], j);
      }
      s->opp_spts_to_fpts[fpt * nS + spt] = v;
    }
  }
  /* setupGradPts (string form, basis == pts == spts), operators.cpp:660-740 */
  for (int d = 0; d < D; d++)
    for (int g = 0; g < nS; g++)
      for (int b = 0; b < nS; b++) {
        double v;
        if (D == 3) {
          int kg = g / (P1 * P1), jg = (g - P1 * P1 * kg) / P1, ig = g - P1 * jg - P1 * P1 * kg;
          int kb = b / (P1 * P1), jb = (b - P1 * P1 * kb) / P1, ib = b - P1 * jb - P1 * P1 * kb;
          if (d == 0) v = dlagrange(x, P1, x[ig], ib) * lagrange(x, P1, x[jg], jb) * lagrange(x, P1, x[kg], kb);
          else if (d == 1) v = dlagrange(x, P1, x[jg], jb) * lagrange(x, P1, x[ig], ib) * lagrange(x, P1, x[kg], kb);
          else v = dlagrange(x, P1, x[kg], kb) * lagrange(x, P1, x[ig], ib) * lagrange(x, P1, x[jg], jb);
        } else {
          int ig = g % P1, jg = g / P1, ib = b % P1, jb = b / P1;
          if (d == 0) v = dlagrange(x, P1, x[ig], ib) * lagrange(x, P1, x[jg], jb);
          else v = dlagrange(x, P1, x[jg], jb) * lagrange(x, P1, x[ig], ib);
        }
        s->opp_grad_spts[((size_t)d * nS + g) * nS + b] = v;
      }
  /* setupExtrapolateFn, operators.cpp:807-889 */
  for (int d = 0; d < D; d++)
    for (int fpt = 0; fpt < nF; fpt++)
      for (int spt = 0; spt < nS; spt++)
        s->opp_extrapolateFn[((size_t)d * nF + fpt) * nS + spt] =
            s->opp_spts_to_fpts[fpt * nS + spt] * s->tNorm[3 * fpt + d];
  /* setupCorrection, operators.cpp:891-914 */
  for (int spt = 0; spt < nS; spt++)
    for (int fpt = 0; fpt < nF; fpt++)
      s->opp_correction[spt * nF + fpt] = div_vcjh(s, fpt, s->loc_spts + 3 * spt);
  /* setupCorrectGradU, operators.cpp:1008-1092 */
  for (int d = 0; d < D; d++)
    for (int fpt = 0; fpt < nF; fpt++)
      for (int spt = 0; spt < nS; spt++)
        s->opp_correctU[((size_t)d * nS + spt) * nF + fpt] = s->opp_correction[spt * nF + fpt] * s->tNorm[3 * fpt + d];
}

/* ------------------------------------------------------------------------ */
/* shape functions, src/funcs.cpp: quad (4 nodes, shape_quad 121-176) and    */
/* hex (shape_hex 178-566, dshape_hex 680-1607): 8 and (N+1)^3 nodes by the  */
/* Lagrange recursion over nested shells, 20 nodes serendipity; Gmsh order    */
/* ------------------------------------------------------------------------ */
/* one node of the Lagrange hex: lattice indices along xi, eta, mu */
typedef struct { int a, b, c; } or_node;

/* the shell recursion of shape_hex (funcs.cpp:213-566), as the list of each
 * node's lattice indices in node order; returns the node count */
static int hex_nodes(int nSide, or_node *nd) {
  int n = 0;
  const int nLevels = nSide / 2, isOdd = nSide % 2, h = nSide / 2;
#define PUT(x, y, z) (nd[n].a = (x), nd[n].b = (y), nd[n].c = (z), n++)
  for (int i = 0; i < nLevels; i++) {
    int i2 = (nSide - 1) - i;
    PUT(i, i, i); PUT(i2, i, i); PUT(i2, i2, i); PUT(i, i2, i);
    PUT(i, i, i2); PUT(i2, i, i2); PUT(i2, i2, i2); PUT(i, i2, i2);
    int nSide2 = nSide - 2 * (i + 1);
    /* edges: 12 blocks of nSide2, written slot by slot */
    or_node *e = nd + n;
    for (int j = 0; j < nSide2; j++) {
#define E(slot, x, y, z) (e[(slot) * nSide2 + j].a = (x), e[(slot) * nSide2 + j].b = (y), e[(slot) * nSide2 + j].c = (z))
      E(0, i + 1 + j, i, i);      E(3, i2, i + 1 + j, i);      E(5, i2 - 1 - j, i2, i);  E(1, i, i + 1 + j, i);
      E(2, i, i, i + 1 + j);      E(4, i2, i, i + 1 + j);      E(6, i2, i2, i + 1 + j);  E(7, i, i2, i + 1 + j);
      E(8, i + 1 + j, i, i2);     E(10, i2, i + 1 + j, i2);    E(11, i2 - 1 - j, i2, i2); E(9, i, i + 1 + j, i2);
#undef E
    }
    n += 12 * nSide2;
    int nLevels2 = nSide2 / 2, isOdd2 = nSide2 % 2;
    /* faces: bottom (mu = i), front (eta = i), left (xi = i), right (xi = i2),
     * back (eta = i2), top (mu = i2); each ring: 4 corners, then 4 edges of nSide3 */
    for (int f = 0; f < 6; f++) {
      for (int j0 = 0; j0 < nLevels2; j0++) {
        int j = j0 + i + 1, j2 = i + 1 + (nSide2 - 1) - j0, nSide3 = nSide2 - 2 * (j0 + 1);
        switch (f) {
        case 0: PUT(j, j, i); PUT(j, j2, i); PUT(j2, j2, i); PUT(j2, j, i); break;
        case 1: PUT(j, i, j); PUT(j2, i, j); PUT(j2, i, j2); PUT(j, i, j2); break;
        case 2: PUT(i, j, j); PUT(i, j, j2); PUT(i, j2, j2); PUT(i, j2, j); break;
        case 3: PUT(i2, j, j); PUT(i2, j2, j); PUT(i2, j2, j2); PUT(i2, j, j2); break;
        case 4: PUT(j2, i2, j); PUT(j, i2, j); PUT(j, i2, j2); PUT(j2, i2, j2); break;
        default: PUT(j, j, i2); PUT(j2, j, i2); PUT(j2, j2, i2); PUT(j, j2, i2); break;
        }
        or_node *q = nd + n;
        for (int k = 0; k < nSide3; k++) {
          int up = j + 1 + k, dn = j2 - 1 - k;
#define Q(slot, x, y, z) (q[(slot) * nSide3 + k].a = (x), q[(slot) * nSide3 + k].b = (y), q[(slot) * nSide3 + k].c = (z))
          switch (f) {
          case 0: Q(0, j, up, i); Q(1, up, j2, i); Q(2, j2, dn, i); Q(3, dn, j, i); break;
          case 1: Q(0, up, i, j); Q(1, j2, i, up); Q(2, dn, i, j2); Q(3, j, i, dn); break;
          case 2: Q(0, i, j, up); Q(1, i, up, j2); Q(2, i, j2, dn); Q(3, i, dn, j); break;
          case 3: Q(0, i2, up, j); Q(1, i2, j2, up); Q(2, i2, dn, j2); Q(3, i2, j, dn); break;
          case 4: Q(0, dn, i2, j); Q(1, j, i2, up); Q(2, up, i2, j2); Q(3, j2, i2, dn); break;
          default: Q(0, up, j, i2); Q(1, j2, up, i2); Q(2, dn, j2, i2); Q(3, j, dn, i2); break;
          }
#undef Q
        }
        n += 4 * nSide3;
      }
      if (isOdd2) {
        switch (f) {
        case 0: PUT(h, h, i); break;
        case 1: PUT(h, i, h); break;
        case 2: PUT(i, h, h); break;
        case 3: PUT(i2, h, h); break;
        case 4: PUT(h, i2, h); break;
        default: PUT(h, h, i2); break;
        }
      }
    }
  }
  if (isOdd) PUT(h, h, h);
#undef PUT
  return n;
}

static const double XL[2] = {-1., 1.};
/* shape values sh[node] and derivatives dsh[node][3] (either may be NULL) */
static void shape_dshape_n(int nDims, int nNodes, const double *r, double *sh, double *dsh) {
  if (nDims == 2) {
    static const int QI[4] = {0, 1, 1, 0}, QJ[4] = {0, 0, 1, 1};
    for (int n = 0; n < 4; n++) {
      double la = lagrange(XL, 2, r[0], QI[n]), lb = lagrange(XL, 2, r[1], QJ[n]);
      if (sh) sh[n] = la * lb;
      if (dsh) {
        dsh[3 * n + 0] = dlagrange(XL, 2, r[0], QI[n]) * lb;
        dsh[3 * n + 1] = la * dlagrange(XL, 2, r[1], QJ[n]);
        dsh[3 * n + 2] = 0;
      }
    }
    return;
  }
  if (nNodes == 20) {
    static const double XI[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, ETA[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                        MU[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double xi = r[0], eta = r[1], mu = r[2];
    for (int i = 0; i < 8; i++) {
      if (sh) sh[i] = .125 * (1 + xi * XI[i]) * (1 + eta * ETA[i]) * (1 + mu * MU[i]) *
                      (xi * XI[i] + eta * ETA[i] + mu * MU[i] - 2);
      if (dsh) {
        dsh[3 * i + 0] = .125 * XI[i] * (1 + eta * ETA[i]) * (1 + mu * MU[i]) *
                         (2 * xi * XI[i] + eta * ETA[i] + mu * MU[i] - 1);
        dsh[3 * i + 1] = .125 * ETA[i] * (1 + xi * XI[i]) * (1 + mu * MU[i]) *
                         (xi * XI[i] + 2 * eta * ETA[i] + mu * MU[i] - 1);
        dsh[3 * i + 2] = .125 * MU[i] * (1 + xi * XI[i]) * (1 + eta * ETA[i]) *
                         (xi * XI[i] + eta * ETA[i] + 2 * mu * MU[i] - 1);
      }
    }
    /* edge nodes (funcs.cpp:197-211, 705-749) */
    if (sh) {
      sh[8] = .25 * (1 - xi * xi) * (1 - eta) * (1 - mu);
      sh[10] = .25 * (1 - xi * xi) * (1 + eta) * (1 - mu);
      sh[16] = .25 * (1 - xi * xi) * (1 - eta) * (1 + mu);
      sh[18] = .25 * (1 - xi * xi) * (1 + eta) * (1 + mu);
      sh[9] = .25 * (1 - eta * eta) * (1 + xi) * (1 - mu);
      sh[11] = .25 * (1 - eta * eta) * (1 - xi) * (1 - mu);
      sh[17] = .25 * (1 - eta * eta) * (1 + xi) * (1 + mu);
      sh[19] = .25 * (1 - eta * eta) * (1 - xi) * (1 + mu);
      sh[12] = .25 * (1 - mu * mu) * (1 - xi) * (1 - eta);
      sh[13] = .25 * (1 - mu * mu) * (1 + xi) * (1 - eta);
      sh[14] = .25 * (1 - mu * mu) * (1 + xi) * (1 + eta);
      sh[15] = .25 * (1 - mu * mu) * (1 - xi) * (1 + eta);
    }
    if (dsh) {
      double *o = dsh;
      o[3 * 8 + 0] = -.5 * xi * (1 - eta) * (1 - mu);  o[3 * 8 + 1] = -.25 * (1 - xi * xi) * (1 - mu);
      o[3 * 8 + 2] = -.25 * (1 - xi * xi) * (1 - eta);
      o[3 * 10 + 0] = -.5 * xi * (1 + eta) * (1 - mu); o[3 * 10 + 1] = .25 * (1 - xi * xi) * (1 - mu);
      o[3 * 10 + 2] = -.25 * (1 - xi * xi) * (1 + eta);
      o[3 * 16 + 0] = -.5 * xi * (1 - eta) * (1 + mu); o[3 * 16 + 1] = -.25 * (1 - xi * xi) * (1 + mu);
      o[3 * 16 + 2] = .25 * (1 - xi * xi) * (1 - eta);
      o[3 * 18 + 0] = -.5 * xi * (1 + eta) * (1 + mu); o[3 * 18 + 1] = .25 * (1 - xi * xi) * (1 + mu);
      o[3 * 18 + 2] = .25 * (1 - xi * xi) * (1 + eta);
      o[3 * 9 + 1] = -.5 * eta * (1 + xi) * (1 - mu);  o[3 * 9 + 0] = .25 * (1 - eta * eta) * (1 - mu);
      o[3 * 9 + 2] = -.25 * (1 - eta * eta) * (1 + xi);
      o[3 * 11 + 1] = -.5 * eta * (1 - xi) * (1 - mu); o[3 * 11 + 0] = -.25 * (1 - eta * eta) * (1 - mu);
      o[3 * 11 + 2] = -.25 * (1 - eta * eta) * (1 - xi);
      o[3 * 17 + 1] = -.5 * eta * (1 + xi) * (1 + mu); o[3 * 17 + 0] = .25 * (1 - eta * eta) * (1 + mu);
      o[3 * 17 + 2] = .25 * (1 - eta * eta) * (1 + xi);
      o[3 * 19 + 1] = -.5 * eta * (1 - xi) * (1 + mu); o[3 * 19 + 0] = -.25 * (1 - eta * eta) * (1 + mu);
      o[3 * 19 + 2] = .25 * (1 - eta * eta) * (1 - xi);
      o[3 * 12 + 2] = -.5 * mu * (1 - xi) * (1 - eta); o[3 * 12 + 0] = -.25 * (1 - mu * mu) * (1 - eta);
      o[3 * 12 + 1] = -.25 * (1 - mu * mu) * (1 - xi);
      o[3 * 13 + 2] = -.5 * mu * (1 + xi) * (1 - eta); o[3 * 13 + 0] = .25 * (1 - mu * mu) * (1 - eta);
      o[3 * 13 + 1] = -.25 * (1 - mu * mu) * (1 + xi);
      o[3 * 14 + 2] = -.5 * mu * (1 + xi) * (1 + eta); o[3 * 14 + 0] = .25 * (1 - mu * mu) * (1 + eta);
      o[3 * 14 + 1] = .25 * (1 - mu * mu) * (1 + xi);
      o[3 * 15 + 2] = -.5 * mu * (1 - xi) * (1 + eta); o[3 * 15 + 0] = -.25 * (1 - mu * mu) * (1 + eta);
      o[3 * 15 + 1] = .25 * (1 - mu * mu) * (1 - xi);
    }
    return;
  }
  int nSide = (int)lround(cbrt((double)nNodes));
  double xl[6];
  for (int i = 0; i < nSide; i++) xl[i] = -1. + i * (2. / (nSide - 1));
  or_node nd[216];
  hex_nodes(nSide, nd);
  for (int n = 0; n < nNodes; n++) {
    double la = lagrange(xl, nSide, r[0], nd[n].a), lb = lagrange(xl, nSide, r[1], nd[n].b),
           lc = lagrange(xl, nSide, r[2], nd[n].c);
    if (sh) sh[n] = la * lb * lc;
    if (dsh) {
      dsh[3 * n + 0] = dlagrange(xl, nSide, r[0], nd[n].a) * lb * lc;
      dsh[3 * n + 1] = la * dlagrange(xl, nSide, r[1], nd[n].b) * lc;
      dsh[3 * n + 2] = la * lb * dlagrange(xl, nSide, r[2], nd[n].c);
    }
  }
}
/* linear elements (8-node hex, 4-node quad) */
static void shape_dshape(int nDims, const double *r, double *sh, double *dsh) {
  shape_dshape_n(nDims, nDims == 3 ? 8 : 4, r, sh, dsh);
}

/* test hook: shape values / derivatives of an nNodes element at r[3] */
void or_shape(int nDims, int nNodes, const double *r, double *sh, double *dsh) {
  shape_dshape_n(nDims, nNodes, r, sh, dsh);
}
/* test hook: lattice indices of the nodes of the nSide^3 Lagrange hex */
int or_hex_lattice(int nSide, int *abc) {
  or_node nd[216];
  if (nSide < 2 || nSide > 6) return -1;
  int n = hex_nodes(nSide, nd);
  for (int i = 0; i < n; i++) {
    abc[3 * i] = nd[i].a;
    abc[3 * i + 1] = nd[i].b;
    abc[3 * i + 2] = nd[i].c;
  }
  return n;
}

/* setPosSptsFpts (solver.cpp:904-951) + calcTransforms (solver.cpp:1227-1356) */
static int setup_geometry(or_solver *s, const double *xv, const int *c2v) {
  int D = s->nDims, nN = s->nNodes, nE = s->nEles, nS = s->nSpts, nF = s->nFpts;
  for (int n = 0; n < nN; n++)
    for (int e = 0; e < nE; e++)
      for (int d = 0; d < D; d++) s->nodes[IDX3(n, e, d, nE, D)] = xv[3 * c2v[e * nN + n] + d];
  double *shs = xcalloc((size_t)nS * nN, 8), *dshs = xcalloc((size_t)D * nS * nN, 8);
  double *dshf = xcalloc((size_t)D * nF * nN, 8);
  double tmp[216], dtmp[3 * 216];
  for (int spt = 0; spt < nS; spt++) {
    shape_dshape_n(D, nN, s->loc_spts + 3 * spt, tmp, dtmp);
    for (int n = 0; n < nN; n++) {
      shs[spt * nN + n] = tmp[n];
      for (int d = 0; d < D; d++) dshs[((size_t)d * nS + spt) * nN + n] = dtmp[3 * n + d];
    }
  }
  for (int fpt = 0; fpt < nF; fpt++) {
    shape_dshape_n(D, nN, s->loc_fpts + 3 * fpt, tmp, dtmp);
    for (int n = 0; n < nN; n++)
      for (int d = 0; d < D; d++) dshf[((size_t)d * nF + fpt) * nN + n] = dtmp[3 * n + d];
  }
  int n = nE * D;
  oracle_dgemm(nS, n, nN, 1.0, shs, s->nodes, 0.0, s->pos_spts);
  for (int d = 0; d < D; d++) {
    oracle_dgemm(nS, n, nN, 1.0, dshs + (size_t)d * nS * nN, s->nodes, 0.0, s->Jac_spts + (size_t)d * nS * nE * D);
    oracle_dgemm(nF, n, nN, 1.0, dshf + (size_t)d * nF * nN, s->nodes, 0.0, s->Jac_fpts + (size_t)d * nF * nE * D);
  }
  free(shs); free(dshs); free(dshf);
#define JS(d1, p, e, d2) s->Jac_spts[(((size_t)(d1) * nS + (p)) * nE + (e)) * D + (d2)]
#define GS(d1, p, e, d2) s->JGinv_spts[(((size_t)(d1) * nS + (p)) * nE + (e)) * D + (d2)]
#define JF(d1, p, e, d2) s->Jac_fpts[(((size_t)(d1) * nF + (p)) * nE + (e)) * D + (d2)]
#define GF(d1, p, e, d2) s->JGinv_fpts[(((size_t)(d1) * nF + (p)) * nE + (e)) * D + (d2)]
  for (int spt = 0; spt < nS; spt++)
    for (int e = 0; e < nE; e++) {
      double det;
      if (D == 2) {
        det = JS(0, spt, e, 0) * JS(1, spt, e, 1) - JS(0, spt, e, 1) * JS(1, spt, e, 0);
        GS(0, spt, e, 0) = JS(1, spt, e, 1);
        GS(0, spt, e, 1) = -JS(1, spt, e, 0);
        GS(1, spt, e, 0) = -JS(0, spt, e, 1);
        GS(1, spt, e, 1) = JS(0, spt, e, 0);
      } else {
        double xr = JS(0, spt, e, 0), xs = JS(1, spt, e, 0), xt = JS(2, spt, e, 0);
        double yr = JS(0, spt, e, 1), ys = JS(1, spt, e, 1), yt = JS(2, spt, e, 1);
        double zr = JS(0, spt, e, 2), zs = JS(1, spt, e, 2), zt = JS(2, spt, e, 2);
        det = xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
        GS(0, spt, e, 0) = ys * zt - yt * zs; GS(0, spt, e, 1) = xt * zs - xs * zt; GS(0, spt, e, 2) = xs * yt - xt * ys;
        GS(1, spt, e, 0) = yt * zr - yr * zt; GS(1, spt, e, 1) = xr * zt - xt * zr; GS(1, spt, e, 2) = xt * yr - xr * yt;
        GS(2, spt, e, 0) = yr * zs - ys * zr; GS(2, spt, e, 1) = xs * zr - xr * zs; GS(2, spt, e, 2) = xr * ys - xs * yr;
      }
      s->detJac_spts[(size_t)spt * nE + e] = det;
      if (det < 0) { set_err("Negative Jacobian at solution points."); return -1; }
    }
  for (int fpt = 0; fpt < nF; fpt++)
    for (int e = 0; e < nE; e++) {
      double det;
      if (D == 2) {
        det = JF(0, fpt, e, 0) * JF(1, fpt, e, 1) - JF(0, fpt, e, 1) * JF(1, fpt, e, 0);
        GF(0, fpt, e, 0) = JF(1, fpt, e, 1);
        GF(0, fpt, e, 1) = -JF(1, fpt, e, 0);
        GF(1, fpt, e, 0) = -JF(0, fpt, e, 1);
        GF(1, fpt, e, 1) = JF(0, fpt, e, 0);
      } else {
        double xr = JF(0, fpt, e, 0), xs = JF(1, fpt, e, 0), xt = JF(2, fpt, e, 0);
        double yr = JF(0, fpt, e, 1), ys = JF(1, fpt, e, 1), yt = JF(2, fpt, e, 1);
        double zr = JF(0, fpt, e, 2), zs = JF(1, fpt, e, 2), zt = JF(2, fpt, e, 2);
        det = xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
        GF(0, fpt, e, 0) = ys * zt - yt * zs; GF(0, fpt, e, 1) = xt * zs - xs * zt; GF(0, fpt, e, 2) = xs * yt - xt * ys;
        GF(1, fpt, e, 0) = yt * zr - yr * zt; GF(1, fpt, e, 1) = xr * zt - xt * zr; GF(1, fpt, e, 2) = xt * yr - xr * yt;
        GF(2, fpt, e, 0) = yr * zs - ys * zr; GF(2, fpt, e, 1) = xs * zr - xr * zs; GF(2, fpt, e, 2) = xr * ys - xs * yr;
      }
      s->detJac_fpts[(size_t)fpt * nE + e] = det;
      double *nrm = s->norm_fpts + ((size_t)fpt * nE + e) * D;
      for (int d1 = 0; d1 < D; d1++) {
        nrm[d1] = 0.;
        for (int d2 = 0; d2 < D; d2++) nrm[d1] += GF(d2, fpt, e, d1) * s->tNorm[3 * fpt + d2];
      }
      double dA = 0;
      for (int d = 0; d < D; d++) dA += nrm[d] * nrm[d];
      dA = sqrt(dA);
      if (fabs(dA) < 1e-10) {
        dA = 0.;
        for (int d = 0; d < D; d++) nrm[d] = 0;
      } else {
        for (int d = 0; d < D; d++) nrm[d] /= dA;
      }
      s->dA_fpts[(size_t)fpt * nE + e] = dA;
    }
#undef JS
#undef GS
#undef JF
#undef GF
  return 0;
}

/* ------------------------------------------------------------------------ */
/* face connectivity (src/geo.cpp:75-267, 322-446, 789-917; intFace.cpp:6-71) */
/* ------------------------------------------------------------------------ */
/* local face -> element vertex lists; matching list ct2fv (geo.cpp:86-91) and
 * orientation list used by compareOrientation (geo.cpp:804-900) */
static const int HEX_FV[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {3, 0, 4, 7}, {2, 1, 5, 6}, {1, 0, 4, 5}, {3, 2, 6, 7}};
static const int HEX_ORIENT[6][4] = {{0, 1, 2, 3}, {5, 4, 7, 6}, {0, 3, 7, 4}, {2, 1, 5, 6}, {1, 0, 4, 5}, {3, 2, 6, 7}};
static const int QUAD_FV[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};

typedef struct { int v[4]; int e, f; } facekey;
static int nfv_g;
static int cmp_facekey(const void *a, const void *b) {
  const facekey *x = a, *y = b;
  for (int i = 0; i < nfv_g; i++)
    if (x->v[i] != y->v[i]) return x->v[i] < y->v[i] ? -1 : 1;
  if (x->e != y->e) return x->e < y->e ? -1 : 1;
  return x->f < y->f ? -1 : (x->f > y->f);
}
static void sort_int(int *a, int n) {
  for (int i = 1; i < n; i++) {
    int t = a[i], j = i - 1;
    while (j >= 0 && a[j] > t) { a[j + 1] = a[j]; j--; }
    a[j + 1] = t;
  }
}
static int cmp_int(const void *a, const void *b) { int x = *(const int *)a, y = *(const int *)b; return (x > y) - (x < y); }

static int setup_faces(or_solver *s, int nVerts, const int *c2v, const int *vmap, int nBounds,
                       const int *bcType, const int *bndOff, const int *bndPts) {
  int D = s->nDims, nE = s->nEles, nN = s->nNodes, nfe = s->nFacesEle, nfv = (D == 3) ? 4 : 2;
  int nk = nE * nfe;
  facekey *keys = xcalloc(nk, sizeof(facekey));
#define CV(e, n) (vmap ? vmap[c2v[(e) * nN + (n)]] : c2v[(e) * nN + (n)])
  for (int e = 0; e < nE; e++)
    for (int f = 0; f < nfe; f++) {
      facekey *k = keys + e * nfe + f;
      for (int i = 0; i < nfv; i++) k->v[i] = (D == 3) ? CV(e, HEX_FV[f][i]) : CV(e, QUAD_FV[f][i]);
      sort_int(k->v, nfv);
      k->e = e; k->f = f;
    }
  nfv_g = nfv;
  qsort(keys, nk, sizeof(facekey), cmp_facekey);
  /* group identical vertex lists: 2 owners -> interior, 1 -> boundary (geo.cpp:131-159) */
  int cap = nk;
  s->fL = xcalloc(cap, sizeof(int)); s->flocL = xcalloc(cap, sizeof(int));
  s->fR = xcalloc(cap, sizeof(int)); s->flocR = xcalloc(cap, sizeof(int));
  s->frot = xcalloc(cap, sizeof(int)); s->fbc = xcalloc(cap, sizeof(int));
  /* two passes: interior faces first, then boundary faces (face list order of
   * geo::setupElesFaces, geo.cpp:369-445) */
  int nF = 0;
  int *isb = xcalloc(nk, sizeof(int));
  for (int pass = 0; pass < 2; pass++) {
    for (int a = 0; a < nk;) {
      int b = a + 1;
      while (b < nk && memcmp(keys[b].v, keys[a].v, nfv * sizeof(int)) == 0) b++;
      int cnt = b - a;
      if (cnt > 2) { set_err("More than 2 cells for face"); free(keys); free(isb); return -1; }
      if (pass == 0 && cnt == 2) {
        /* keys sorted by (vertices, e, f): first occurrence = left (geo.cpp:222-233) */
        s->fL[nF] = keys[a].e; s->flocL[nF] = keys[a].f;
        s->fR[nF] = keys[a + 1].e; s->flocR[nF] = keys[a + 1].f;
        s->fbc[nF] = OR_BC_NONE;
        nF++;
      } else if (pass == 1 && cnt == 1) {
        int bc = OR_BC_NONE;
        /* boundary matching: all face vertices in a boundary's point set (geo.cpp:163-185) */
        for (int bnd = 0; bnd < nBounds && bc == OR_BC_NONE; bnd++) {
          int ok = 1;
          for (int i = 0; i < nfv && ok; i++) {
            int v = keys[a].v[i], found = 0;
            for (int q = bndOff[bnd]; q < bndOff[bnd + 1]; q++) {
              int w = vmap ? vmap[bndPts[q]] : bndPts[q];
              if (w == v) { found = 1; break; }
            }
            ok = found;
          }
          if (ok) bc = bcType[bnd];
        }
        if (bc == OR_BC_NONE) { set_err("Boundary face not matched to any boundary"); free(keys); free(isb); return -1; }
        if (bc == OR_PERIODIC) { set_err("Boundary condition not recognized."); free(keys); free(isb); return -1; }
        s->fL[nF] = keys[a].e; s->flocL[nF] = keys[a].f;
        s->fR[nF] = -1; s->flocR[nF] = -1; s->fbc[nF] = bc;
        nF++;
      }
      a = b;
    }
    if (pass == 0) s->nIntFaces = nF;
  }
  s->nFaces = nF;
  free(keys); free(isb);
  /* relRot: compareOrientation, geo.cpp:789-917 */
  for (int f = 0; f < s->nIntFaces; f++) {
    if (D == 2) { s->frot[f] = 1; continue; }
    int e1 = s->fL[f], f1 = s->flocL[f], e2 = s->fR[f], f2 = s->flocR[f];
    int v2 = CV(e2, HEX_ORIENT[f2][0]), rot = -1;
    for (int i = 0; i < 4; i++)
      if (CV(e1, HEX_ORIENT[f1][i]) == v2) { rot = i; break; }
    if (rot < 0) { set_err("Internal faces improperly matched."); return -1; }
    s->frot[f] = rot;
  }
#undef CV
  /* fptR: intFace::setupRightState, intFace.cpp:29-65 */
  int nfp = s->nFptsFace, P = s->order;
  s->fptR = xcalloc((size_t)nF * nfp, sizeof(int));
  for (int f = 0; f < s->nIntFaces; f++) {
    int *m = s->fptR + (size_t)f * nfp;
    if (D == 2) {
      int start = s->flocR[f] * nfp + nfp, end = s->flocR[f] * nfp, q = 0;
      for (int i = start - 1; i >= end; i--) m[q++] = i;
    } else {
      int start = nfp * s->flocR[f], end = nfp * (s->flocR[f] + 1);
      for (int i = 0; i < nfp; i++) {
        int ifpt = i % (P + 1), jfpt = i / (P + 1);
        switch (s->frot[f]) {
        case 0: m[i] = start + ifpt * (P + 1) + jfpt; break;
        case 1: m[i] = start + P - ifpt + jfpt * (P + 1); break;
        case 2: m[i] = end - 1 - (ifpt * (P + 1) + jfpt); break;
        default: m[i] = end - (P + 1) * (jfpt + 1) + ifpt; break;
        }
      }
    }
  }
  (void)nVerts; (void)cmp_int;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* creation                                                                   */
/* ------------------------------------------------------------------------ */
or_solver *or_create(const or_params *p, int nVerts, const double *xv, int nEles, int nNodes, const int *c2v,
                     const int *vmap, int nBounds, const int *bcType, const int *bndOff, const int *bndPts) {
  if (p->nDims != 2 && p->nDims != 3) { set_err("nDims must be 2 or 3"); return NULL; }
  if (p->order < 0 || p->order > 10) { set_err("order out of range"); return NULL; }
  {
    int ok = p->nDims == 2 ? nNodes == 4 : (nNodes == 8 || nNodes == 20 || nNodes == 27 || nNodes == 64 ||
                                            nNodes == 125 || nNodes == 216);
    if (!ok) { set_err("hexes of 8, 20, 27, 64, 125 or 216 nodes, or 4-node quads"); return NULL; }
  }
  if (p->nDims == 3 && p->equation == OR_NAVIER_STOKES && p->riemannType == 1) {
    set_err("Roe not implemented in 3D"); /* face.cpp:362-363 */
    return NULL;
  }
  or_solver *s = xcalloc(1, sizeof(or_solver));
  s->p = *p;
  int D = s->nDims = p->nDims, P = s->order = p->order, P1 = s->P1 = P + 1;
  s->nSpts = (D == 3) ? P1 * P1 * P1 : P1 * P1;
  s->nFptsFace = (D == 3) ? P1 * P1 : P1;
  s->nFacesEle = 2 * D;
  s->nFpts = s->nFacesEle * s->nFptsFace;
  s->nFields = (p->equation == OR_ADVECTION_DIFFUSION) ? 1 : D + 2;
  s->nEles = nEles;
  s->nNodes = nNodes;
  /* RK tables, input.cpp:424-461 */
  memset(s->RKa, 0, sizeof s->RKa); memset(s->RKb, 0, sizeof s->RKb);
  switch (p->timeType) {
  case 0: s->nRK = 1; s->RKa[0] = 0; s->RKb[0] = 1; break;
  case 2: s->nRK = 2; s->RKa[0] = 0; s->RKa[1] = 1.; s->RKb[0] = 1. / 2.; s->RKb[1] = 1. / 2.; break;
  case 4: s->nRK = 4; s->RKa[0] = 0.; s->RKa[1] = .5; s->RKa[2] = .5; s->RKa[3] = 1.;
    s->RKb[0] = 1. / 6.; s->RKb[1] = 1. / 3.; s->RKb[2] = 1. / 3.; s->RKb[3] = 1. / 6.; break;
  case 5: s->nRK = 4; s->RKa[0] = 0.; s->RKa[1] = 1. / 4.; s->RKa[2] = 1. / 3.; s->RKa[3] = 1. / 2.; s->RKa[4] = 1.0; break;
  case 6: s->nRK = 4;
    if (P <= 2) { s->RKa[1] = .153; s->RKa[2] = .442; s->RKa[3] = .931; s->RKa[4] = 1.0; }
    else { s->RKa[1] = .158; s->RKa[2] = .477; s->RKa[3] = 1.; s->RKa[4] = 1.; }
    break;
  default: set_err("Time-Stepping type not supported."); free(s); return NULL;
  }
  int nS = s->nSpts, nF = s->nFpts, nE = nEles, nK = s->nFields;
  s->spts1D = xcalloc(P1, 8);
  if (pts1d(p->sptsType, P, s->spts1D)) { free(s); return NULL; }
  s->wts = xcalloc(nS, 8);
  s->loc_spts = xcalloc(3 * nS, 8); s->loc_fpts = xcalloc(3 * nF, 8); s->tNorm = xcalloc(3 * nF, 8);
  s->opp_spts_to_fpts = xcalloc((size_t)nF * nS, 8);
  s->opp_grad_spts = xcalloc((size_t)D * nS * nS, 8);
  s->opp_extrapolateFn = xcalloc((size_t)D * nF * nS, 8);
  s->opp_correction = xcalloc((size_t)nS * nF, 8);
  s->opp_correctU = xcalloc((size_t)D * nS * nF, 8);
  setup_points(s);
  setup_operators(s);
  s->nodes = xcalloc((size_t)nNodes * nE * D, 8);
  s->pos_spts = xcalloc((size_t)nS * nE * D, 8);
  s->Jac_spts = xcalloc((size_t)D * nS * nE * D, 8);
  s->JGinv_spts = xcalloc((size_t)D * nS * nE * D, 8);
  s->detJac_spts = xcalloc((size_t)nS * nE, 8);
  s->Jac_fpts = xcalloc((size_t)D * nF * nE * D, 8);
  s->JGinv_fpts = xcalloc((size_t)D * nF * nE * D, 8);
  s->detJac_fpts = xcalloc((size_t)nF * nE, 8);
  s->norm_fpts = xcalloc((size_t)nF * nE * D, 8);
  s->dA_fpts = xcalloc((size_t)nF * nE, 8);
  if (setup_geometry(s, xv, c2v)) { or_destroy(s); return NULL; }
  if (setup_faces(s, nVerts, c2v, vmap, nBounds, bcType, bndOff, bndPts)) { or_destroy(s); return NULL; }
  s->U_spts = xcalloc((size_t)nS * nE * nK, 8);
  s->U_fpts = xcalloc((size_t)nF * nE * nK, 8);
  s->F_spts = xcalloc((size_t)D * nS * nE * nK, 8);
  s->Fn_fpts = xcalloc((size_t)nF * nE * nK, 8);
  s->disFn_fpts = xcalloc((size_t)nF * nE * nK, 8);
  s->U0 = xcalloc((size_t)nS * nE * nK, 8);
  for (int r = 0; r < s->nRK; r++) s->divF[r] = xcalloc((size_t)nS * nE * nK, 8);
  s->faceFn = xcalloc((size_t)s->nFaces * s->nFptsFace * nK, 8);
  if (p->viscous) {
    s->dU_spts = xcalloc((size_t)D * nS * nE * nK, 8);
    s->dU_fpts = xcalloc((size_t)D * nF * nE * nK, 8);
    s->dUc_fpts = xcalloc((size_t)nF * nE * nK, 8);
    s->faceUC = xcalloc((size_t)s->nFaces * s->nFptsFace * nK, 8);
  }
  return s;
}

void or_destroy(or_solver *s) {
  if (!s) return;
  double *ptrs[] = {s->spts1D, s->wts, s->loc_spts, s->loc_fpts, s->tNorm, s->opp_spts_to_fpts, s->opp_grad_spts,
                    s->opp_extrapolateFn, s->opp_correction, s->opp_correctU, s->nodes, s->pos_spts, s->Jac_spts,
                    s->JGinv_spts, s->detJac_spts, s->Jac_fpts, s->JGinv_fpts, s->detJac_fpts, s->norm_fpts,
                    s->dA_fpts, s->U_spts, s->U_fpts, s->F_spts, s->Fn_fpts, s->disFn_fpts, s->U0, s->dU_spts,
                    s->dU_fpts, s->dUc_fpts, s->faceFn, s->faceUC};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
  for (int r = 0; r < 4; r++) free(s->divF[r]);
  free(s->dtEle);
  free(s->waveSp);
  int *iptrs[] = {s->fL, s->flocL, s->fR, s->flocR, s->frot, s->fbc, s->fptR};
  for (size_t i = 0; i < sizeof iptrs / sizeof iptrs[0]; i++) free(iptrs[i]);
  free(s);
}

void or_sizes(const or_solver *s, int *o) {
  o[0] = s->nSpts; o[1] = s->nFpts; o[2] = s->nEles; o[3] = s->nFields;
  o[4] = s->nDims; o[5] = s->nFaces; o[6] = s->nIntFaces; o[7] = s->nRK;
}

int or_get_operator(const or_solver *s, int which, double *out) {
  size_t nS = s->nSpts, nF = s->nFpts, D = s->nDims;
  switch (which) {
  case 0: memcpy(out, s->opp_spts_to_fpts, nF * nS * 8); return 0;
  case 1: memcpy(out, s->opp_grad_spts, D * nS * nS * 8); return 0;
  case 2: memcpy(out, s->opp_extrapolateFn, D * nF * nS * 8); return 0;
  case 3: memcpy(out, s->opp_correction, nS * nF * 8); return 0;
  case 4: memcpy(out, s->opp_correctU, D * nS * nF * 8); return 0;
  case 5: memcpy(out, s->loc_spts, nS * 3 * 8); return 0;
  case 6: memcpy(out, s->loc_fpts, nF * 3 * 8); return 0;
  case 7: memcpy(out, s->spts1D, (s->order + 1) * 8); return 0;
  case 8: memcpy(out, s->wts, nS * 8); return 0;
  }
  return -1;
}

int or_get_geometry(const or_solver *s, int which, double *out) {
  size_t nS = s->nSpts, nF = s->nFpts, D = s->nDims, nE = s->nEles;
  switch (which) {
  case 0: memcpy(out, s->JGinv_spts, D * nS * nE * D * 8); return 0;
  case 1: memcpy(out, s->detJac_spts, nS * nE * 8); return 0;
  case 2: memcpy(out, s->norm_fpts, nF * nE * D * 8); return 0;
  case 3: memcpy(out, s->dA_fpts, nF * nE * 8); return 0;
  case 4: memcpy(out, s->pos_spts, nS * nE * D * 8); return 0;
  case 5: memcpy(out, s->Jac_spts, D * nS * nE * D * 8); return 0;
  }
  return -1;
}

int or_get_faces(const or_solver *s, int *o6, int *fptR) {
  for (int f = 0; f < s->nFaces; f++) {
    o6[6 * f + 0] = s->fL[f]; o6[6 * f + 1] = s->flocL[f]; o6[6 * f + 2] = s->fR[f];
    o6[6 * f + 3] = s->flocR[f]; o6[6 * f + 4] = s->frot[f]; o6[6 * f + 5] = s->fbc[f];
  }
  if (fptR) memcpy(fptR, s->fptR, (size_t)s->nFaces * s->nFptsFace * sizeof(int));
  return 0;
}

void or_set_state(or_solver *s, const double *U) { memcpy(s->U_spts, U, (size_t)s->nSpts * s->nEles * s->nFields * 8); }
void or_get_state(const or_solver *s, double *U) { memcpy(U, s->U_spts, (size_t)s->nSpts * s->nEles * s->nFields * 8); }

int or_get_array(const or_solver *s, int which, double *out) {
  size_t nS = s->nSpts, nF = s->nFpts, D = s->nDims, nE = s->nEles, nK = s->nFields;
  const double *src = NULL;
  size_t n = 0;
  switch (which) {
  case 0: src = s->U_spts; n = nS * nE * nK; break;
  case 1: src = s->U_fpts; n = nF * nE * nK; break;
  case 2: src = s->F_spts; n = D * nS * nE * nK; break;
  case 3: src = s->Fn_fpts; n = nF * nE * nK; break;
  case 4: src = s->disFn_fpts; n = nF * nE * nK; break;
  case 5: src = s->dU_spts; n = D * nS * nE * nK; break;
  case 6: src = s->dU_fpts; n = D * nF * nE * nK; break;
  case 7: src = s->dUc_fpts; n = nF * nE * nK; break;
  case 8: src = s->U0; n = nS * nE * nK; break;
  case 9: src = s->waveSp; n = nF * nE; break;
  default:
    if (which >= 10 && which < 10 + s->nRK) { src = s->divF[which - 10]; n = nS * nE * nK; }
  }
  if (!src) return -1;
  memcpy(out, src, n * 8);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* pointwise physics, src/flux.cpp                                           */
/* ------------------------------------------------------------------------ */
/* inviscidFlux, flux.cpp:37-88 */
static void inviscid_flux(const or_params *p, const double *U, double F[3][5]) {
  if (p->equation == OR_ADVECTION_DIFFUSION) {
    F[0][0] = p->advectVx * U[0];
    F[1][0] = p->advectVy * U[0];
    if (p->nDims == 3) F[2][0] = p->advectVz * U[0];
  } else if (p->nDims == 2) {
    double rho = U[0], u = U[1] / rho, v = U[2] / rho, E = U[3];
    double pr = (p->gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    F[0][0] = rho * u;         F[1][0] = rho * v;
    F[0][1] = rho * u * u + pr; F[1][1] = rho * u * v;
    F[0][2] = rho * v * u;     F[1][2] = rho * v * v + pr;
    F[0][3] = (E + pr) * u;    F[1][3] = (E + pr) * v;
  } else {
    double rho = U[0], u = U[1] / rho, v = U[2] / rho, w = U[3] / rho, E = U[4];
    double pr = (p->gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v + w * w));
    F[0][0] = rho * u;         F[1][0] = rho * v;         F[2][0] = rho * w;
    F[0][1] = rho * u * u + pr; F[1][1] = rho * u * v;     F[2][1] = rho * u * w;
    F[0][2] = rho * v * u;     F[1][2] = rho * v * v + pr; F[2][2] = rho * v * w;
    F[0][3] = rho * w * u;     F[1][3] = rho * w * v;     F[2][3] = rho * w * w + pr;
    F[0][4] = (E + pr) * u;    F[1][4] = (E + pr) * v;    F[2][4] = (E + pr) * w;
  }
}

/* viscousFlux, flux.cpp:90-226.  viscFix=0 keeps the reference's index
 * defects (tauxz/tauyz at 197-198, Fvis[0][3] at 216); viscFix=1 uses the
 * physically correct terms. gradU[dim][field]. */
static void viscous_flux(const or_params *p, const double *U, const double gradU[3][5], double Fv[3][5]) {
  int D = p->nDims;
  double rho = U[0], u = U[1] / rho, v = U[2] / rho;
  double e = U[D + 1] / rho - 0.5 * (u * u + v * v);
  double w;
  if (D == 3) { w = U[3] / rho; e -= 0.5 * (w * w); } else w = 0.;
  double dRho_dx = gradU[0][0], dRhoU_dx = gradU[0][1], dRhoV_dx = gradU[0][2], dE_dx = gradU[0][D + 1];
  double dRho_dy = gradU[1][0], dRhoU_dy = gradU[1][1], dRhoV_dy = gradU[1][2], dE_dy = gradU[1][D + 1];
  double dRho_dz = 0, dRhoU_dz = 0, dRhoV_dz = 0, dRhoW_dx = 0, dRhoW_dy = 0, dRhoW_dz = 0, dE_dz = 0;
  if (D == 3) {
    dRho_dz = gradU[2][0]; dRhoU_dz = gradU[2][1]; dRhoV_dz = gradU[2][2];
    dRhoW_dx = gradU[0][3]; dRhoW_dy = gradU[1][3]; dRhoW_dz = gradU[2][3]; dE_dz = gradU[2][4];
  }
  double mu = p->mu_inf;
  if (!p->fixVis) {
    double rt_ratio = (p->gamma - 1.0) * e / (p->rt_inf);
    mu *= pow(rt_ratio, 1.5) * (1. + (p->c_sth)) / (rt_ratio + (p->c_sth));
  }
  double du_dx = (dRhoU_dx - dRho_dx * u) / rho, du_dy = (dRhoU_dy - dRho_dy * u) / rho;
  double dv_dx = (dRhoV_dx - dRho_dx * v) / rho, dv_dy = (dRhoV_dy - dRho_dy * v) / rho;
  double du_dz = 0, dv_dz = 0, dw_dx = 0, dw_dy = 0, dw_dz = 0;
  if (D == 3) {
    du_dz = (dRhoU_dz - dRho_dz * u) / rho;
    dv_dz = (dRhoV_dz - dRho_dz * v) / rho;
    dw_dx = (dRhoW_dx - dRho_dx * w) / rho;
    dw_dy = (dRhoW_dy - dRho_dy * w) / rho;
    dw_dz = (dRhoW_dz - dRho_dz * w) / rho;
  }
  double dK_dx, dK_dy, dK_dz;
  if (D == 2) {
    dK_dx = 0.5 * (u * u + v * v) * dRho_dx + rho * (u * du_dx + v * dv_dx);
    dK_dy = 0.5 * (u * u + v * v) * dRho_dy + rho * (u * du_dy + v * dv_dy);
    dK_dz = 0;
  } else {
    dK_dx = 0.5 * (u * u + v * v + w * w) * dRho_dx + rho * (u * du_dx + v * dv_dx + w * dw_dx);
    dK_dy = 0.5 * (u * u + v * v + w * w) * dRho_dy + rho * (u * du_dy + v * dv_dy + w * dw_dy);
    dK_dz = 0.5 * (u * u + v * v + w * w) * dRho_dz + rho * (u * du_dz + v * dv_dz + w * dw_dz);
  }
  double de_dx = (dE_dx - dK_dx - dRho_dx * e) / rho;
  double de_dy = (dE_dy - dK_dy - dRho_dy * e) / rho;
  double de_dz = 0;
  if (D == 3) de_dz = (dE_dz - dK_dz - dRho_dz * e) / rho;
  double diag = (du_dx + dv_dy + dw_dz) / 3.0;
  double tauxx = 2.0 * mu * (du_dx - diag), tauyy = 2.0 * mu * (dv_dy - diag);
  double tauxy = mu * (du_dy + dv_dx);
  double tauxz = 0, tauyz = 0, tauzz = 0;
  if (D == 3) {
    if (p->viscFix) {
      tauxz = mu * (du_dz + dw_dx);
      tauyz = mu * (dv_dz + dw_dy);
    } else {
      tauxz = mu * (du_dz + dv_dx);
      tauyz = mu * (du_dz + dv_dy);
    }
    tauzz = 2.0 * mu * (dw_dz - diag);
  }
  double kq = (mu / p->prandtl) * (p->gamma);
  Fv[0][0] = 0.0; Fv[0][1] = -tauxx; Fv[0][2] = -tauxy;
  Fv[0][D + 1] = -(u * tauxx + v * tauxy + w * tauxz + kq * de_dx);
  Fv[1][0] = 0.0; Fv[1][1] = -tauxy; Fv[1][2] = -tauyy;
  Fv[1][D + 1] = -(u * tauxy + v * tauyy + w * tauyz + kq * de_dy);
  if (D == 3) {
    Fv[0][3] = p->viscFix ? -tauxz : -tauzz;
    Fv[1][3] = -tauyz;
    Fv[2][0] = 0.0; Fv[2][1] = -tauxz; Fv[2][2] = -tauyz; Fv[2][3] = -tauzz;
    Fv[2][4] = -(u * tauxz + v * tauyz + w * tauzz + kq * de_dz);
  }
}

/* viscousFluxAD, flux.cpp:361-366 */
static void viscous_flux_ad(const or_params *p, const double gradU[3][5], double Fv[3][5]) {
  Fv[0][0] = -p->diffD * gradU[0][0];
  Fv[1][0] = -p->diffD * gradU[1][0];
  if (p->nDims == 3) Fv[2][0] = -p->diffD * gradU[2][0];
}

/* ------------------------------------------------------------------------ */
/* common fluxes, src/face.cpp                                               */
/* ------------------------------------------------------------------------ */
/* rusanovFlux, face.cpp:289-356 (static mesh) */
static void rusanov(const or_params *p, const double *UL, const double *UR, const double *n, double *Fn,
                    double *waveSp) {
  double FL[3][5], FR[3][5];
  int D = p->nDims, K = D + 2;
  inviscid_flux(p, UL, FL);
  inviscid_flux(p, UR, FR);
  double FnL[5] = {0, 0, 0, 0, 0}, FnR[5] = {0, 0, 0, 0, 0};
  double rhoL = UL[0], rhoR = UR[0];
  double uL = UL[1] / rhoL, uR = UR[1] / rhoR, vL = UL[2] / rhoL, vR = UR[2] / rhoR;
  double pL, pR, vnL = 0., vnR = 0.;
  if (D == 2) {
    pL = (p->gamma - 1.0) * (UL[3] - 0.5 * rhoL * (uL * uL + vL * vL));
    pR = (p->gamma - 1.0) * (UR[3] - 0.5 * rhoR * (uR * uR + vR * vR));
  } else {
    double wL = UL[3] / rhoL, wR = UR[3] / rhoR;
    pL = (p->gamma - 1.0) * (UL[4] - 0.5 * rhoL * (uL * uL + vL * vL + wL * wL));
    pR = (p->gamma - 1.0) * (UR[4] - 0.5 * rhoR * (uR * uR + vR * vR + wR * wR));
  }
  for (int d = 0; d < D; d++) {
    vnL += n[d] * UL[d + 1] / rhoL;
    vnR += n[d] * UR[d + 1] / rhoR;
    for (int i = 0; i < K; i++) {
      FnL[i] += n[d] * FL[d][i];
      FnR[i] += n[d] * FR[d][i];
    }
  }
  double csqL = fmax(p->gamma * pL / rhoL, 0.0), csqR = fmax(p->gamma * pR / rhoR, 0.0);
  double eigL = fabs(vnL) + sqrt(csqL), eigR = fabs(vnR) + sqrt(csqR);
  double eig = fmax(eigL, eigR);
  for (int i = 0; i < K; i++) Fn[i] = 0.5 * (FnL[i] + FnR[i] - eig * (UR[i] - UL[i]));
  /* wave speed for calculation of allowable dt (face.cpp:354) */
  if (waveSp) *waveSp = fmax(eigL, eigR);
}

/* roeFlux, face.cpp:358-484 (2D; entropy fix 436-446) */
static void roe(const or_params *p, const double *UL, const double *UR, const double *n, double *Fn) {
  double g = p->gamma;
  int D = p->nDims;
  double vL[3] = {0, 0, 0}, vR[3] = {0, 0, 0}, um[3] = {0, 0, 0}, du[5] = {0, 0, 0, 0, 0};
  for (int i = 0; i < D; i++) { vL[i] = UL[i + 1] / UL[0]; vR[i] = UR[i + 1] / UR[0]; }
  double pL, pR;
  if (D == 2) {
    pL = (g - 1.0) * (UL[3] - (0.5 * UL[0] * (vL[0] * vL[0] + vL[1] * vL[1])));
    pR = (g - 1.0) * (UR[3] - (0.5 * UR[0] * (vR[0] * vR[0] + vR[1] * vR[1])));
  } else {
    pL = (g - 1.0) * (UL[4] - (0.5 * UL[0] * (vL[0] * vL[0] + vL[1] * vL[1] + vL[2] * vL[2])));
    pR = (g - 1.0) * (UR[4] - (0.5 * UR[0] * (vR[0] * vR[0] + vR[1] * vR[1] + vR[2] * vR[2])));
  }
  double hL = (UL[D + 1] + pL) / UL[0], hR = (UR[D + 1] + pR) / UR[0];
  double sq_rho = sqrt(UR[0] / UL[0]), rrho = 1. / (sq_rho + 1.);
  double usq = 0., unm = 0., vgn = 0.;
  for (int i = 0; i < D; i++) {
    um[i] = rrho * (vL[i] + sq_rho * vR[i]);
    usq += 0.5 * um[i] * um[i];
    unm += um[i] * n[i];
  }
  double hm = rrho * (hL + sq_rho * hR);
  double am_sq = (g - 1.) * (hm - usq), am = sqrt(am_sq);
  double rhoUnL = 0., rhoUnR = 0.;
  for (int i = 0; i < D; i++) { rhoUnL += UL[i + 1] * n[i]; rhoUnR += UR[i + 1] * n[i]; }
  Fn[0] = rhoUnL + rhoUnR;
  Fn[1] = rhoUnL * vL[0] + rhoUnR * vR[0] + (pL + pR) * n[0];
  Fn[2] = rhoUnL * vL[1] + rhoUnR * vR[1] + (pL + pR) * n[1];
  Fn[3] = rhoUnL * hL + rhoUnR * hR;
  for (int i = 0; i < D + 2; i++) du[i] = UR[i] - UL[i];
  double lambda0 = fabs(unm - vgn), lambdaP = fabs(unm - vgn + am), lambdaM = fabs(unm - vgn - am);
  double eps = 0.5 * (fabs(rhoUnL / UL[0] - rhoUnR / UR[0]) + fabs(sqrt(g * pL / UL[0]) - sqrt(g * pR / UR[0])));
  if (lambda0 < 2. * eps) lambda0 = 0.25 * lambda0 * lambda0 / eps + eps;
  if (lambdaP < 2. * eps) lambdaP = 0.25 * lambdaP * lambdaP / eps + eps;
  if (lambdaM < 2. * eps) lambdaM = 0.25 * lambdaM * lambdaM / eps + eps;
  double a2 = 0.5 * (lambdaP + lambdaM) - lambda0;
  double a3 = 0.5 * (lambdaP - lambdaM) / am;
  double a1 = a2 * (g - 1.) / am_sq;
  double a4 = a3 * (g - 1.);
  double a5 = usq * du[0] - um[0] * du[1] - um[1] * du[2] + du[3];
  double a6 = unm * du[0] - n[0] * du[1] - n[1] * du[2];
  double aL1 = a1 * a5 - a3 * a6, bL1 = a4 * a5 - a2 * a6;
  Fn[0] -= lambda0 * du[0] + aL1;
  Fn[1] -= lambda0 * du[1] + aL1 * um[0] + bL1 * n[0];
  Fn[2] -= lambda0 * du[2] + aL1 * um[1] + bL1 * n[1];
  Fn[3] -= lambda0 * du[3] + aL1 * hm + bL1 * unm;
  for (int i = 0; i < D + 2; i++) Fn[i] *= 0.5;
}

/* laxFriedrichsFlux (advection), face.cpp:486-514 */
static void lax_friedrichs(const or_params *p, const double *UL, const double *UR, const double *n, double *Fn,
                           double *waveSp) {
  double uAvg = 0.5 * (UL[0] + UR[0]), uDiff = UL[0] - UR[0];
  double vNorm = p->advectVx * n[0] + p->advectVy * n[1];
  if (p->nDims == 3) vNorm += p->advectVz * n[2];
  Fn[0] = vNorm * uAvg + 0.5 * p->lambda * fabs(vNorm) * uDiff;
  if (waveSp) *waveSp = fabs(vNorm); /* face.cpp:504-508 (static mesh) */
}

/* boundFace::applyBCs, boundFace.cpp:33-277 (static mesh, vG = 0) */
static void apply_bc(const or_params *p, int bc, const double *UL, const double *n, double *UR) {
  int D = p->nDims, K = (p->equation == OR_ADVECTION_DIFFUSION) ? 1 : D + 2;
  if (p->equation == OR_ADVECTION_DIFFUSION) {
    /* the reference leaves UR at its zero initial value (face.cpp:41, boundFace.cpp:273-274) */
    for (int k = 0; k < K; k++) UR[k] = 0.0;
    return;
  }
  double rhoR = 0, pR, ER = 0, TR;
  double vL[3] = {0, 0, 0}, vR[3] = {0, 0, 0}, vG[3] = {0, 0, 0};
  double vBound[3] = {p->uBound, p->vBound, p->wBound};
  double gamma = p->gamma, gmo = gamma - 1.;
  double rhoL = UL[0], eL = UL[D + 1];
  for (int i = 0; i < D; i++) vL[i] = UL[i + 1] / UL[0];
  double vSq = 0;
  for (int i = 0; i < D; i++) vSq += (vL[i] * vL[i]);
  double pL = (gamma - 1.0) * (eL - 0.5 * rhoL * vSq);
  if (bc == OR_SUB_IN) {
    rhoR = p->rhoBound;
    for (int i = 0; i < D; i++) vR[i] = vBound[i];
    pR = pL;
    vSq = 0;
    for (int i = 0; i < D; i++) vSq += (vR[i] * vR[i]);
    ER = (pR / (gamma - 1.0)) + 0.5 * rhoR * vSq;
  } else if (bc == OR_SUB_OUT) {
    rhoR = rhoL;
    for (int i = 0; i < D; i++) vR[i] = vL[i];
    pR = p->pBound;
    vSq = 0.;
    for (int i = 0; i < D; i++) vSq += (vR[i] * vR[i]);
    ER = (pR / (gamma - 1.0)) + 0.5 * rhoR * vSq;
  } else if (bc == OR_SUP_IN) {
    rhoR = p->rhoBound;
    vR[0] = p->uBound; vR[1] = p->vBound; vR[2] = p->wBound;
    pR = p->pBound;
    vSq = 0.;
    for (int i = 0; i < D; i++) vSq += (vR[i] * vR[i]);
    ER = (pR / (gamma - 1.0)) + 0.5 * rhoR * vSq;
  } else if (bc == OR_SUP_OUT) {
    rhoR = rhoL;
    for (int i = 0; i < D; i++) vR[i] = vL[i];
    ER = eL;
  } else if (bc == OR_SLIP_WALL || bc == OR_SYMMETRY) {
    rhoR = rhoL;
    double vnL = 0.;
    for (int i = 0; i < D; i++) vnL += (vL[i] - vG[i]) * n[i];
    for (int i = 0; i < D; i++) vR[i] = vL[i] - (2.0) * vnL * n[i];
    ER = eL;
  } else if (bc == OR_ISOTHERMAL_NOSLIP_MOVING) {
    pR = pL;
    TR = p->TWall;
    rhoR = pR / (p->RGas * TR);
    vR[0] = vG[0] + p->uWall;
    vR[1] = vG[1] + p->vWall;
    if (D == 3) vR[2] = vG[2] + p->wWall;
    vSq = 0.;
    for (int i = 0; i < D; i++) vSq += (vR[i] * vR[i]);
    ER = (pR / (gamma - 1.0)) + 0.5 * rhoR * vSq;
  } else if (bc == OR_ISOTHERMAL_NOSLIP) {
    pR = pL;
    TR = p->TWall;
    rhoR = pR / (p->RGas * TR);
    for (int i = 0; i < D; i++) vR[i] = vG[i];
    vSq = 0.;
    for (int i = 0; i < D; i++) vSq += (vR[i] * vR[i]);
    ER = (pR / (gamma - 1.0)) + 0.5 * rhoR * vSq;
  } else if (bc == OR_ADIABATIC_NOSLIP) {
    rhoR = rhoL;
    pR = pL;
    for (int i = 0; i < D; i++) vR[i] = vG[i];
    vSq = 0.;
    for (int i = 0; i < D; i++) vSq += (vR[i] * vR[i]);
    ER = (pR / (gamma - 1.0)) + 0.5 * rhoR * vSq;
  } else if (bc == OR_CHAR_INOUT) {
    double vnL = 0., vnB = 0.;
    for (int i = 0; i < D; i++) vnL += vL[i] * n[i];
    for (int i = 0; i < D; i++) vnB += vBound[i] * n[i];
    double cL = sqrt(gamma * pL / rhoL), cB = sqrt(gamma * p->pBound / p->rhoBound);
    double R_L, R_B;
    if (fabs(vnB) >= cB && vnL >= 0) R_L = vnB + 2. / gmo * cB;
    else R_L = vnL + 2. / gmo * cL;
    if (fabs(vnB) >= cB && vnL < 0) R_B = vnL - 2. / gmo * cL;
    else R_B = vnB - 2. / gmo * cB;
    double rhoFac = .25 * (gamma - 1.) * (R_L - R_B);
    double rhoBG = rhoFac * rhoFac / gamma;
    if (vnL < 0) rhoBG *= pow(p->rhoBound, gamma) / p->pBound;
    else rhoBG *= pow(rhoL, gamma) / pL;
    rhoR = pow(rhoBG, 1. / gmo);
    pR = rhoFac * rhoFac * rhoR / gamma;
    double VB = 0.5 * (R_L + R_B);
    for (int i = 0; i < D; i++) {
      if (vnL < 0) vR[i] = vBound[i] + (VB - vnB) * n[i];
      else vR[i] = vL[i] + (VB - vnL) * n[i];
    }
    double vMag2 = 0;
    for (int i = 0; i < D; i++) vMag2 += vR[i] * vR[i];
    ER = pR / gmo + .5 * rhoR * vMag2;
  }
  UR[0] = rhoR;
  for (int i = 0; i < D; i++) UR[i + 1] = rhoR * vR[i];
  UR[D + 1] = ER;
}

/* boundFace::applyViscousBCs, boundFace.cpp:279-320 (gradients [dim][field]) */
static void apply_viscous_bc(const or_params *p, int bc, const double *UL, const double *UR, const double gL[3][5],
                             double gR[3][5]) {
  int D = p->nDims, K = D + 2;
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 5; k++) gR[d][k] = gL[d][k];
  if (bc != OR_ADIABATIC_NOSLIP) return;
  (void)K;
  double rhovSq = 0.;
  for (int d = 0; d < D; d++) rhovSq += UL[d + 1] * UL[d + 1];
  double pL = (p->gamma - 1.) * (UL[D + 1] - 0.5 * rhovSq / UL[0]);
  double pR = pL;
  double e = pR / ((p->gamma - 1.) * UR[0]);
  double gradVel[3][3];
  for (int d1 = 0; d1 < D; d1++)
    for (int d2 = 0; d2 < D; d2++)
      gradVel[d1][d2] = (gR[d1][d2 + 1] - gR[d1][0] * UR[d2 + 1] / UR[0]) / UR[0];
  double vSq = 0.;
  for (int d = 0; d < D; d++) vSq += UR[d + 1] * UR[d + 1] / (UR[0] * UR[0]);
  for (int d1 = 0; d1 < D; d1++) {
    gR[d1][D + 1] = (e + 0.5 * vSq) * gR[d1][0];
    for (int d2 = 0; d2 < D; d2++) gR[d1][D + 1] += UR[d2 + 1] * gradVel[d2][d1];
  }
}

/* ------------------------------------------------------------------------ */
/* residual, src/solver.cpp:252-291                                           */
/* ------------------------------------------------------------------------ */
#define UF(s_, fpt, e, k) (s_)->U_fpts[IDX3(fpt, e, k, (s_)->nEles, (s_)->nFields)]

/* solver::calcInviscidFlux_faces -> face::calcInviscidFlux (face.cpp:132-175),
 * intFace right state / scatter (intFace.cpp:85-126), LDG common solution
 * (face.cpp:582-607) */
static void inviscid_faces(or_solver *s) {
  const or_params *p = &s->p;
  int D = s->nDims, K = s->nFields, nfp = s->nFptsFace, nE = s->nEles;
#pragma omp parallel for schedule(static)
  for (int f = 0; f < s->nFaces; f++) {
    int eL = s->fL[f], eR = s->fR[f], bc = s->fbc[f], isBnd = (bc != OR_BC_NONE);
    for (int i = 0; i < nfp; i++) {
      int fptL = s->flocL[f] * nfp + i;
      int fptR = isBnd ? -1 : s->fptR[(size_t)f * nfp + i];
      double UL[5], UR[5], n[3] = {0, 0, 0}, Fn[5];
      for (int k = 0; k < K; k++) UL[k] = UF(s, fptL, eL, k);
      for (int d = 0; d < D; d++) n[d] = s->norm_fpts[((size_t)fptL * nE + eL) * D + d];
      double dAL = s->dA_fpts[(size_t)fptL * nE + eL];
      if (isBnd) apply_bc(p, bc, UL, n, UR);
      else for (int k = 0; k < K; k++) UR[k] = UF(s, fptR, eR, k);
      /* the left element's wave speed at this point (face::getPointers,
       * face.cpp:74-80); roeFlux stores none */
      double *ws = s->waveSp ? s->waveSp + (size_t)fptL * nE + eL : NULL;
      if (p->equation == OR_ADVECTION_DIFFUSION) lax_friedrichs(p, UL, UR, n, Fn, ws);
      else if (isBnd || p->riemannType == 0) rusanov(p, UL, UR, n, Fn, ws);
      else roe(p, UL, UR, n, Fn);
      double *keep = s->faceFn + ((size_t)f * nfp + i) * K;
      for (int k = 0; k < K; k++) {
        keep[k] = Fn[k];
        s->Fn_fpts[IDX3(fptL, eL, k, nE, K)] = Fn[k] * dAL;
      }
      if (p->viscous) {
        double UC[5];
        if (isBnd) {
          for (int k = 0; k < K; k++) UC[k] = 0.5 * (UL[k] + UR[k]);
        } else {
          double pen = p->penFact;
          if (D == 2) { if (n[0] + n[1] < 0) pen = -p->penFact; }
          else if (n[0] + n[1] + sqrt(2.) * n[2] < 0) pen = -p->penFact;
          for (int k = 0; k < K; k++) UC[k] = 0.5 * (UL[k] + UR[k]) - pen * (UL[k] - UR[k]);
        }
        double *uc = s->faceUC + ((size_t)f * nfp + i) * K;
        for (int k = 0; k < K; k++) {
          uc[k] = UC[k];
          s->dUc_fpts[IDX3(fptL, eL, k, nE, K)] = UC[k] - UL[k];
        }
        if (!isBnd)
          for (int k = 0; k < K; k++) s->dUc_fpts[IDX3(fptR, eR, k, nE, K)] = UC[k] - UR[k];
      }
      if (!isBnd) {
        double dAR = s->dA_fpts[(size_t)fptR * nE + eR];
        for (int k = 0; k < K; k++) s->Fn_fpts[IDX3(fptR, eR, k, nE, K)] = -Fn[k] * dAR;
      }
    }
  }
}

/* Common LDG viscous flux added to the inviscid normal flux Fn, in the left
 * element's frame (face::calcViscousFlux, face.cpp:177-287).  isBnd: 0
 * interior, 1 boundary, 2 adiabatic wall (flux from the ghost state). */
static void ldg_viscous_normal(const or_params *p, int isBnd, const double *UL, const double *UR,
                               const double gL[3][5], const double gR[3][5], const double *n, double *Fn) {
  int D = p->nDims, K = (p->equation == OR_NAVIER_STOKES) ? D + 2 : 1;
  double FL[3][5] = {{0}}, FR[3][5] = {{0}};
  if (p->equation == OR_NAVIER_STOKES) {
    double Fc[3][5];
    if (isBnd) {
      if (isBnd > 1) {
        viscous_flux(p, UR, gR, FR);
        for (int d = 0; d < D; d++)
          for (int k = 0; k < K; k++) Fc[d][k] = FR[d][k] + p->tau * n[d] * (UL[k] - UR[k]);
      } else {
        viscous_flux(p, UL, gL, FL);
        for (int d = 0; d < D; d++)
          for (int k = 0; k < K; k++) Fc[d][k] = FL[d][k] + p->tau * n[d] * (UL[k] - UR[k]);
      }
    } else {
      viscous_flux(p, UL, gL, FL);
      viscous_flux(p, UR, gR, FR);
      double pen = p->penFact;
      if (D == 2) { if (n[0] + n[1] < 0) pen = -p->penFact; }
      else if (n[0] + n[1] + sqrt(2.) * n[2] < 0) pen = -p->penFact;
      double nx = n[0], ny = n[1], nz = (D == 3) ? n[2] : 0;
      for (int k = 0; k < K; k++) {
        double dF0 = FL[0][k] - FR[0][k], dF1 = FL[1][k] - FR[1][k];
        double dU = UL[k] - UR[k];
        if (D == 2) {
          Fc[0][k] = 0.5 * (FL[0][k] + FR[0][k]) + pen * nx * (nx * (dF0) + ny * (dF1)) + p->tau * nx * (dU);
          Fc[1][k] = 0.5 * (FL[1][k] + FR[1][k]) + pen * ny * (nx * (dF0) + ny * (dF1)) + p->tau * ny * (dU);
        } else {
          /* face.cpp:245 uses tempFL[1] here; viscFix=1 uses tempFL[2] */
          double dF2 = (p->viscFix ? FL[2][k] : FL[1][k]) - FR[2][k];
          double nd = nx * (dF0) + ny * (dF1) + nz * (dF2);
          Fc[0][k] = 0.5 * (FL[0][k] + FR[0][k]) + pen * nx * nd + p->tau * nx * (dU);
          Fc[1][k] = 0.5 * (FL[1][k] + FR[1][k]) + pen * ny * nd + p->tau * ny * (dU);
          Fc[2][k] = 0.5 * (FL[2][k] + FR[2][k]) + pen * nz * nd + p->tau * nz * (dU);
        }
      }
    }
    for (int d = 0; d < D; d++)
      for (int k = 0; k < K; k++) Fn[k] += Fc[d][k] * n[d];
  } else {
    /* advection-diffusion: centralFlux of viscousFluxAD (face.cpp:270-279, flux.cpp:376-386) */
    viscous_flux_ad(p, gL, FL);
    viscous_flux_ad(p, gR, FR);
    double t = 0;
    for (int d = 0; d < D; d++) t += 0.5 * (FL[d][0] + FR[d][0]) * n[d];
    Fn[0] += t;
  }
}

/* face::calcViscousFlux, face.cpp:177-287 (+ intFace/boundFace gradient hooks) */
static void viscous_faces(or_solver *s) {
  const or_params *p = &s->p;
  int D = s->nDims, K = s->nFields, nfp = s->nFptsFace, nE = s->nEles, nF = s->nFpts;
#pragma omp parallel for schedule(static)
  for (int f = 0; f < s->nFaces; f++) {
    int eL = s->fL[f], eR = s->fR[f], bc = s->fbc[f];
    int isBnd = (bc != OR_BC_NONE) ? ((bc == OR_ADIABATIC_NOSLIP) ? 2 : 1) : 0;
    for (int i = 0; i < nfp; i++) {
      int fptL = s->flocL[f] * nfp + i;
      int fptR = isBnd ? -1 : s->fptR[(size_t)f * nfp + i];
      double UL[5], UR[5], n[3] = {0, 0, 0}, gL[3][5] = {{0}}, gR[3][5] = {{0}};
      for (int k = 0; k < K; k++) UL[k] = UF(s, fptL, eL, k);
      for (int d = 0; d < D; d++) n[d] = s->norm_fpts[((size_t)fptL * nE + eL) * D + d];
      double dAL = s->dA_fpts[(size_t)fptL * nE + eL];
      for (int d = 0; d < D; d++)
        for (int k = 0; k < K; k++) gL[d][k] = s->dU_fpts[(((size_t)d * nF + fptL) * nE + eL) * K + k];
      if (isBnd) {
        apply_bc(p, bc, UL, n, UR);
        if (p->equation == OR_NAVIER_STOKES) apply_viscous_bc(p, bc, UL, UR, gL, gR);
      } else {
        for (int k = 0; k < K; k++) UR[k] = UF(s, fptR, eR, k);
        for (int d = 0; d < D; d++)
          for (int k = 0; k < K; k++) gR[d][k] = s->dU_fpts[(((size_t)d * nF + fptR) * nE + eR) * K + k];
      }
      double *Fn = s->faceFn + ((size_t)f * nfp + i) * K;
      ldg_viscous_normal(p, isBnd, UL, UR, (const double(*)[5])gL, (const double(*)[5])gR, n, Fn);
      for (int k = 0; k < K; k++) s->Fn_fpts[IDX3(fptL, eL, k, nE, K)] = Fn[k] * dAL;
      if (!isBnd) {
        double dAR = s->dA_fpts[(size_t)fptR * nE + eR];
        for (int k = 0; k < K; k++) s->Fn_fpts[IDX3(fptR, eR, k, nE, K)] = -Fn[k] * dAR;
      }
    }
  }
}

void or_calc_residual(or_solver *s, int step) {
  const or_params *p = &s->p;
  int D = s->nDims, K = s->nFields, nS = s->nSpts, nF = s->nFpts, nE = s->nEles;
  int n = nE * K;
  size_t slab = (size_t)nS * nE * K;
  /* extrapolateU, solver.cpp:401-412 */
  oracle_dgemm(nF, n, nS, 1.0, s->opp_spts_to_fpts, s->U_spts, 0.0, s->U_fpts);
  /* calcGradU_spts, solver.cpp:808-821 */
  if (p->viscous)
    for (int d = 0; d < D; d++)
      oracle_dgemm(nS, n, nS, 1.0, s->opp_grad_spts + (size_t)d * nS * nS, s->U_spts, 0.0, s->dU_spts + (size_t)d * slab);
  /* calcInviscidFlux_spts, solver.cpp:514-539 */
#pragma omp parallel for schedule(static)
  for (int spt = 0; spt < nS; spt++)
    for (int e = 0; e < nE; e++) {
      double tF[3][5];
      inviscid_flux(p, s->U_spts + IDX3(spt, e, 0, nE, K), tF);
      for (int d1 = 0; d1 < D; d1++)
        for (int k = 0; k < K; k++) {
          double *Fp = s->F_spts + (size_t)d1 * slab + IDX3(spt, e, k, nE, K);
          if (p->viscous) { *Fp = tF[d1][k]; continue; }
          *Fp = 0.;
          for (int d2 = 0; d2 < D; d2++)
            *Fp += s->JGinv_spts[(((size_t)d1 * nS + spt) * nE + e) * D + d2] * tF[d2][k];
        }
    }
  inviscid_faces(s);
#define GI(a, b) s->JGinv_spts[(((size_t)(a) * nS + spt) * nE + e) * D + (b)]
  if (p->viscous) {
    /* correctGradU, solver.cpp:823-879 */
    for (int d = 0; d < D; d++)
      oracle_dgemm(nS, n, nF, 1.0, s->opp_correctU + (size_t)d * nS * nF, s->dUc_fpts, 1.0, s->dU_spts + (size_t)d * slab);
#pragma omp parallel for schedule(static)
    for (int spt = 0; spt < nS; spt++)
      for (int e = 0; e < nE; e++) {
        double invDet = 1. / s->detJac_spts[(size_t)spt * nE + e];
        for (int k = 0; k < K; k++) {
          size_t o = IDX3(spt, e, k, nE, K);
          if (D == 2) {
            double ur = s->dU_spts[o], us = s->dU_spts[slab + o];
            s->dU_spts[o] = invDet * (ur * GI(0, 0) + us * GI(1, 0));
            s->dU_spts[slab + o] = invDet * (ur * GI(0, 1) + us * GI(1, 1));
          } else {
            double ur = s->dU_spts[o], us = s->dU_spts[slab + o], ut = s->dU_spts[2 * slab + o];
            s->dU_spts[o] = invDet * (ur * GI(0, 0) + us * GI(1, 0) + ut * GI(2, 0));
            s->dU_spts[slab + o] = invDet * (ur * GI(0, 1) + us * GI(1, 1) + ut * GI(2, 1));
            s->dU_spts[2 * slab + o] = invDet * (ur * GI(0, 2) + us * GI(1, 2) + ut * GI(2, 2));
          }
        }
      }
    /* extrapolateGradU, solver.cpp:881-895 */
    size_t fslab = (size_t)nF * nE * K;
    for (int d = 0; d < D; d++)
      oracle_dgemm(nF, n, nS, 1.0, s->opp_spts_to_fpts, s->dU_spts + (size_t)d * slab, 0.0, s->dU_fpts + (size_t)d * fslab);
    /* calcViscousFlux_spts, solver.cpp:562-596 */
#pragma omp parallel for schedule(static)
    for (int spt = 0; spt < nS; spt++)
      for (int e = 0; e < nE; e++) {
        double g[3][5] = {{0}}, tF[3][5] = {{0}};
        size_t o0 = IDX3(spt, e, 0, nE, K);
        for (int d = 0; d < D; d++)
          for (int k = 0; k < K; k++) g[d][k] = s->dU_spts[(size_t)d * slab + o0 + k];
        if (p->equation == OR_NAVIER_STOKES) viscous_flux(p, s->U_spts + o0, (const double(*)[5])g, tF);
        else viscous_flux_ad(p, (const double(*)[5])g, tF);
        for (int d = 0; d < D; d++)
          for (int k = 0; k < K; k++) tF[d][k] += s->F_spts[(size_t)d * slab + o0 + k];
        for (int d1 = 0; d1 < D; d1++)
          for (int k = 0; k < K; k++) {
            double acc = 0.;
            for (int d2 = 0; d2 < D; d2++) acc += GI(d1, d2) * tF[d2][k];
            s->F_spts[(size_t)d1 * slab + o0 + k] = acc;
          }
      }
#undef GI
    viscous_faces(s);
  }
  /* extrapolateNormalFlux (static), solver.cpp:764-786 */
  oracle_dgemm(nF, n, nS, 1.0, s->opp_extrapolateFn, s->F_spts, 0.0, s->disFn_fpts);
  for (int d = 1; d < D; d++)
    oracle_dgemm(nF, n, nS, 1.0, s->opp_extrapolateFn + (size_t)d * nF * nS, s->F_spts + (size_t)d * slab, 1.0, s->disFn_fpts);
  /* calcDivF_spts, solver.cpp:707-727 */
  double *C = s->divF[step];
  oracle_dgemm(nS, n, nS, 1.0, s->opp_grad_spts, s->F_spts, 0.0, C);
  for (int d = 1; d < D; d++)
    oracle_dgemm(nS, n, nS, 1.0, s->opp_grad_spts + (size_t)d * nS * nS, s->F_spts + (size_t)d * slab, 1.0, C);
  /* correctDivFlux, solver.cpp:789-806 */
  size_t nfv = (size_t)nF * nE * K;
  for (size_t i = 0; i < nfv; i++) s->disFn_fpts[i] = s->Fn_fpts[i] - s->disFn_fpts[i];
  oracle_dgemm(nS, n, nF, 1.0, s->opp_correction, s->disFn_fpts, 1.0, C);
}

/* ------------------------------------------------------------------------ */
/* time integration, src/solver.cpp:212-250, 307-399 (dtType 0, no PMG)       */
/* ------------------------------------------------------------------------ */
/* timeStepA / timeStepB: params->dt, or eles[e]->dt for dtType 2
 * (solver.cpp:307-385) */
void or_time_step_a(or_solver *s, int step, double RKval) {
  int nS = s->nSpts, nE = s->nEles, K = s->nFields;
#pragma omp parallel for schedule(static)
  for (int spt = 0; spt < nS; spt++)
    for (int e = 0; e < nE; e++) {
      double dt = (s->dtType == 2) ? s->dtEle[e] : s->p.dt;
      for (int k = 0; k < K; k++) {
        size_t o = IDX3(spt, e, k, nE, K);
        s->U_spts[o] = s->U0[o] - RKval * s->divF[step][o] / s->detJac_spts[(size_t)spt * nE + e] * dt;
      }
    }
}
void or_time_step_b(or_solver *s, int step) {
  int nS = s->nSpts, nE = s->nEles, K = s->nFields;
  double b = s->RKb[step];
#pragma omp parallel for schedule(static)
  for (int spt = 0; spt < nS; spt++)
    for (int e = 0; e < nE; e++) {
      double dt = (s->dtType == 2) ? s->dtEle[e] : s->p.dt;
      for (int k = 0; k < K; k++) {
        size_t o = IDX3(spt, e, k, nE, K);
        s->U_spts[o] -= b * s->divF[step][o] / s->detJac_spts[(size_t)spt * nE + e] * dt;
      }
    }
}

double or_get_dt(const or_solver *s, int e) {
  return (e >= 0 && s->dtEle && s->dtType == 2) ? s->dtEle[e] : s->p.dt;
}

/* getCFLLimit, global.cpp:45-68 */
static double cfl_limit(int order) {
  static const double lim[6] = {1.393, 0.464, 0.235, 0.1453, 0.100, 0.0736};
  return (order >= 0 && order <= 5) ? lim[order] : -1.0;
}

/* ele::calcWaveSpFpts (ele.cpp:1031-1084, static grids): the wave speed at
 * every flux point from the traces U_fpts, |v.n| + c (Euler) or |a.n|
 * (advection) with the element's own normal. */
static void calc_wave_sp_fpts(or_solver *s) {
  const or_params *p = &s->p;
  int D = s->nDims, K = s->nFields, nF = s->nFpts, nE = s->nEles;
  for (int e = 0; e < nE; e++)
    for (int fpt = 0; fpt < nF; fpt++) {
      const double *n = s->norm_fpts + ((size_t)fpt * nE + e) * D;
      double ws;
      if (p->equation == OR_ADVECTION_DIFFUSION) {
        double vN = p->advectVx * n[0] + p->advectVy * n[1];
        if (D == 3) vN += p->advectVz * n[2];
        ws = fabs(vN);
      } else {
        const double *U = s->U_fpts + IDX3(fpt, e, 0, nE, K);
        double rho = U[0], u = U[1] / rho, v = U[2] / rho, w = (D == 3) ? U[3] / rho : 0.;
        double rhoVSq = rho * (u * u + v * v + w * w);
        double pr = (p->gamma - 1) * (U[D + 1] - 0.5 * rhoVSq);
        double vN = u * n[0] + v * n[1];
        if (D == 3) vN += w * n[2];
        double csq = fmax(p->gamma * pr / rho, 0.0);
        ws = fabs(vN) + sqrt(csq);
      }
      s->waveSp[(size_t)fpt * nE + e] = ws;
    }
}

/* ele::calcDt (ele.cpp:1086-1105) from waveSp, then solver::calcDt's minimum
 * (solver.cpp:298-305): dtType 1 sets params->dt to it, dtType 2 also keeps
 * each element's dt for the stage updates.  Returns the minimum. */
static double calc_dt_ws(or_solver *s) {
  int D = s->nDims, nS = s->nSpts, nF = s->nFpts, nE = s->nEles, nfp = s->nFptsFace, P1 = s->P1;
  double wS[64];
  for (int m = 0; m < nfp; m++) {
    const double *w1 = GL_WTS[s->order];
    wS[m] = (D == 3) ? w1[m / P1] * w1[m % P1] : w1[m];
  }
  if (!s->dtEle) s->dtEle = xcalloc(nE, 8);
  double dtMin = INFINITY;
  for (int e = 0; e < nE; e++) {
    double vol = 0;
    for (int spt = 0; spt < nS; spt++) vol += s->detJac_spts[(size_t)spt * nE + e] * s->wts[spt];
    double intWave = 0;
    for (int fpt = 0; fpt < nF; fpt++)
      intWave += s->waveSp[(size_t)fpt * nE + e] * wS[fpt % nfp] * s->dA_fpts[(size_t)fpt * nE + e];
    s->dtEle[e] = s->CFL * cfl_limit(s->order) * 2 * vol / intWave;
    dtMin = fmin(dtMin, s->dtEle[e]);
  }
  if (s->dtType != 0) s->p.dt = dtMin;
  return dtMin;
}

/* solver::calcDt (solver.cpp:293-305): on the first iteration the wave
 * speeds come from the traces of the current state (calcWaveSpFpts; there
 * U_fpts still holds initializeSolution's extrapolation of the initial state,
 * solver.cpp:1497-1503), afterwards from what the face fluxes of the last
 * residual stored. */
static double calc_dt(or_solver *s) {
  if (s->dtFirst) {
    oracle_dgemm(s->nFpts, s->nEles * s->nFields, s->nSpts, 1.0, s->opp_spts_to_fpts, s->U_spts, 0.0, s->U_fpts);
    calc_wave_sp_fpts(s);
    s->dtFirst = 0;
  }
  return calc_dt_ws(s);
}

/* CFL time stepping from the current state (solver::initializeSolution's
 * extrapolateU + calcWaveSpFpts, solver.cpp:1497-1503, and the dt the first
 * update computes): dtType 1 global, 2 local; returns the minimum.  The next
 * update is the first: it forms the wave speeds from the state again. */
double or_calc_dt(or_solver *s, int dtType, double CFL) {
  int K = s->nFields, nS = s->nSpts, nF = s->nFpts, nE = s->nEles;
  oracle_dgemm(nF, nE * K, nS, 1.0, s->opp_spts_to_fpts, s->U_spts, 0.0, s->U_fpts);
  if (!s->waveSp) s->waveSp = xcalloc((size_t)nF * nE, 8);
  s->dtType = dtType;
  s->CFL = CFL;
  calc_wave_sp_fpts(s);
  double m = calc_dt_ws(s);
  s->dtFirst = 1;
  return m;
}
void or_copy_u_u0(or_solver *s) { memcpy(s->U0, s->U_spts, (size_t)s->nSpts * s->nEles * s->nFields * 8); }
void or_copy_u0_u(or_solver *s) { memcpy(s->U_spts, s->U0, (size_t)s->nSpts * s->nEles * s->nFields * 8); }

void or_update(or_solver *s) {
  int nRK = s->nRK;
  for (int step = 0; step < nRK - 1; step++) {
    if (step == 0 && s->dtType != 0) calc_dt(s);
    if (step == 0) or_copy_u_u0(s);
    or_calc_residual(s, step);
    or_time_step_a(s, step, s->RKa[step + 1]);
  }
  or_calc_residual(s, nRK - 1);
  if (s->p.timeType < 5) {
    if (nRK > 1) or_copy_u0_u(s);
    else if (s->dtType != 0) calc_dt(s);
    for (int step = 0; step < nRK; step++) or_time_step_b(s, step);
  } else {
    or_time_step_a(s, nRK - 1, s->RKa[nRK]);
  }
}

/* ele::getNormResidual (ele.cpp:1882-1903) summed over elements and finished
 * as writeResidual does (output.cpp:846-892) */
void or_norm_residual(const or_solver *s, int normType, double *out) {
  int nS = s->nSpts, nE = s->nEles, K = s->nFields;
  for (int k = 0; k < K; k++) out[k] = 0;
  const double *dF = s->divF[0];
  for (int e = 0; e < nE; e++) {
    double res[5] = {0, 0, 0, 0, 0};
    for (int spt = 0; spt < nS; spt++)
      for (int k = 0; k < K; k++) {
        double v = dF[IDX3(spt, e, k, nE, K)], J = s->detJac_spts[(size_t)spt * nE + e];
        if (normType == 1) res[k] += fabs(v) * s->wts[spt];
        else if (normType == 2) res[k] += v * v / J * s->wts[spt];
        else res[k] = fmax(fabs(v) / J, res[k]);
      }
    for (int k = 0; k < K; k++) out[k] = (normType == 3) ? fmax(out[k], res[k]) : out[k] + res[k];
  }
  if (normType == 2)
    for (int k = 0; k < K; k++) out[k] = sqrt(out[k]);
}

/* ------------------------------------------------------------------------ */
/* error norms: solver::integrateError (solver.cpp:1507-1574) with           */
/* ele::calcEleError (ele.cpp:792-810), calcError (funcs.cpp:1875-2056) and  */
/* ele::calcPos (ele.cpp:389-404)                                            */
/* ------------------------------------------------------------------------ */
static const double OR_PI = 3.141592653589793; /* 4 atan(1), global.hpp:37 */

/* calcError (funcs.cpp:1875-2056): the error of state U at physical point pos
 * against the analytic solution of the test case (the solution itself when
 * testCase == 0), then |.| (errorNorm 1) or squared (errorNorm 2) */
static void calc_error(const or_solver *s, const or_error_params *ep, const double *U, const double *pos,
                       double *err) {
  const int D = s->nDims, K = s->nFields;
  if (ep->testCase == 0) {
    for (int k = 0; k < K; k++) err[k] = U[k];
    return;
  }
  for (int k = 0; k < K; k++) err[k] = 0.0;
  const double px = pos[0], py = pos[1], pz = D == 3 ? pos[2] : 0.0;
  if (s->p.equation == OR_NAVIER_STOKES) {
    const double gamma = s->p.gamma;
    if (ep->icType == 0) {
      /* uniform free stream: entropy error (funcs.cpp:1891-1903) */
      double pfs = ep->pIC, rhofsg = pow(ep->rhoIC, gamma);
      double rho = U[0], u = U[1] / rho, v = U[2] / rho, w = (D == 2) ? 0 : U[3] / rho;
      double p = (gamma - 1.) * (U[D + 1] - .5 * rho * (u * u + v * v + w * w));
      err[0] = p * rhofsg / (pow(rho, gamma) * pfs) - 1.;
    } else if (ep->icType == 1) {
      /* isentropic vortex of strength 5 convected along (1, 1) (1904-1946) */
      double eps = 5.0;
      double xmin = ep->xmin, xmax = ep->xmax, ymin = ep->ymin, ymax = ep->ymax;
      double x = fmod((px - ep->time), (xmax - xmin));
      double y = fmod((py - ep->time), (ymax - ymin));
      if (x > xmax) x -= (xmax - xmin);
      if (y > ymax) y -= (ymax - ymin);
      if (x < xmin) x += (xmax - xmin);
      if (y < ymin) y += (ymax - ymin);
      double f = 1.0 - (x * x + y * y);
      double rho = fmax(pow(1. - eps * eps * (gamma - 1.) / (8. * gamma * OR_PI * OR_PI) * exp(f), 1.0 / (gamma - 1.0)),
                        1e-10);
      double vx = 1. - eps * y / (2. * OR_PI) * exp(f / 2.);
      double vy = 1. + eps * x / (2. * OR_PI) * exp(f / 2.);
      double p = pow(rho, gamma);
      err[0] = rho;
      err[1] = rho * vx;
      err[2] = rho * vy;
      if (D == 3) err[3] = 0.;
      err[D + 1] = p / (gamma - 1) + (0.5 * rho * (vx * vx + vy * vy));
      for (int i = 0; i < K; i++) err[i] = U[i] - err[i];
    } else if (ep->icType == 2) {
      /* Liang & Miyaji's vortex (1947-2000) */
      double eps = 1.0, rc = 1.0, Minf = .3, Uinf = 1, rhoInf = 1, theta = ep->vortexAngle;
      double Pinf = pow(Minf, -2) / gamma;
      double eM = (eps * Minf) * (eps * Minf);
      double xmin = ep->xmin, xmax = ep->xmax, ymin = ep->ymin, ymax = ep->ymax;
      double advX = Uinf * cos(theta) * ep->time, advY = Uinf * sin(theta) * ep->time;
      double x = (fmod((px - advX), (xmax - xmin)));
      double y = (fmod((py - advY), (ymax - ymin)));
      if (x > xmax) x -= (xmax - xmin);
      if (x < xmin) x += (xmax - xmin);
      if (y > ymax) y -= (ymax - ymin);
      if (y < ymin) y += (ymax - ymin);
      double f = -(x * x + y * y) / (rc * rc);
      double vx = Uinf * (cos(theta) - y * eps / rc * exp(f / 2.));
      double vy = Uinf * (sin(theta) + x * eps / rc * exp(f / 2.));
      double rho = rhoInf * pow(1. - (gamma - 1.) / 2. * eM * exp(f), gamma / (gamma - 1.0));
      double p = Pinf * pow(1. - (gamma - 1.) / 2. * eM * exp(f), gamma / (gamma - 1.0));
      err[0] = rho;
      err[1] = rho * vx;
      err[2] = rho * vy;
      if (D == 3) err[3] = 0.;
      err[D + 1] = p / (gamma - 1) + (0.5 * rho * (vx * vx + vy * vy));
      for (int i = 0; i < K; i++) err[i] = U[i] - err[i];
    }
  } else {
    /* advection-diffusion (2001-2044) */
    double xmin = ep->xmin, xmax = ep->xmax, ymin = ep->ymin, ymax = ep->ymax;
    double x = fabs(fmod((px - ep->time), (xmax - xmin)));
    double y = fabs(fmod((py - ep->time), (ymax - ymin)));
    if (x > (xmax - xmin) / 2) x -= (xmax - xmin);
    if (y > (ymax - ymin) / 2) y -= (ymax - ymin);
    if (ep->icType == 0) {
      double r2 = x * x + y * y;
      err[0] = exp(-r2);
    } else if (ep->icType == 1) {
      err[0] = 1 + sin(2. * OR_PI * (px + 5. - ep->time) / 10.);
    } else if (ep->icType == 2) {
      err[0] = cos(2 * OR_PI * px / 6.) * cos(2 * OR_PI * py / 6.) * cos(2 * OR_PI * pz / 6.);
    }
    for (int i = 0; i < K; i++) err[i] = U[i] - err[i];
  }
  if (ep->errorNorm == 1)
    for (int k = 0; k < K; k++) err[k] = fabs(err[k]);
  else if (ep->errorNorm == 2)
    for (int k = 0; k < K; k++) err[k] *= err[k];
}

int or_integrate_error(const or_solver *s, const or_error_params *ep, double *out) {
  const int D = s->nDims, nS = s->nSpts, nE = s->nEles, K = s->nFields, nN = s->nNodes, P1 = s->P1;
  double LpErr[5] = {0, 0, 0, 0, 0}, err[5], pos[3];
  if (ep->errorNorm == 0) {
    /* the error at the solution points against the integral of the solution
     * (solver.cpp:1510-1526; quadrature weights of the solution order) */
    double intU[5] = {0, 0, 0, 0, 0};
    for (int ic = 0; ic < nE; ic++)
      for (int spt = 0; spt < nS; spt++) {
        const double *U = s->U_spts + (size_t)K * ((size_t)spt * nE + ic);
        for (int d = 0; d < D; d++) pos[d] = s->pos_spts[IDX3(spt, ic, d, nE, D)];
        if (ep->testCase) calc_error(s, ep, U, pos, err);
        else for (int k = 0; k < K; k++) err[k] = U[k];
        const double wJ = s->detJac_spts[(size_t)spt * nE + ic];
        for (int k = 0; k < K; k++) {
          LpErr[k] += err[k] * s->wts[spt] * wJ;
          intU[k] += U[k] * s->wts[spt] * wJ;
        }
      }
    for (int k = 0; k < K; k++) LpErr[k] = fabs(LpErr[k] - intU[k]);
  } else {
    /* quadrature points of order quadOrder (setupInterpolateSptsQpts,
     * operators.cpp:305-316; getQptWeights, points.cpp:347-374) */
    const int qo = ep->quadOrder, nq1 = qo + 1;
    if (qo < 0 || qo > 10) { set_err("quadrature order %d not tabulated", qo); return -1; }
    const int nQ = D == 3 ? nq1 * nq1 * nq1 : nq1 * nq1;
    const double *q1 = GL_PTS[qo], *w1 = GL_WTS[qo];
    double *loc = xcalloc((size_t)nQ * 3, 8), *wq = xcalloc(nQ, 8);
    for (int q = 0; q < nQ; q++) {
      int i = q % nq1, j = (q / nq1) % nq1, k = q / (nq1 * nq1);
      loc[3 * q] = q1[i];
      loc[3 * q + 1] = q1[j];
      loc[3 * q + 2] = D == 3 ? q1[k] : 0.0;
      wq[q] = D == 3 ? w1[k] * w1[j] * w1[i] : w1[j] * w1[i];
    }
    /* opp_spts_to_qpts (setupInterpolateSptsIpts, operators.cpp:259-303) */
    double *A = xcalloc((size_t)nQ * nS, 8);
    for (int q = 0; q < nQ; q++)
      for (int spt = 0; spt < nS; spt++) {
        int ks = spt / (P1 * P1), js = (spt - P1 * P1 * ks) / P1, is = spt - P1 * js - P1 * P1 * ks;
        if (D == 2) {
          is = spt % P1;
          js = spt / P1;
        }
        double v = lagrange(s->spts1D, P1, loc[3 * q], is) * lagrange(s->spts1D, P1, loc[3 * q + 1], js);
        if (D == 3) v = v * lagrange(s->spts1D, P1, loc[3 * q + 2], ks);
        A[(size_t)q * nS + spt] = v;
      }
    double *Uq = xcalloc((size_t)nQ * nE * K, 8), *Jq = xcalloc((size_t)nQ * nE, 8);
    oracle_dgemm(nQ, nE * K, nS, 1.0, A, s->U_spts, 0.0, Uq);
    oracle_dgemm(nQ, nE, nS, 1.0, A, s->detJac_spts, 0.0, Jq);
    double sh[216];
    for (int ic = 0; ic < nE; ic++)
      for (int q = 0; q < nQ; q++) {
        /* ele::calcPos: the shape functions at the point times the nodes */
        shape_dshape_n(D, nN, loc + 3 * q, sh, NULL);
        pos[0] = pos[1] = pos[2] = 0;
        for (int n = 0; n < nN; n++)
          for (int d = 0; d < D; d++) pos[d] += sh[n] * s->nodes[IDX3(n, ic, d, nE, D)];
        calc_error(s, ep, Uq + (size_t)K * ((size_t)q * nE + ic), pos, err);
        for (int k = 0; k < K; k++) LpErr[k] += err[k] * wq[q] * Jq[(size_t)q * nE + ic];
      }
    free(loc); free(wq); free(A); free(Uq); free(Jq);
  }
  if (ep->errorNorm == 2)
    for (int k = 0; k < K; k++) LpErr[k] = sqrt(fabs(LpErr[k]));
  for (int k = 0; k < K; k++) out[k] = LpErr[k];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* test hooks: pointwise routines exposed so that tests can check them        */
/* against independent textbook implementations                              */
/* ------------------------------------------------------------------------ */
void or_point_inviscid_flux(const or_params *p, const double *U, double *F /* [3][5] */) {
  double t[3][5] = {{0}};
  inviscid_flux(p, U, t);
  memcpy(F, t, sizeof t);
}
/* kind: 0 Rusanov, 1 Roe, 2 Lax-Friedrichs (advection) */
void or_point_common_flux(const or_params *p, int kind, const double *UL, const double *UR, const double *n,
                          double *Fn) {
  if (kind == 0) rusanov(p, UL, UR, n, Fn, NULL);
  else if (kind == 1) roe(p, UL, UR, n, Fn);
  else lax_friedrichs(p, UL, UR, n, Fn, NULL);
}
void or_point_boundary_state(const or_params *p, int bc, const double *UL, const double *n, double *UR) {
  apply_bc(p, bc, UL, n, UR);
}
void or_point_viscous_flux(const or_params *p, const double *U, const double *grad /* [3][5] */,
                           double *F /* [3][5] */) {
  double g[3][5], t[3][5] = {{0}};
  memcpy(g, grad, sizeof g);
  if (p->equation == OR_NAVIER_STOKES) viscous_flux(p, U, (const double(*)[5])g, t);
  else viscous_flux_ad(p, (const double(*)[5])g, t);
  memcpy(F, t, sizeof t);
}
void or_point_viscous_common(const or_params *p, int isBnd, const double *UL, const double *UR,
                             const double *gL /* [3][5] */, const double *gR /* [3][5] */, const double *n,
                             double *Fn) {
  double a[3][5], b[3][5];
  memcpy(a, gL, sizeof a);
  memcpy(b, gR, sizeof b);
  ldg_viscous_normal(p, isBnd, UL, UR, (const double(*)[5])a, (const double(*)[5])b, n, Fn);
}
void or_point_viscous_bc_gradient(const or_params *p, int bc, const double *UL, const double *UR,
                                  const double *gL /* [3][5] */, double *gR /* [3][5] */) {
  double a[3][5], b[3][5];
  memcpy(a, gL, sizeof a);
  apply_viscous_bc(p, bc, UL, UR, (const double(*)[5])a, b);
  memcpy(gR, b, sizeof b);
}
double or_point_dvcjh(double r, int mode, int order, double eta) { return dvcjh_1d(r, mode, order, eta); }
double or_point_eta(int scheme, int order) { return compute_eta(scheme, order); }
