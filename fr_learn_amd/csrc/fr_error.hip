// fr_error.hip -- error norms of the current state on the device
// (solver::integrateError, solver.cpp:1507-1574; ele::calcEleError,
// ele.cpp:792-810; calcError, funcs.cpp:1875-2056; ele::calcPos,
// ele.cpp:389-404).
//
// One thread per (element, point): errorNorm 1 / 2 at the (quadOrder+1)^D
// Gauss points, the state and detJ interpolated from the solution points by
// the tensor Lagrange products of opp_spts_to_qpts summed in solution-point
// order (the dense product the reference's dgemm forms); errorNorm 0 at the
// solution points, with the integral of the state beside it.  Each block sums
// its threads' terms in a tree; the host adds the block sums in block order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fr_ctx.hpp"

namespace frdev {

struct ErrArgs {
  const double *U;     // [ele][field][spt]
  const double *detJ;  // [ele][spt]
  const double *X;     // element nodes [ele][node][3]
  const double *Lq;    // errorNorm 1, 2: Lagrange basis of the solution points at the 1D Gauss points [nq1][P1]
  const double *sh;    // shape functions at the points [pt][nN]
  const double *w;     // quadrature weight of each point [pt]
  double *part;        // block sums [block][2 * NF]
  int NF, NS, nN, nE, nPts, nq1, P1;
  int testCase, errorNorm, icType, equation;
  double time, rhoIC, pIC, vortexAngle, xmin, xmax, ymin, ymax, gamma;
};

// calcError (funcs.cpp:1875-2056): the error of state U at physical point
// pos against the test case's analytic solution (U itself when testCase is
// 0), then |.| (errorNorm 1) or squared (errorNorm 2).
template <int D>
__device__ void calcError(const ErrArgs &a, const double *U, const double *pos, double *err) {
  const int K = a.NF;
  constexpr double pi = 3.141592653589793;  // 4 atan(1), global.hpp:37
  if (a.testCase == 0) {
    for (int k = 0; k < K; k++) err[k] = U[k];
    return;
  }
  for (int k = 0; k < K; k++) err[k] = 0.0;
  const double px = pos[0], py = pos[1], pz = D == 3 ? pos[2] : 0.0;
  if (a.equation == 1) {
    const double gamma = a.gamma;
    if (a.icType == 0) {
      // uniform free stream: entropy error
      const double pfs = a.pIC, rhofsg = pow(a.rhoIC, gamma);
      const double rho = U[0], u = U[1] / rho, v = U[2] / rho, w = (D == 2) ? 0 : U[3] / rho;
      const double p = (gamma - 1.) * (U[D + 1] - .5 * rho * (u * u + v * v + w * w));
      err[0] = p * rhofsg / (pow(rho, gamma) * pfs) - 1.;
    } else if (a.icType == 1 || a.icType == 2) {
      double rho, vx, vy, p;
      if (a.icType == 1) {
        // isentropic vortex of strength 5 convected along (1, 1)
        const double eps = 5.0;
        double x = fmod((px - a.time), (a.xmax - a.xmin));
        double y = fmod((py - a.time), (a.ymax - a.ymin));
        if (x > a.xmax) x -= (a.xmax - a.xmin);
        if (y > a.ymax) y -= (a.ymax - a.ymin);
        if (x < a.xmin) x += (a.xmax - a.xmin);
        if (y < a.ymin) y += (a.ymax - a.ymin);
        const double f = 1.0 - (x * x + y * y);
        rho = fmax(pow(1. - eps * eps * (gamma - 1.) / (8. * gamma * pi * pi) * exp(f), 1.0 / (gamma - 1.0)), 1e-10);
        vx = 1. - eps * y / (2. * pi) * exp(f / 2.);
        vy = 1. + eps * x / (2. * pi) * exp(f / 2.);
        p = pow(rho, gamma);
      } else {
        // the vortex of Liang & Miyaji
        const double eps = 1.0, rc = 1.0, Minf = .3, Uinf = 1, rhoInf = 1, theta = a.vortexAngle;
        const double Pinf = pow(Minf, -2) / gamma;
        const double eM = (eps * Minf) * (eps * Minf);
        const double advX = Uinf * cos(theta) * a.time, advY = Uinf * sin(theta) * a.time;
        double x = (fmod((px - advX), (a.xmax - a.xmin)));
        double y = (fmod((py - advY), (a.ymax - a.ymin)));
        if (x > a.xmax) x -= (a.xmax - a.xmin);
        if (x < a.xmin) x += (a.xmax - a.xmin);
        if (y > a.ymax) y -= (a.ymax - a.ymin);
        if (y < a.ymin) y += (a.ymax - a.ymin);
        const double f = -(x * x + y * y) / (rc * rc);
        vx = Uinf * (cos(theta) - y * eps / rc * exp(f / 2.));
        vy = Uinf * (sin(theta) + x * eps / rc * exp(f / 2.));
        rho = rhoInf * pow(1. - (gamma - 1.) / 2. * eM * exp(f), gamma / (gamma - 1.0));
        p = Pinf * pow(1. - (gamma - 1.) / 2. * eM * exp(f), gamma / (gamma - 1.0));
      }
      err[0] = rho;
      err[1] = rho * vx;
      err[2] = rho * vy;
      if (D == 3) err[3] = 0.;
      err[D + 1] = p / (gamma - 1) + (0.5 * rho * (vx * vx + vy * vy));
      for (int i = 0; i < K; i++) err[i] = U[i] - err[i];
    }
  } else {
    double x = fabs(fmod((px - a.time), (a.xmax - a.xmin)));
    double y = fabs(fmod((py - a.time), (a.ymax - a.ymin)));
    if (x > (a.xmax - a.xmin) / 2) x -= (a.xmax - a.xmin);
    if (y > (a.ymax - a.ymin) / 2) y -= (a.ymax - a.ymin);
    if (a.icType == 0) err[0] = exp(-(x * x + y * y));
    else if (a.icType == 1) err[0] = 1 + sin(2. * pi * (px + 5. - a.time) / 10.);
    else if (a.icType == 2) err[0] = cos(2 * pi * px / 6.) * cos(2 * pi * py / 6.) * cos(2 * pi * pz / 6.);
    for (int i = 0; i < K; i++) err[i] = U[i] - err[i];
  }
  if (a.errorNorm == 1)
    for (int k = 0; k < K; k++) err[k] = fabs(err[k]);
  else if (a.errorNorm == 2)
    for (int k = 0; k < K; k++) err[k] *= err[k];
}

template <int D>
__global__ __launch_bounds__(256) void k_error(ErrArgs a) {
  __shared__ double red[256][10];
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int NF = a.NF;
  double acc[10];
  for (int k = 0; k < 10; k++) acc[k] = 0.0;
  if (idx < (long)a.nE * a.nPts) {
    const int e = (int)(idx / a.nPts), q = (int)(idx - (long)e * a.nPts);
    const double *Ue = a.U + (size_t)e * NF * a.NS;
    double U[5], J, pos[3] = {0, 0, 0}, err[5];
    if (a.errorNorm == 0) {
      for (int k = 0; k < NF; k++) U[k] = Ue[k * a.NS + q];
      J = a.detJ[(size_t)e * a.NS + q];
    } else {
      // U and detJ at Gauss point q: opp_spts_to_qpts row q times the
      // solution-point values, in solution-point order
      const int n1 = a.nq1, P1 = a.P1;
      const int qi = q % n1, qj = (q / n1) % n1, qk = q / (n1 * n1);
      for (int k = 0; k < NF; k++) U[k] = 0.0;
      J = 0.0;
      for (int s = 0; s < a.NS; s++) {
        const int si = s % P1, sj = (s / P1) % P1, sk = s / (P1 * P1);
        double A = a.Lq[qi * P1 + si] * a.Lq[qj * P1 + sj];
        if (D == 3) A = A * a.Lq[qk * P1 + sk];
        for (int k = 0; k < NF; k++) U[k] += A * Ue[k * a.NS + s];
        J += A * a.detJ[(size_t)e * a.NS + s];
      }
    }
    // ele::calcPos: the shape functions at the point times the nodes
    const double *sh = a.sh + (size_t)q * a.nN;
    const double *Xe = a.X + (size_t)e * a.nN * 3;
    for (int n = 0; n < a.nN; n++)
      for (int d = 0; d < D; d++) pos[d] += sh[n] * Xe[n * 3 + d];
    if (a.errorNorm == 0 && a.testCase == 0) {
      for (int k = 0; k < NF; k++) err[k] = U[k];
    } else {
      calcError<D>(a, U, pos, err);
    }
    const double w = a.w[q];
    for (int k = 0; k < NF; k++) acc[k] = err[k] * w * J;
    if (a.errorNorm == 0)
      for (int k = 0; k < NF; k++) acc[NF + k] = U[k] * w * J;
  }
  for (int k = 0; k < 10; k++) red[threadIdx.x][k] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 10; k++) red[threadIdx.x][k] += red[threadIdx.x + s][k];
    __syncthreads();
  }
  if (threadIdx.x < 2 * NF) a.part[(size_t)blockIdx.x * 2 * NF + threadIdx.x] = red[0][threadIdx.x];
}

}  // namespace frdev

using namespace frimpl;

extern "C" int fr_integrate_error(fr_ctx *c, const fr_error_params *p, double *out) {
  if (!c || !p || !out) return fail(FR_EINVAL, "null argument");
  if (c->tri) return fail(FR_EINVAL, "error norms are available for quads and hexes");
  if (p->errorNorm < 0 || p->errorNorm > 2) return fail(FR_EINVAL, "errorNorm must be 0, 1 or 2");
  if (p->testCase != 0 && p->testCase != 1)
    return fail(FR_EINVAL, "testCase must be 0 (the solution) or 1 (the analytic error)");
  const int D = c->D, NF = c->NF, NS = c->NS, nE = c->nE, P1 = c->P + 1;
  const int nN = (int)(c->X.size() / ((size_t)nE * 3));
  // the points: the solution points (errorNorm 0) or the Gauss points of
  // quadOrder with their tensor weights (getQptWeights, points.cpp:347-374)
  std::vector<double> loc, w, Lq;
  int nq1 = 0;
  if (p->errorNorm == 0) {
    loc = c->ops.locSpts;
    w = c->ops.wts;
  } else {
    const std::vector<double> q1 = frhost::pts1D(0, p->quadOrder), w1 = frhost::gaussWeights1D(p->quadOrder);
    if (q1.empty()) return fail(FR_EINVAL, "Gauss-Ledgendre point locations for this quadOrder not implemented.");
    nq1 = p->quadOrder + 1;
    const int nQ = D == 3 ? nq1 * nq1 * nq1 : nq1 * nq1;
    loc.assign((size_t)nQ * 3, 0.0);
    w.resize(nQ);
    for (int q = 0; q < nQ; q++) {
      const int i = q % nq1, j = (q / nq1) % nq1, k = q / (nq1 * nq1);
      loc[3 * q] = q1[i];
      loc[3 * q + 1] = q1[j];
      loc[3 * q + 2] = D == 3 ? q1[k] : 0.0;
      w[q] = D == 3 ? w1[k] * w1[j] * w1[i] : w1[j] * w1[i];
    }
    Lq.resize((size_t)nq1 * P1);
    for (int i = 0; i < nq1; i++)
      for (int m = 0; m < P1; m++) Lq[(size_t)i * P1 + m] = frhost::lagrange(c->ops.spts1D, q1[i], m);
  }
  const int nPts = (int)w.size();
  std::vector<double> sh((size_t)nPts * nN);
  for (int q = 0; q < nPts; q++) frhost::shapeFunctions(D, &loc[3 * q], &sh[(size_t)q * nN], nN);
  // tables and the element nodes to the device (kept between calls)
  const size_t nTab = Lq.size() + sh.size() + w.size();
  int rc;
  if (c->errTabCap < nTab) {
    if (c->errTab) (void)hipFree(c->errTab);
    c->errTab = nullptr;
    c->errTabCap = 0;
    if ((rc = dmalloc(&c->errTab, nTab))) return rc;
    c->errTabCap = nTab;
  }
  if (!c->Xd) {
    if ((rc = dmalloc(&c->Xd, c->X.size()))) return rc;
    HIP_TRY(hipMemcpyAsync(c->Xd, c->X.data(), c->X.size() * 8, hipMemcpyHostToDevice, c->stream));
  }
  std::vector<double> tab;
  tab.reserve(nTab);
  tab.insert(tab.end(), Lq.begin(), Lq.end());
  tab.insert(tab.end(), sh.begin(), sh.end());
  tab.insert(tab.end(), w.begin(), w.end());
  HIP_TRY(hipMemcpyAsync(c->errTab, tab.data(), nTab * 8, hipMemcpyHostToDevice, c->stream));
  const long n = (long)nE * nPts;
  const int nb = (int)((n + 255) / 256);
  if ((rc = needPart(c, (size_t)nb * 2 * NF))) return rc;
  frdev::ErrArgs a{};
  a.U = c->U;
  a.detJ = c->detJ;
  a.X = c->Xd;
  a.Lq = c->errTab;
  a.sh = c->errTab + Lq.size();
  a.w = a.sh + sh.size();
  a.part = c->part;
  a.NF = NF; a.NS = NS; a.nN = nN; a.nE = nE; a.nPts = nPts; a.nq1 = nq1; a.P1 = P1;
  a.testCase = p->testCase; a.errorNorm = p->errorNorm; a.icType = p->icType; a.equation = c->cfg.equation;
  a.time = p->time; a.rhoIC = p->rhoIC; a.pIC = p->pIC; a.vortexAngle = p->vortexAngle;
  a.xmin = p->xmin; a.xmax = p->xmax; a.ymin = p->ymin; a.ymax = p->ymax; a.gamma = c->cfg.gamma;
  if (D == 3) hipLaunchKernelGGL(frdev::k_error<3>, dim3(nb), dim3(256), 0, c->stream, a);
  else hipLaunchKernelGGL(frdev::k_error<2>, dim3(nb), dim3(256), 0, c->stream, a);
  HIP_TRY(hipGetLastError());
  std::vector<double> h((size_t)nb * 2 * NF);
  HIP_TRY(hipMemcpyAsync(h.data(), c->part, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  double Lp[5] = {0, 0, 0, 0, 0}, intU[5] = {0, 0, 0, 0, 0};
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < NF; k++) {
      Lp[k] += h[(size_t)b * 2 * NF + k];
      intU[k] += h[(size_t)b * 2 * NF + NF + k];
    }
  for (int k = 0; k < NF; k++) {
    double v = (p->errorNorm == 0) ? std::fabs(Lp[k] - intU[k]) : Lp[k];
    if (p->errorNorm == 2) v = std::sqrt(std::fabs(v));
    out[k] = v;
  }
  return 0;
}
