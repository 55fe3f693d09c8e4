// fr_api.hip -- implementation of the C ABI in include/fr.h: context,
// device buffers, kernel dispatch and the solver verbs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "fr_adjoint.hpp"
#include "fr_ctx.hpp"

using namespace frimpl;

#ifndef FR_STAGE_REVERSE
#define FR_STAGE_REVERSE 1  // stage launches alternate their element-block order
#endif


namespace {

void freeAll(fr_ctx *c) {
  void *bufs[] = {c->U,       c->U0,      c->detJ,    c->JG64,    c->fgeo64, c->ownN,  c->wts,
                  c->oned64,  c->part,    c->U32,     c->JG32,    c->fgeo32, c->oned32, c->Us,
                  c->Uf[0],   c->Uf[1],   c->divF[0], c->divF[1], c->divF[2], c->divF[3], c->fnbr,
                  c->finfo,   c->dUf,     c->dtG,     c->dtE,     c->adjA,   c->adjTr, c->adjTr2,
                  c->adjLam,  c->adjOut,  c->adjNu,   c->adjW[0], c->adjW[1], c->adjS[0], c->adjS[1],
                  c->adjS[2], c->adjS[3], c->adjQf,   c->adjB,    c->JGc64,  c->detJc, c->fgeoc64,
                  c->JGc32,   c->fgeoc32, c->adjUo,   c->adjQo,   c->adjCj,   c->Fv,   c->waveSp,
                  c->dtNum,   c->wsurf,   c->dtPart,  c->geoN64,  c->geoN32,  c->errTab, c->Xd};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &v : c->ev)
    for (auto &pr : v) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  if (c->own) (void)hipStreamDestroy(c->own);
}

template <typename T>
const T *stateIn(fr_ctx *c, int which) {
  if (which == 1) return (const T *)c->Us;
  return std::is_same<T, float>::value ? (const T *)c->U32 : (const T *)c->U;
}

// Make Uf[cur] hold the traces of state `which` (0 = U_spts, 1 = Us).
template <typename T>
int ensureTraces(fr_ctx *c, int which) {
  // traces written by launches of another element blocking lack that
  // blocking's in-block faces
  if (c->traceOf == which && (c->traceEpb == 0 || c->traceEpb == c->epb)) return 0;
  const T *U = stateIn<T>(c, which);
  const T *o1 = std::is_same<T, float>::value ? (const T *)c->oned32 : (const T *)c->oned64;
  if (c->tri) {
    int rc = launchTriTrace<T>(c, U, (T *)c->Uf[c->cur], o1);
    if (rc) return rc;
    c->traceOf = which;
    c->traceEpb = 0;
    return 0;
  }
  hipError_t err = hipSuccess;
  bool ok = dispatch(c->D, c->P, c->NF, [&](auto Dc, auto Pc, auto NFc) {
    constexpr int D = decltype(Dc)::value, P = decltype(Pc)::value, NF = decltype(NFc)::value;
    using G = frdev::ElemGeo<D, P>;
    TimeScope ts(c, 1);
    hipLaunchKernelGGL((frdev::k_trace<D, P, NF, T>), dim3(nblk((long)c->nE * G::NFP)), dim3(256), 0, c->stream, U,
                       (T *)c->Uf[c->cur], o1, c->nE);
    err = hipGetLastError();
  });
  if (!ok) return fail(FR_EINVAL, "no kernel instance for this (nDims, order, nFields)");
  if (err != hipSuccess) return fail(FR_EHIP, std::string("trace launch: ") + hipGetErrorString(err));
  c->traceOf = which;
  c->traceEpb = 0;
  return 0;
}

// Stage-kernel arguments: input state `in` (0 = U_spts, 1 = Us), traces in
// Uf[cur], mode / stage / RK coefficient as in k_stage.
template <typename T>
frdev::StageArgs<T> makeStageArgs(fr_ctx *c, int in, int cur, int mode, int step, double rk) {
  const bool f32 = std::is_same<T, float>::value;
  frdev::StageArgs<T> a{};
  a.Uin = stateIn<T>(c, in);
  a.UfIn = (const T *)c->Uf[cur];
  a.UfOut = (T *)c->Uf[cur ^ 1];
  const int gm = c->geoActive();
  const bool aff = gm == 1;
  a.affine = aff ? 1 : 0;
  a.nodal = gm == 2 ? 1 : gm == 3 ? 2 : 0;
  a.geoN = f32 ? (const T *)c->geoN32 : (const T *)c->geoN64;
  if (aff) {
    a.JG = f32 ? (const T *)c->JGc32 : (const T *)c->JGc64;
    a.fgeo = f32 ? (const T *)c->fgeoc32 : (const T *)c->fgeoc64;
  } else {
    a.JG = f32 ? (const T *)c->JG32 : (const T *)c->JG64;
    a.fgeo = f32 ? (const T *)c->fgeo32 : (const T *)c->fgeo64;
  }
  a.fnbr = c->fnbr;
  a.finfo = c->finfo;
  a.oned = f32 ? (const T *)c->oned32 : (const T *)c->oned64;
  a.divF = (T *)c->divF[step];
  a.Ubase = c->U;
  a.UbaseT = f32 ? (const T *)c->U32 : (const T *)c->U;
  a.detJ = aff ? c->detJc : c->detJ;
  a.Unext = (T *)c->Us;
  a.Uout = c->U;
  a.Uout32 = f32 ? c->U32 : nullptr;
  a.dprev0 = (const T *)c->divF[0];
  a.dprev1 = (const T *)c->divF[c->nRK > 1 ? 1 : 0];
  a.dprev2 = (const T *)c->divF[c->nRK > 2 ? 2 : 0];
  a.c0 = c->RKb[0];
  a.c1 = c->RKb[1];
  a.c2 = c->RKb[2];
  a.c3 = c->RKb[3];
  a.rk = rk;
  a.dtE = dtSource(c).first;
  a.dtStride = dtSource(c).second;
  // CFL time stepping: the face fluxes store their wave speeds for the next
  // update's time step (not while the adjoint replays given time steps)
  a.waveSp = (c->dtType != 0 && !c->dtFrozen && !c->tri) ? c->waveSp : nullptr;
  a.dUf = (const T *)c->dUf;
  a.Fv = (const T *)c->Fv;
  // Inviscid quads and hexes store no intermediate stage state: a stage that
  // would read Us forms it from the base and the previous stage's residual
  // (the viscous gradient kernel and the triangle kernels read Us itself).
  if (!c->cfg.viscous && !c->tri) {
    if (in == 1 && step > 0) {
      a.rprev = (const T *)c->divF[step - 1];
      a.rkPrev = c->RKa[step];
    }
    if (mode == 1) a.Unext = nullptr;
  }
  // the final stage of a multi-stage update leaves divF_spts[nRK-1] out
  // unless asked to keep it (stage 1's residual, divF_spts[0], is the one the
  // residual norms read)
  if (mode >= 2 && step > 0 && !c->keepFinal) a.divF = nullptr;
  a.mode = mode;
  a.nprev = step;
  a.nEles = c->nE;
  a.ph = c->ph;
  return a;
}

// One launch of k_stage.  in: 0 = U_spts, 1 = Us.  mode as in k_stage.
template <typename T>
int launchStage(fr_ctx *c, int in, int mode, int step, double rk) {
  int rc = ensureTraces<T>(c, in);
  if (rc) return rc;
  frdev::StageArgs<T> a = makeStageArgs<T>(c, in, c->cur, mode, step, rk);
  if (FR_STAGE_REVERSE) {  // alternate the block order from launch to launch
    a.rev = c->revNext;
    c->revNext ^= 1;
  }
  int rc2 = launchStageArgs<T>(c, a, mode);
  if (rc2) return rc2;
  if (mode != 0) {
    c->cur ^= 1;
    c->traceOf = (mode == 1) ? 1 : 0;
    c->traceEpb = c->tri ? 0 : c->epb;  // the triangle kernels write every face
  }
  if (step == c->nRK - 1) c->finalDropped = a.divF == nullptr;
  return 0;
}

int residual(fr_ctx *c, int step) {
  if (step < 0 || step >= c->nRK) return fail(FR_EINVAL, "step out of range");
  // the residual's face fluxes store new wave speeds (face.cpp:354), which
  // the next update's time step then uses
  if (c->dtType != 0 && !c->dtFirst && !c->dtFrozen && !c->tri) c->dtReady = false;
  return c->prec == FR_PREC_FP32 ? launchStage<float>(c, 0, 0, step, 0.0) : launchStage<double>(c, 0, 0, step, 0.0);
}

// solver::update (solver.cpp:212-250): every stage is one fused launch that
// evaluates the residual of the stage state, applies the stage update and
// extrapolates the new state to the flux points.  U_spts holds the step's
// starting state (the reference's U0) until the final stage overwrites it.
template <typename T>
int updateT(fr_ctx *c) {
  int rc;
  for (int step = 0; step < c->nRK - 1; step++)
    if ((rc = launchStage<T>(c, step == 0 ? 0 : 1, 1, step, c->RKa[step + 1]))) return rc;
  const int last = c->nRK - 1;
  const int in = (c->nRK == 1) ? 0 : 1;
  if (c->cfg.timeType < 5) return launchStage<T>(c, in, 2, last, 0.0);
  return launchStage<T>(c, in, 3, last, c->RKa[c->nRK]);
}

// timeStepA as a separate verb (the fused update does not use it)
template <typename TD>
int rkA(fr_ctx *c, int step, double rk) {
  long n = (long)c->nState();
  TimeScope ts(c, 2);
  hipLaunchKernelGGL((frdev::k_rk_a<TD>), dim3(nblk(n)), dim3(256), 0, c->stream, c->U, c->U0,
                     (const TD *)c->divF[step], c->detJ, c->U32, rk, dtSource(c).first, dtSource(c).second, 0,
                     c->NF, c->NS, n, c->ph.eOff);
  HIP_TRY(hipGetLastError());
  c->traceOf = -1;
  return 0;
}

// Forward Euler with a CFL time step after the first update: the reference
// computes dt after the step's residual, from the wave speeds that residual
// stored (solver.cpp:236-240), so the residual, the time step and the update
// (timeStepB) are three launches here instead of one fused stage.
int updateEulerCfl(fr_ctx *c) {
  int rc = residual(c, 0);
  if (rc) return rc;
  if ((rc = dtPrepare(c, c->stream))) return rc;
  long n = (long)c->nState();
  if (c->prec == FR_PREC_FP32)
    hipLaunchKernelGGL(frdev::k_rk_b1<float>, dim3(nblk(n)), dim3(256), 0, c->stream, c->U,
                       (const float *)c->divF[0], c->detJ, c->U32, c->RKb[0], dtSource(c).first,
                       dtSource(c).second, c->NF, c->NS, n, c->ph.eOff);
  else
    hipLaunchKernelGGL(frdev::k_rk_b1<double>, dim3(nblk(n)), dim3(256), 0, c->stream, c->U,
                       (const double *)c->divF[0], c->detJ, c->U32, c->RKb[0], dtSource(c).first,
                       dtSource(c).second, c->NF, c->NS, n, c->ph.eOff);
  HIP_TRY(hipGetLastError());
  c->traceOf = -1;
  c->dtReady = false;
  return 0;
}

// solver::update: the time step first (solver.cpp:218-219), then the stages.
int updateOnce(fr_ctx *c) {
  int rc;
  if (c->dtType != 0 && !c->dtFrozen) {
    if (c->nRK == 1 && !c->dtFirst && !c->tri) return updateEulerCfl(c);
    if ((rc = dtPrepare(c, c->stream))) return rc;
  }
  rc = c->prec == FR_PREC_FP32 ? updateT<float>(c) : updateT<double>(c);
  if (c->dtType != 0 && !c->dtFrozen) {
    c->dtFirst = false;
    c->dtReady = false;
  }
  return rc;
}

// reference layout [pt][ele][field] <-> device [device ele][field][pt]
void refToDev(const fr_ctx *c, const double *src, double *dst, int NF, int NP) {
  const int nE = c->nE;
  for (int p = 0; p < NP; p++)
    for (int e = 0; e < nE; e++) {
      const size_t de = (size_t)c->iperm[e];
      for (int k = 0; k < NF; k++) dst[(de * NF + k) * NP + p] = src[((size_t)p * nE + e) * NF + k];
    }
}
void devToRef(const fr_ctx *c, const double *src, double *dst, int NF, int NP) {
  const int nE = c->nE;
  for (int p = 0; p < NP; p++)
    for (int e = 0; e < nE; e++) {
      const size_t de = (size_t)c->iperm[e];
      for (int k = 0; k < NF; k++) dst[((size_t)p * nE + e) * NF + k] = src[(de * NF + k) * NP + p];
    }
}

int syncU32(fr_ctx *c) {
  if (c->prec != FR_PREC_FP32) return 0;
  long n = (long)c->nState();
  hipLaunchKernelGGL(frdev::k_state_to_float, dim3(nblk(n)), dim3(256), 0, c->stream, c->U, c->U32, c->NF,
                     c->NS, c->ph.eOff, n);
  HIP_TRY(hipGetLastError());
  return 0;
}

// The energy offset of the FP32 states (Phys::eOff) for state U (device
// layout [ele][field][spt] or the reference layout): the mean pressure over the
// solution points divided by gamma - 1.  Inviscid Euler quads and hexes in
// FP32 only; zero otherwise.  A change bumps the settings generation.
void setEnergyOffset(fr_ctx *c, const double *U, int layout) {
  double eOff = 0.0;
  if (c->prec == FR_PREC_FP32 && c->cfg.equation == 1 && !c->cfg.viscous && !c->tri) {
    const int NF = c->NF, NS = c->NS, D = c->D;
    const size_t nPts = (size_t)c->nE * NS;
    double sum = 0.0;
    for (size_t pt = 0; pt < nPts; pt++) {
      double u[5];
      if (layout == FR_LAYOUT_REF) {
        for (int k = 0; k < NF; k++) u[k] = U[pt * NF + k];
      } else {
        const size_t e = pt / NS, s = pt % NS;
        for (int k = 0; k < NF; k++) u[k] = U[(e * NF + k) * NS + s];
      }
      double m2 = 0.0;
      for (int d = 0; d < D; d++) m2 += u[d + 1] * u[d + 1];
      sum += (c->cfg.gamma - 1.0) * (u[NF - 1] - 0.5 * m2 / u[0]);
    }
    const double pMean = sum / (double)nPts;
    if (std::isfinite(pMean)) eOff = pMean / (c->cfg.gamma - 1.0);
  }
  if (eOff != c->ph.eOff) {
    c->ph.eOff = eOff;
    c->ph.pOff = eOff * (c->cfg.gamma - 1.0);
    c->ph.hOff = eOff * c->cfg.gamma;
    c->gen++;
  }
}

// A verb changed the state: a first update after fr_calc_dt takes its wave
// speeds from the state, so its time step must be formed again.
void stateChanged(fr_ctx *c) {
  if (c->dtFirst) c->dtReady = false;
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
// The members' stage arguments in device memory, which the captured graphs
// read: built at fr_batch_create, and rebuilt in place (same buffers, so the
// graphs stay valid) when a member's settings changed since.
template <typename T>
int batchArgs(fr_batch *b) {
  const size_t n = b->ctx.size();
  for (int p = 0; p < 2; p++) {
    std::vector<frdev::StageArgs<T>> h((size_t)b->nRK * n);
    for (int s = 0; s < b->nRK; s++)
      for (size_t i = 0; i < n; i++) {
        int mode;
        double rk;
        stageMode(b->ctx[i], s, mode, rk);
        h[(size_t)s * n + i] = makeStageArgs<T>(b->ctx[i], s == 0 ? 0 : 1, p ^ (s & 1), mode, s, rk);
      }
    int rc;
    if (!b->args[p] && (rc = dmalloc((frdev::StageArgs<T> **)&b->args[p], h.size()))) return rc;
    HIP_TRY(hipMemcpy(b->args[p], h.data(), h.size() * sizeof(frdev::StageArgs<T>), hipMemcpyHostToDevice));
  }
  std::vector<frdev::StageArgs<T>> h(n);
  for (size_t i = 0; i < n; i++) {
    h[i] = makeStageArgs<T>(b->ctx[i], 0, 0, 0, 0, 0.0);
    h[i].waveSp = nullptr;  // the tuning launches leave the CFL wave speeds alone
  }
  int rc;
  if (!b->argsTune && (rc = dmalloc((frdev::StageArgs<T> **)&b->argsTune, n))) return rc;
  HIP_TRY(hipMemcpy(b->argsTune, h.data(), n * sizeof(frdev::StageArgs<T>), hipMemcpyHostToDevice));
  b->gen.resize(n);
  for (size_t i = 0; i < n; i++) b->gen[i] = b->ctx[i]->gen;
  return 0;
}

// A member's settings must suit the batched step: a fixed or CFL time step
// (forward Euler with a CFL time step takes its dt after the step's residual,
// which the fused batched stage cannot).
int batchMemberCheck(const fr_ctx *c) {
  if (c->dtType != 0 && c->nRK == 1)
    return fail(FR_EINVAL, "batched forward Euler needs a fixed time step (fr_set_dt)");
  return 0;
}

// Before a batched step: every instance holds the traces of U_spts in
// Uf[parity] (the stage kernels leave only the cross-block traces in HBM,
// which is what the next step reads: complete, or written by launches of the
// batch's own element blocking).
int batchPrepare(fr_batch *b, int parity) {
  for (fr_ctx *c : b->ctx) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int epb = batchEpb(b, c, b->lay[batchGroup(c)]);
    if (c->traceOf != 0 || c->cur != parity || (c->traceEpb != 0 && c->traceEpb != epb)) {
      c->cur = parity;
      c->traceOf = -1;
      int rc = (c->prec == FR_PREC_FP32) ? ensureTraces<float>(c, 0) : ensureTraces<double>(c, 0);
      if (rc) return rc;
      HIP_TRY(hipStreamSynchronize(c->stream));
    }
  }
  return 0;
}

void batchFree(fr_batch *b) {
  // the live members forget this batch
  for (size_t k = 0; k < b->ctx.size(); k++)
    if (k < b->alive.size() && b->alive[k]) {
      auto &v = b->ctx[k]->batches;
      for (size_t i = 0; i < v.size(); i++)
        if (v[i] == b) {
          v.erase(v.begin() + i);
          break;
        }
    }
  for (int p = 0; p < 2; p++) {
    if (b->exec[p]) (void)hipGraphExecDestroy(b->exec[p]);
    if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
    if (b->args[p]) (void)hipFree(b->args[p]);
  }
  if (b->argsTune) (void)hipFree(b->argsTune);
  for (int grp = 0; grp < 16; grp++) {
    if (b->ps[grp]) (void)hipStreamDestroy(b->ps[grp]);
    if (b->join[grp]) (void)hipEventDestroy(b->join[grp]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (int l = 0; l < 2; l++)
    if (b->items[l]) (void)hipFree(b->items[l]);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

// ---------------------------------------------------------------------------
// adjoint (FP64 contexts, inviscid and advection)
// ---------------------------------------------------------------------------
int adjCheck(fr_ctx *c) {
  if (c->prec != FR_PREC_FP64) return fail(FR_EINVAL, "the adjoint needs an FP64 context");
  if (c->tri) return fail(FR_EINVAL, "the adjoint is not available for triangles");
  const size_t nSt = c->nState(), nTr = (size_t)c->nE * c->NF * c->NFP;
  int rc;
  if (!c->adjA) {
    if ((rc = dmalloc(&c->adjA, nTr)) || (rc = dmalloc(&c->adjTr, nTr)) || (rc = dmalloc(&c->adjTr2, nTr)) ||
        (rc = dmalloc(&c->adjLam, nSt)) || (rc = dmalloc(&c->adjOut, nSt)) || (rc = dmalloc(&c->adjNu, nSt)) ||
        (rc = dmalloc(&c->adjW[0], nSt)) || (rc = dmalloc(&c->adjW[1], nSt)))
      return rc;
    for (int s = 1; s < c->nRK; s++)
      if ((rc = dmalloc(&c->adjS[s], nSt))) return rc;
    if (c->cfg.viscous && ((rc = dmalloc(&c->adjQf, nTr * c->D)) || (rc = dmalloc(&c->adjB, nTr)) ||
                           (rc = dmalloc(&c->adjUo, nTr)) || (rc = dmalloc(&c->adjQo, nTr * c->D)) ||
                           (rc = dmalloc(&c->adjCj, nTr))))
      return rc;
  }
  return 0;
}

// The stage states of one RK step from U (timeStepA, solver.cpp:307-345):
// S[0] = U, S[s+1] = U - RKa[s+1] R(S[s]) dt / detJ, with the stage kernel.
int stageStates(fr_ctx *c, const double **S) {
  S[0] = c->U;
  for (int s = 0; s + 1 < c->nRK; s++) {
    dispatch(c->D, c->P, c->NF, [&](auto Dc, auto Pc, auto NFc) {
      constexpr int D = decltype(Dc)::value, P = decltype(Pc)::value, NF = decltype(NFc)::value;
      const long nq = (long)c->nE * frdev::ElemGeo<D, P>::NFP;
      hipLaunchKernelGGL((frdev::k_trace<D, P, NF, double>), dim3(nblk(nq)), dim3(256), 0, c->stream, S[s],
                         c->adjTr, c->oned64, c->nE);
    });
    HIP_TRY(hipGetLastError());
    // the update's own arguments, redirected to the adjoint's buffers
    frdev::StageArgs<double> a = makeStageArgs<double>(c, 0, 0, 1, s, c->RKa[s + 1]);
    a.Uin = S[s];
    a.UfIn = c->adjTr;
    a.UfOut = c->adjTr2;
    a.divF = nullptr;
    a.Unext = c->adjS[s + 1];
    a.Uout = nullptr;
    a.Uout32 = nullptr;
    a.waveSp = nullptr;  // the recomputation leaves the CFL wave speeds alone
    int rc = launchStageArgs<double>(c, a, 1);
    if (rc) return rc;
    S[s + 1] = c->adjS[s + 1];
  }
  return 0;
}

// lam0 = (d U_{n+1} / d U_n)^T lam1 for one update at the current U:
// nu_s = -c_s (dt/detJ) lam1 - RKa[s+1] (dt/detJ) w_{s+1},
// w_s = (dR/dU at S[s])^T nu_s, lam0 = lam1 + sum_s w_s, where c_s is the
// weight of R_s in the final update (RKb, or RKa[nRK] on the last stage of
// the Jameson schemes).
int vjpStepDev(fr_ctx *c, const double *lam1, double *lam0) {
  const double *S[4];
  int rc = stageStates(c, S);
  if (rc) return rc;
  const long n = (long)c->nState();
  HIP_TRY(hipMemcpyAsync(lam0, lam1, n * 8, hipMemcpyDeviceToDevice, c->stream));
  const double *wNext = nullptr;
  for (int s = c->nRK - 1; s >= 0; s--) {
    const double cs = (c->cfg.timeType < 5) ? c->RKb[s] : (s == c->nRK - 1 ? c->RKa[c->nRK] : 0.0);
    const double an = (s + 1 < c->nRK) ? c->RKa[s + 1] : 0.0;
    hipLaunchKernelGGL(frdev::k_dtj_combine, dim3(nblk(n)), dim3(256), 0, c->stream, lam1, wNext, c->detJ,
                       dtSource(c).first, dtSource(c).second, -cs, -an, c->NF, c->NS, n, c->adjNu);
    HIP_TRY(hipGetLastError());
    double *w = c->adjW[s & 1];
    if ((rc = vjpResidualDev(c, S[s], c->adjNu, w))) return rc;
    hipLaunchKernelGGL(frdev::k_accumulate, dim3(nblk(n)), dim3(256), 0, c->stream, lam0, w, n);
    HIP_TRY(hipGetLastError());
    wNext = w;
  }
  return 0;
}

int createImpl(const fr_config *cfg, const fr_mesh *m, int precision, fr_ctx *c) {
  c->cfg = *cfg;
  c->prec = precision;
  if (precision != FR_PREC_FP64 && precision != FR_PREC_FP32) return fail(FR_EINVAL, "precision must be 0 or 1");
  if (cfg->equation != 0 && cfg->equation != 1) return fail(FR_EINVAL, "equation must be 0 or 1");
  if (cfg->nDims == 3 && cfg->equation == 1 && cfg->riemannType == 1)
    return fail(FR_EINVAL, "Roe not implemented in 3D");
  if (cfg->riemannType != 0 && cfg->riemannType != 1) return fail(FR_EINVAL, "riemannType must be 0 or 1");
  if (cfg->viscous != 0 && cfg->viscous != 1) return fail(FR_EINVAL, "viscous must be 0 or 1");
  if (!m || m->nEles <= 0) return fail(FR_EMESH, "Cannot setup elements array - nEles = 0");
  if (!(cfg->nDims == 3 && frhost::hexNodeCount(m->nNodes)) && !(cfg->nDims == 2 && (m->nNodes == 4 || m->nNodes == 3)))
    return fail(FR_EINVAL, "hexes of 8, 20, 27, 64, 125 or 216 nodes, linear quads (4) or triangles (3)");
  c->tri = (m->nNodes == 3);
  for (int i = 0; i < m->nEles * m->nNodes; i++)
    if (m->c2v[i] < 0 || m->c2v[i] >= m->nVerts) return fail(FR_EMESH, "c2v vertex id out of range");
  // RK tables (input.cpp:424-461)
  switch (cfg->timeType) {
    case 0: c->nRK = 1; c->RKb[0] = 1; break;
    case 2: c->nRK = 2; c->RKa[1] = 1.; c->RKb[0] = 1. / 2.; c->RKb[1] = 1. / 2.; break;
    case 4:
      c->nRK = 4;
      c->RKa[1] = .5; c->RKa[2] = .5; c->RKa[3] = 1.;
      c->RKb[0] = 1. / 6.; c->RKb[1] = 1. / 3.; c->RKb[2] = 1. / 3.; c->RKb[3] = 1. / 6.;
      break;
    case 5: c->nRK = 4; c->RKa[1] = 1. / 4.; c->RKa[2] = 1. / 3.; c->RKa[3] = 1. / 2.; c->RKa[4] = 1.0; break;
    case 6:
      c->nRK = 4;
      if (cfg->order <= 2) { c->RKa[1] = .153; c->RKa[2] = .442; c->RKa[3] = .931; c->RKa[4] = 1.0; }
      else { c->RKa[1] = .158; c->RKa[2] = .477; c->RKa[3] = 1.; c->RKa[4] = 1.; }
      break;
    default: return fail(FR_EINVAL, "Time-Stepping type not supported.");
  }
  std::string err =
      c->tri ? frhost::buildTriOperators(*cfg, c->ops) : frhost::buildOperators(*cfg, c->ops, m->nNodes);
  if (!err.empty()) return fail(FR_EINVAL, err);
  err = frhost::buildConnectivity(*cfg, *m, c->ops, c->conn);
  if (!err.empty()) return fail(FR_EMESH, err);
  c->D = cfg->nDims;
  c->P = cfg->order;
  c->NS = c->ops.nSpts;
  c->NFP = c->ops.nFpts;
  c->NFPF = c->ops.nFptsFace;
  c->NFACE = c->ops.nFacesEle;
  c->NF = (cfg->equation == 0) ? 1 : c->D + 2;
  c->nE = m->nEles;
  if ((double)c->nE * c->NF * c->NFP >= 2147483647.0) return fail(FR_EINVAL, "mesh too large for 32-bit face indices");
  frdev::Phys &ph = c->ph;
  ph.gamma = cfg->gamma; ph.lambda = cfg->lambda; ph.diffD = cfg->diffD;
  ph.advect[0] = cfg->advectVx; ph.advect[1] = cfg->advectVy; ph.advect[2] = cfg->advectVz;
  ph.rhoBound = cfg->rhoBound; ph.uBound = cfg->uBound; ph.vBound = cfg->vBound; ph.wBound = cfg->wBound;
  ph.pBound = cfg->pBound; ph.TWall = cfg->TWall; ph.RGas = cfg->RGas;
  ph.uWall = cfg->uWall; ph.vWall = cfg->vWall; ph.wWall = cfg->wWall;
  ph.penFact = cfg->penFact; ph.tau = cfg->tau; ph.mu_inf = cfg->mu_inf; ph.prandtl = cfg->prandtl;
  ph.rt_inf = cfg->rt_inf; ph.c_sth = cfg->c_sth;
  ph.equation = cfg->equation; ph.riemannType = cfg->riemannType; ph.fixVis = cfg->fixVis; ph.viscFix = cfg->viscFix;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FR_EHIP, "no HIP device available");
  HIP_TRY(hipGetDevice(&c->dev));
  HIP_TRY(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
  c->stream = c->own;

  const int D = c->D, nE = c->nE, NS = c->NS, NFP = c->NFP, nN = m->nNodes;
  const size_t nSt = c->nState();
  int rc;
  if ((rc = dmalloc(&c->U, nSt)) || (rc = dmalloc(&c->U0, nSt)) || (rc = dmalloc(&c->detJ, (size_t)nE * NS)) ||
      (rc = dmalloc(&c->JG64, (size_t)nE * D * D * NS)) || (rc = dmalloc(&c->fgeo64, (size_t)nE * (D + 1) * NFP)) ||
      (rc = dmalloc(&c->ownN, (size_t)nE * D * NFP)) || (rc = dmalloc(&c->wts, NS)) ||
      (rc = dmalloc(&c->oned64, c->ops.oned.size())) || (rc = dmalloc(&c->fnbr, (size_t)nE * c->NFACE)) ||
      (rc = dmalloc(&c->finfo, (size_t)nE * c->NFACE)))
    return rc;
  HIP_TRY(hipMemset(c->U, 0, nSt * 8));
  HIP_TRY(hipMemset(c->U0, 0, nSt * 8));
  const size_t ts = c->tsize();
  for (int b = 0; b < 2; b++) HIP_TRY(hipMalloc(&c->Uf[b], (size_t)nE * c->NF * NFP * ts));
  HIP_TRY(hipMalloc(&c->Us, nSt * ts));
  if ((rc = dmalloc(&c->dtG, 1))) return rc;
  HIP_TRY(hipMemcpy(c->dtG, &cfg->dt, sizeof(double), hipMemcpyHostToDevice));
  if (cfg->viscous) {
    HIP_TRY(hipMalloc(&c->dUf, (size_t)nE * D * c->NF * NFP * ts));
    HIP_TRY(hipMalloc(&c->Fv, (size_t)nE * c->NF * NFP * ts));
  }
  HIP_TRY(hipMemset(c->Us, 0, nSt * ts));
  for (int s = 0; s < c->nRK; s++) {
    HIP_TRY(hipMalloc(&c->divF[s], nSt * ts));
    HIP_TRY(hipMemset(c->divF[s], 0, nSt * ts));
  }
  if (c->prec == FR_PREC_FP32) {
    if ((rc = dmalloc(&c->U32, nSt)) || (rc = dmalloc(&c->JG32, (size_t)nE * D * D * NS)) ||
        (rc = dmalloc(&c->fgeo32, (size_t)nE * (D + 1) * NFP)) || (rc = dmalloc(&c->oned32, c->ops.oned.size())))
      return rc;
    HIP_TRY(hipMemset(c->U32, 0, nSt * 4));
  }
  // device element order: blocks of EPB face-connected elements
  dispatch(D, c->P, c->NF, [&](auto Dc, auto Pc, auto) {
    c->epb = frdev::ElemGeo<decltype(Dc)::value, decltype(Pc)::value>::EPB;
  });
  if (c->tri) c->epb = 256 / frdev::triTPE(c->P);  // frdev::TriGeo<P>::EPB
  {
    std::vector<double> cen((size_t)nE * 3, 0.0);
    for (int e = 0; e < nE; e++)
      for (int n = 0; n < nN; n++)
        for (int d = 0; d < 3; d++) cen[3 * (size_t)e + d] += m->xv[3 * (size_t)m->c2v[(size_t)e * nN + n] + d] / nN;
    c->perm = frhost::blockOrder(c->conn, cen, nE, c->NFACE, c->epb);
    c->iperm.assign(nE, -1);
    for (int i = 0; i < nE; i++) c->iperm[c->perm[i]] = i;
    for (int e = 0; e < nE; e++)
      if (c->iperm[e] < 0) return fail(FR_EMESH, "element order is not a permutation");
  }
  // geometry on device
  {
    std::vector<double> &X = c->X;
    X.assign((size_t)nE * nN * 3, 0.0);
    for (int de = 0; de < nE; de++) {
      const size_t e = (size_t)c->perm[de];
      for (int n = 0; n < nN; n++)
        for (int d = 0; d < 3; d++) X[((size_t)de * nN + n) * 3 + d] = m->xv[3 * (size_t)m->c2v[e * nN + n] + d];
    }
    double *dX, *dS, *dF, *dT;
    int *dNeg;
    if ((rc = dmalloc(&dX, X.size())) || (rc = dmalloc(&dS, c->ops.dshapeSpts.size())) ||
        (rc = dmalloc(&dF, c->ops.dshapeFpts.size())) || (rc = dmalloc(&dT, c->ops.tNorm.size())) ||
        (rc = dmalloc(&dNeg, 1)))
      return rc;
    int32_t *pairEF;
    if ((rc = dmalloc(&pairEF, (size_t)nE * NFP))) return rc;
    HIP_TRY(hipMemcpy(dX, X.data(), X.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dS, c->ops.dshapeSpts.data(), c->ops.dshapeSpts.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dF, c->ops.dshapeFpts.data(), c->ops.dshapeFpts.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dT, c->ops.tNorm.data(), c->ops.tNorm.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(dNeg, 0, sizeof(int)));
    {
      // per-element tables in device order, neighbour ids renumbered
      const int NFACE = c->NFACE;
      std::vector<int32_t> pe((size_t)nE * NFP), fn((size_t)nE * NFACE);
      std::vector<int8_t> fi((size_t)nE * NFACE);
      for (int de = 0; de < nE; de++) {
        const size_t e = (size_t)c->perm[de];
        for (int q = 0; q < NFP; q++) {
          const int32_t v = c->conn.nbr[e * NFP + q];
          pe[(size_t)de * NFP + q] = v < 0 ? v : c->iperm[v / NFP] * NFP + v % NFP;
        }
        for (int f = 0; f < NFACE; f++) {
          const int32_t v = c->conn.fnbr[e * NFACE + f];
          fn[(size_t)de * NFACE + f] = v < 0 ? v : c->iperm[v];
          fi[(size_t)de * NFACE + f] = c->conn.finfo[e * NFACE + f];
        }
      }
      HIP_TRY(hipMemcpy(pairEF, pe.data(), pe.size() * 4, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(c->finfo, fi.data(), fi.size(), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(c->fnbr, fn.data(), fn.size() * 4, hipMemcpyHostToDevice));
    }
    long np = (long)nE * (NS + NFP);
    if (D == 3)
      hipLaunchKernelGGL(frdev::k_metrics<3>, dim3(nblk(np)), dim3(256), 0, c->stream, dX, dS, dF, dT, NS, NFP, nN,
                         nE, c->JG64, c->detJ, c->ownN, c->fgeo64, dNeg);
    else
      hipLaunchKernelGGL(frdev::k_metrics<2>, dim3(nblk(np)), dim3(256), 0, c->stream, dX, dS, dF, dT, NS, NFP, nN,
                         nE, c->JG64, c->detJ, c->ownN, c->fgeo64, dNeg);
    HIP_TRY(hipGetLastError());
    long nq = (long)nE * NFP;
    if (D == 3)
      hipLaunchKernelGGL(frdev::k_fgeo<3>, dim3(nblk(nq)), dim3(256), 0, c->stream, c->ownN, pairEF, c->finfo, NFP,
                         c->NFPF, nE, c->fgeo64);
    else
      hipLaunchKernelGGL(frdev::k_fgeo<2>, dim3(nblk(nq)), dim3(256), 0, c->stream, c->ownN, pairEF, c->finfo, NFP,
                         c->NFPF, nE, c->fgeo64);
    HIP_TRY(hipGetLastError());
    int neg = 0;
    HIP_TRY(hipMemcpyAsync(&neg, dNeg, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(dX); (void)hipFree(dS); (void)hipFree(dF); (void)hipFree(dT); (void)hipFree(dNeg);
    (void)hipFree(pairEF);
    if (neg) return fail(FR_EJACOBIAN, "Negative Jacobian at solution points.");
  }
  // affine meshes: compact per-element / per-face geometry for the stage
  // kernels (every linear triangle is affine)
  {
    const size_t nJ = (size_t)nE * D * D, nG = (size_t)nE * (D + 1) * c->NFACE;
    int8_t *dOk;
    if ((rc = dmalloc(&c->JGc64, nJ)) || (rc = dmalloc(&c->detJc, nE)) || (rc = dmalloc(&c->fgeoc64, nG)) ||
        (rc = dmalloc(&dOk, nE)))
      return rc;
    const double tol = 1e-13;
    if (D == 3)
      hipLaunchKernelGGL(frdev::k_affine<3>, dim3(nblk(nE)), dim3(256), 0, c->stream, c->JG64, c->detJ, c->fgeo64,
                         NS, NFP, c->NFPF, c->NFACE, nE, tol, c->JGc64, c->detJc, c->fgeoc64, dOk);
    else
      hipLaunchKernelGGL(frdev::k_affine<2>, dim3(nblk(nE)), dim3(256), 0, c->stream, c->JG64, c->detJ, c->fgeo64,
                         NS, NFP, c->NFPF, c->NFACE, nE, tol, c->JGc64, c->detJc, c->fgeoc64, dOk);
    HIP_TRY(hipGetLastError());
    std::vector<int8_t> ok(nE);
    HIP_TRY(hipMemcpyAsync(ok.data(), dOk, nE, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(dOk);
    c->affineMesh = true;
    for (int e = 0; e < nE; e++) c->affineMesh = c->affineMesh && ok[e];
    if (c->affineMesh && c->prec == FR_PREC_FP32) {
      if ((rc = dmalloc(&c->JGc32, nJ)) || (rc = dmalloc(&c->fgeoc32, nG))) return rc;
      hipLaunchKernelGGL(frdev::k_to_float, dim3(nblk((long)nJ)), dim3(256), 0, c->stream, c->JGc64, c->JGc32,
                         (long)nJ);
      hipLaunchKernelGGL(frdev::k_to_float, dim3(nblk((long)nG)), dim3(256), 0, c->stream, c->fgeoc64, c->fgeoc32,
                         (long)nG);
      HIP_TRY(hipGetLastError());
    }
  }
  // nodal geometry of linear quads and hexes: the coefficients of each
  // element's map (frdev::LinearMap) from its corner nodes, as sums of node
  // differences (node 0 subtracted first) so that they keep the precision of
  // the element's size
  if (!c->tri && ((D == 3 && nN == 8) || (D == 2 && nN == 4))) {
    const int NC = D == 3 ? frdev::LinearMap<3>::NC : frdev::LinearMap<2>::NC;
    static const int SI[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, SJ[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                     SK[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    std::vector<double> cf((size_t)nE * NC, 0.0);
    for (int de = 0; de < nE; de++) {
      const double *Xe = &c->X[(size_t)de * nN * 3];
      double *o = &cf[(size_t)de * NC];
      for (int n = 1; n < nN; n++)
        for (int d = 0; d < D; d++) {
          const double y = Xe[n * 3 + d] - Xe[d];
          if (D == 3) {
            const int s[7] = {SI[n], SJ[n], SK[n], SI[n] * SJ[n], SI[n] * SK[n], SJ[n] * SK[n], SI[n] * SJ[n] * SK[n]};
            for (int k = 0; k < 7; k++) o[3 * k + d] += s[k] * y;
          } else {
            const int s[3] = {SI[n], SJ[n], SI[n] * SJ[n]};
            for (int k = 0; k < 3; k++) o[2 * k + d] += s[k] * y;
          }
        }
      for (int k = 0; k < NC; k++) o[k] *= (D == 3) ? 0.125 : 0.25;
    }
    if ((rc = dmalloc(&c->geoN64, cf.size()))) return rc;
    HIP_TRY(hipMemcpy(c->geoN64, cf.data(), cf.size() * 8, hipMemcpyHostToDevice));
    if (c->prec == FR_PREC_FP32) {
      std::vector<float> cf32(cf.begin(), cf.end());
      if ((rc = dmalloc(&c->geoN32, cf32.size()))) return rc;
      HIP_TRY(hipMemcpy(c->geoN32, cf32.data(), cf32.size() * 4, hipMemcpyHostToDevice));
    }
  }
  HIP_TRY(hipMemcpy(c->wts, c->ops.wts.data(), NS * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->oned64, c->ops.oned.data(), c->ops.oned.size() * 8, hipMemcpyHostToDevice));
  if (c->prec == FR_PREC_FP32) {
    std::vector<float> o32(c->ops.oned.begin(), c->ops.oned.end());
    HIP_TRY(hipMemcpy(c->oned32, o32.data(), o32.size() * 4, hipMemcpyHostToDevice));
    long n1 = (long)nE * D * D * NS, n2 = (long)nE * (D + 1) * NFP;
    hipLaunchKernelGGL(frdev::k_to_float, dim3(nblk(n1)), dim3(256), 0, c->stream, c->JG64, c->JG32, n1);
    hipLaunchKernelGGL(frdev::k_to_float, dim3(nblk(n2)), dim3(256), 0, c->stream, c->fgeo64, c->fgeo32, n2);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

// Buffers of CFL time stepping (allocated by the first fr_calc_dt).
int cflBuffers(fr_ctx *c) {
  int rc;
  const size_t nb = nblk(c->nE);
  if ((!c->dtE && (rc = dmalloc(&c->dtE, (size_t)c->nE))) ||
      (!c->waveSp && (rc = dmalloc(&c->waveSp, (size_t)c->nE * c->NFP))) ||
      (!c->dtNum && (rc = dmalloc(&c->dtNum, (size_t)c->nE))) || (!c->dtPart && (rc = dmalloc(&c->dtPart, nb))))
    return rc;
  if (!c->wsurf) {
    if ((rc = dmalloc(&c->wsurf, c->ops.wtsFace.size()))) return rc;
    HIP_TRY(hipMemcpy(c->wsurf, c->ops.wtsFace.data(), c->ops.wtsFace.size() * 8, hipMemcpyHostToDevice));
  }
  return 0;
}

}  // namespace

namespace frimpl {

int needPart(fr_ctx *c, size_t n) {
  if (c->partCap >= n) return 0;
  if (c->part) (void)hipFree(c->part);
  c->part = nullptr;
  c->partCap = 0;
  int rc = dmalloc(&c->part, n);
  if (!rc) c->partCap = n;
  return rc;
}

int dtPrepare(fr_ctx *c, hipStream_t st) {
  if (c->dtType == 0 || c->dtReady || c->dtFrozen) return 0;
  // getCFLLimit (global.cpp:45-68)
  static const double lim[6] = {1.393, 0.464, 0.235, 0.1453, 0.100, 0.0736};
  const double factor = c->cfl * lim[c->P] * 2;
  const unsigned nb = nblk(c->nE);
  hipError_t err = hipSuccess;
  if (c->tri) {
    // triangles (no reference counterpart): the wave speeds of the current
    // state at every update
    int rc = launchTriDt(c, c->wsurf, factor, st);
    if (rc) return rc;
    hipLaunchKernelGGL(frdev::k_block_min, dim3(nb), dim3(256), 0, st, c->dtE, c->nE, c->dtPart);
  } else if (c->dtFirst) {
    // the first update: ele::calcWaveSpFpts from the state (solver.cpp:294-297)
    dispatch(c->D, c->P, c->NF, [&](auto Dc, auto Pc, auto NFc) {
      constexpr int D = decltype(Dc)::value, P = decltype(Pc)::value, NF = decltype(NFc)::value;
      hipLaunchKernelGGL((frdev::k_dt<D, P, NF>), dim3(nb), dim3(256), 0, st, c->U, c->ownN, c->fgeo64, c->detJ,
                         c->wts, c->wsurf, c->oned64, factor, c->ph, c->nE, c->dtE, c->waveSp, c->dtNum);
    });
    if ((err = hipGetLastError()) != hipSuccess) return fail(FR_EHIP, std::string("k_dt: ") + hipGetErrorString(err));
    hipLaunchKernelGGL(frdev::k_block_min, dim3(nb), dim3(256), 0, st, c->dtE, c->nE, c->dtPart);
  } else {
    // later updates: the wave speeds the last residual's face fluxes stored
    dispatch(c->D, c->P, c->NF, [&](auto Dc, auto Pc, auto) {
      using G = frdev::ElemGeo<decltype(Dc)::value, decltype(Pc)::value>;
      hipLaunchKernelGGL((frdev::k_dt_ws<G::NFP, G::NFPF>), dim3(nb), dim3(256), 0, st, c->waveSp, c->fgeo64,
                         c->wsurf, c->dtNum, c->D, c->nE, c->dtE, c->dtPart);
    });
  }
  hipLaunchKernelGGL(frdev::k_dt_min, dim3(1), dim3(256), 0, st, c->dtPart, (int)nb, c->dtG);
  if ((err = hipGetLastError()) != hipSuccess) return fail(FR_EHIP, std::string("time step: ") + hipGetErrorString(err));
  c->dtReady = true;
  return 0;
}

}  // namespace frimpl

extern "C" {

const char *fr_last_error(void) { return g_err.c_str(); }

int fr_create(const fr_config *cfg, const fr_mesh *mesh, int precision, fr_ctx **out) {
  if (!cfg || !mesh || !out) return fail(FR_EINVAL, "null argument");
  *out = nullptr;
  fr_ctx *c = new fr_ctx();
  int rc = createImpl(cfg, mesh, precision, c);
  if (rc) {
    std::string msg = g_err;
    freeAll(c);
    delete c;
    g_err = msg;
    return rc;
  }
  *out = c;
  return 0;
}

int fr_destroy(fr_ctx *c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  // a batch holding this context can no longer step
  for (fr_batch *b : c->batches) {
    b->broken = true;
    for (size_t k = 0; k < b->ctx.size(); k++)
      if (b->ctx[k] == c) b->alive[k] = false;
  }
  freeAll(c);
  delete c;
  return 0;
}

int fr_get_sizes(fr_ctx *c, int *o) {
  if (!c || !o) return fail(FR_EINVAL, "null argument");
  o[0] = c->NS; o[1] = c->NFP; o[2] = c->nE; o[3] = c->NF;
  o[4] = c->D; o[5] = (int)c->conn.faces.size(); o[6] = c->conn.nIntFaces; o[7] = c->nRK;
  return 0;
}

int fr_get_operator(fr_ctx *c, int which, double *out) {
  if (!c || !out) return fail(FR_EINVAL, "null argument");
  const std::vector<double> *v = nullptr;
  switch (which) {
    case 0: v = &c->ops.spts_to_fpts; break;
    case 1: v = &c->ops.grad_spts; break;
    case 2: v = &c->ops.extrapolateFn; break;
    case 3: v = &c->ops.correction; break;
    case 4: v = &c->ops.correctU; break;
    default: return fail(FR_EINVAL, "unknown operator id");
  }
  std::memcpy(out, v->data(), v->size() * 8);
  return 0;
}

int fr_get_geometry(fr_ctx *c, int which, double *out) {
  if (!c || !out) return fail(FR_EINVAL, "null argument");
  const int D = c->D, nE = c->nE, NS = c->NS, NFP = c->NFP;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (which == 0) {
    std::vector<double> h((size_t)nE * D * D * NS);
    HIP_TRY(hipMemcpy(h.data(), c->JG64, h.size() * 8, hipMemcpyDeviceToHost));
    for (int d1 = 0; d1 < D; d1++)
      for (int s = 0; s < NS; s++)
        for (int e = 0; e < nE; e++)
          for (int d2 = 0; d2 < D; d2++)
            out[(((size_t)d1 * NS + s) * nE + e) * D + d2] = h[((size_t)c->iperm[e] * D * D + d1 * D + d2) * NS + s];
  } else if (which == 1) {
    std::vector<double> h((size_t)nE * NS);
    HIP_TRY(hipMemcpy(h.data(), c->detJ, h.size() * 8, hipMemcpyDeviceToHost));
    for (int s = 0; s < NS; s++)
      for (int e = 0; e < nE; e++) out[(size_t)s * nE + e] = h[(size_t)c->iperm[e] * NS + s];
  } else if (which == 2) {
    std::vector<double> h((size_t)nE * D * NFP);
    HIP_TRY(hipMemcpy(h.data(), c->ownN, h.size() * 8, hipMemcpyDeviceToHost));
    devToRef(c, h.data(), out, D, NFP);
  } else if (which == 3) {
    std::vector<double> h((size_t)nE * (D + 1) * NFP);
    HIP_TRY(hipMemcpy(h.data(), c->fgeo64, h.size() * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < NFP; q++)
      for (int e = 0; e < nE; e++) out[(size_t)q * nE + e] = h[((size_t)c->iperm[e] * (D + 1) + D) * NFP + q];
  } else if (which == 4) {
    // pos_spts(spt, ele, dim) = sum_n shape(spt, n) * node(n, ele, dim)
    // (solver::setPosSptsFpts, solver.cpp:904-951)
    const int nN = c->ops.nNodes;
    for (int s = 0; s < NS; s++)
      for (int e = 0; e < nE; e++)
        for (int d = 0; d < D; d++) {
          double acc = 0.0;
          for (int n = 0; n < nN; n++)
            acc += c->ops.shapeSpts[(size_t)s * nN + n] * c->X[((size_t)c->iperm[e] * nN + n) * 3 + d];
          out[((size_t)s * nE + e) * D + d] = acc;
        }
  } else {
    return fail(FR_EINVAL, "unknown geometry id");
  }
  return 0;
}

int fr_get_element_order(fr_ctx *c, int *order) {
  if (!c || !order) return fail(FR_EINVAL, "null argument");
  std::memcpy(order, c->perm.data(), c->perm.size() * sizeof(int));
  return 0;
}

int fr_get_faces(fr_ctx *c, int *o6) {
  if (!c || !o6) return fail(FR_EINVAL, "null argument");
  for (size_t f = 0; f < c->conn.faces.size(); f++) {
    const frhost::FaceRec &r = c->conn.faces[f];
    o6[6 * f + 0] = r.eL; o6[6 * f + 1] = r.fL; o6[6 * f + 2] = r.eR;
    o6[6 * f + 3] = r.fR; o6[6 * f + 4] = r.rot; o6[6 * f + 5] = r.bc;
  }
  return 0;
}

int fr_set_state(fr_ctx *c, const double *U, int layout) {
  if (!c || !U) return fail(FR_EINVAL, "null argument");
  const size_t n = c->nState();
  if (layout != FR_LAYOUT_REF && layout != FR_LAYOUT_DEVICE) return fail(FR_EINVAL, "unknown layout");
  setEnergyOffset(c, U, layout);
  stateChanged(c);
  if (layout == FR_LAYOUT_REF) {
    std::vector<double> h(n);
    refToDev(c, U, h.data(), c->NF, c->NS);
    HIP_TRY(hipMemcpyAsync(c->U, h.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  } else if (layout == FR_LAYOUT_DEVICE) {
    HIP_TRY(hipMemcpyAsync(c->U, U, n * 8, hipMemcpyHostToDevice, c->stream));
  } else {
    return fail(FR_EINVAL, "unknown layout");
  }
  int rc = syncU32(c);
  if (rc) return rc;
  c->traceOf = -1;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int fr_get_state(fr_ctx *c, double *U, int layout) {
  if (!c || !U) return fail(FR_EINVAL, "null argument");
  const size_t n = c->nState();
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (layout == FR_LAYOUT_REF) {
    std::vector<double> h(n);
    HIP_TRY(hipMemcpy(h.data(), c->U, n * 8, hipMemcpyDeviceToHost));
    devToRef(c, h.data(), U, c->NF, c->NS);
  } else {
    HIP_TRY(hipMemcpy(U, c->U, n * 8, hipMemcpyDeviceToHost));
  }
  return 0;
}

int fr_calc_residual(fr_ctx *c, int step) {
  if (!c) return fail(FR_EINVAL, "null argument");
  return residual(c, step);
}

int fr_get_residual(fr_ctx *c, int step, double *out, int layout) {
  if (!c || !out) return fail(FR_EINVAL, "null argument");
  if (step < 0 || step >= c->nRK) return fail(FR_EINVAL, "step out of range");
  if (step == c->nRK - 1 && c->finalDropped)
    return fail(FR_EINVAL, "divF_spts of the final stage was not kept (fr_keep_final_residual)");
  const size_t n = c->nState();
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<double> h(n);
  if (c->prec == FR_PREC_FP32) {
    std::vector<float> f(n);
    HIP_TRY(hipMemcpy(f.data(), c->divF[step], n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) h[i] = f[i];
  } else {
    HIP_TRY(hipMemcpy(h.data(), c->divF[step], n * 8, hipMemcpyDeviceToHost));
  }
  if (layout == FR_LAYOUT_REF) devToRef(c, h.data(), out, c->NF, c->NS);
  else std::memcpy(out, h.data(), n * 8);
  return 0;
}

// Debug view of the residual's intermediate arrays (solver.hpp:55-61), for
// the current state: one residual evaluation (as fr_calc_residual(ctx, 0))
// with the stage kernel's (and on the viscous path the gradient kernel's)
// debug outputs switched on, into buffers allocated for the call.
int fr_get_array(fr_ctx *c, int which, double *out) {
  if (!c || !out) return fail(FR_EINVAL, "null argument");
  if (which < 2 || which > 6) return fail(FR_EINVAL, "which: 2 F_spts, 3 Fn_fpts, 4 disFn_fpts, 5 dU_spts, 6 dU_fpts");
  if (c->prec != FR_PREC_FP64) return fail(FR_EINVAL, "fr_get_array: FP64 contexts only");
  if (c->tri) return fail(FR_EINVAL, "fr_get_array: quads and hexes only");
  if (which >= 5 && !c->cfg.viscous) return fail(FR_EINVAL, "fr_get_array: the gradient exists on the viscous path only");
  const int D = c->D, NF = c->NF, NS = c->NS, NFP = c->NFP, nE = c->nE;
  const size_t nSp = (size_t)nE * D * NF * NS, nFp = (size_t)nE * NF * NFP;
  double *dF = nullptr, *dFn = nullptr, *dJ = nullptr, *dUs = nullptr;
  auto release = [&]() {
    for (double *p : {dF, dFn, dJ, dUs})
      if (p) (void)hipFree(p);
  };
  if (hipMalloc(&dF, nSp * 8) != hipSuccess || hipMalloc(&dFn, nFp * 8) != hipSuccess ||
      hipMalloc(&dJ, nFp * 8) != hipSuccess || (c->cfg.viscous && hipMalloc(&dUs, nSp * 8) != hipSuccess)) {
    release();
    return fail(FR_EHIP, "fr_get_array: device allocation failed");
  }
  // the residual's face fluxes store new wave speeds, as in residual()
  if (c->dtType != 0 && !c->dtFirst && !c->dtFrozen) c->dtReady = false;
  int rc = ensureTraces<double>(c, 0);
  if (!rc) {
    frdev::StageArgs<double> a = makeStageArgs<double>(c, 0, c->cur, 0, 0, 0.0);
    a.dbgF = dF;
    a.dbgFn = dFn;
    a.dbgJ = dJ;
    rc = launchStageDbg(c, a, dUs);
  }
  hipError_t err = hipStreamSynchronize(c->stream);
  std::vector<double> h;
  if (!rc && err == hipSuccess) {
    const double *src = which == 2 ? dF : which == 3 ? dFn : which == 4 ? dJ : which == 5 ? dUs : (const double *)c->dUf;
    h.resize(which == 6 ? (size_t)nE * D * NF * NFP : (which == 2 || which == 5) ? nSp : nFp);
    err = hipMemcpy(h.data(), src, h.size() * 8, hipMemcpyDeviceToHost);
  }
  release();
  if (rc) return rc;
  if (err != hipSuccess) return fail(FR_EHIP, std::string("fr_get_array: ") + hipGetErrorString(err));
  if (which == 3 || which == 4) {
    devToRef(c, h.data(), out, NF, NFP);
    return 0;
  }
  // (dim, point, ele, field) from the device's [ele][dim][field][point]
  const int NP = which == 6 ? NFP : NS;
  for (int d = 0; d < D; d++)
    for (int p = 0; p < NP; p++)
      for (int e = 0; e < nE; e++) {
        const size_t de = (size_t)c->iperm[e];
        for (int k = 0; k < NF; k++)
          out[(((size_t)d * NP + p) * nE + e) * NF + k] = h[((de * D + d) * NF + k) * NP + p];
      }
  return 0;
}

int fr_get_traces(fr_ctx *c, double *out) {
  if (!c || !out) return fail(FR_EINVAL, "null argument");
  const size_t n = (size_t)c->nE * c->NF * c->NFP;
  // the stage kernels exchange the traces of faces inside an element block
  // through LDS and leave them out of the trace buffer: extrapolate afresh
  c->traceOf = -1;
  int rc = (c->prec == FR_PREC_FP32) ? ensureTraces<float>(c, 0) : ensureTraces<double>(c, 0);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<double> h(n);
  if (c->prec == FR_PREC_FP32) {
    std::vector<float> f(n);
    HIP_TRY(hipMemcpy(f.data(), c->Uf[c->cur], n * 4, hipMemcpyDeviceToHost));
    // the FP32 energy traces are offset (Phys::eOff)
    for (size_t i = 0; i < n; i++)
      h[i] = (double)f[i] + (((int)((i / c->NFP) % c->NF) == c->NF - 1) ? c->ph.eOff : 0.0);
  } else {
    HIP_TRY(hipMemcpy(h.data(), c->Uf[c->cur], n * 8, hipMemcpyDeviceToHost));
  }
  devToRef(c, h.data(), out, c->NF, c->NFP);
  return 0;
}

int fr_update(fr_ctx *c, int nSteps) {
  if (!c) return fail(FR_EINVAL, "null argument");
  for (int i = 0; i < nSteps; i++) {
    int rc = updateOnce(c);
    if (rc) return rc;
  }
  return 0;
}

int fr_time_step_a(fr_ctx *c, int step, double rkval) {
  if (!c) return fail(FR_EINVAL, "null argument");
  if (step < 0 || step >= c->nRK) return fail(FR_EINVAL, "step out of range");
  stateChanged(c);
  return c->prec == FR_PREC_FP32 ? rkA<float>(c, step, rkval) : rkA<double>(c, step, rkval);
}

int fr_time_step_b(fr_ctx *c, int step) {
  if (!c) return fail(FR_EINVAL, "null argument");
  if (step < 0 || step >= c->nRK) return fail(FR_EINVAL, "step out of range");
  long n = (long)c->nState();
  stateChanged(c);
  if (c->prec == FR_PREC_FP32)
    hipLaunchKernelGGL(frdev::k_rk_b1<float>, dim3(nblk(n)), dim3(256), 0, c->stream, c->U,
                       (const float *)c->divF[step], c->detJ, c->U32, c->RKb[step], dtSource(c).first,
                       dtSource(c).second, c->NF, c->NS, n, c->ph.eOff);
  else
    hipLaunchKernelGGL(frdev::k_rk_b1<double>, dim3(nblk(n)), dim3(256), 0, c->stream, c->U,
                       (const double *)c->divF[step], c->detJ, c->U32, c->RKb[step], dtSource(c).first,
                       dtSource(c).second, c->NF, c->NS, n, c->ph.eOff);
  HIP_TRY(hipGetLastError());
  c->traceOf = -1;
  return 0;
}

int fr_copy_u_u0(fr_ctx *c) {
  if (!c) return fail(FR_EINVAL, "null argument");
  HIP_TRY(hipMemcpyAsync(c->U0, c->U, c->nState() * 8, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

int fr_copy_u0_u(fr_ctx *c) {
  if (!c) return fail(FR_EINVAL, "null argument");
  stateChanged(c);
  HIP_TRY(hipMemcpyAsync(c->U, c->U0, c->nState() * 8, hipMemcpyDeviceToDevice, c->stream));
  c->traceOf = -1;
  return syncU32(c);
}

int fr_reduce(fr_ctx *c, int kind, double *out) {
  if (!c || !out) return fail(FR_EINVAL, "null argument");
  const int nb = (int)nblk(c->nE);
  int rc;
  if (kind >= 1 && kind <= 3) {
    if ((rc = needPart(c, (size_t)nb * 5))) return rc;
    if (c->prec == FR_PREC_FP32)
      hipLaunchKernelGGL(frdev::k_norm<float>, dim3(nb), dim3(256), 0, c->stream, (const float *)c->divF[0],
                         c->detJ, c->wts, kind, c->NF, c->NS, c->nE, c->part);
    else
      hipLaunchKernelGGL(frdev::k_norm<double>, dim3(nb), dim3(256), 0, c->stream, (const double *)c->divF[0],
                         c->detJ, c->wts, kind, c->NF, c->NS, c->nE, c->part);
    HIP_TRY(hipGetLastError());
    std::vector<double> h((size_t)nb * 5);
    HIP_TRY(hipMemcpyAsync(h.data(), c->part, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->NF; k++) {
      double r = 0;
      for (int b = 0; b < nb; b++) r = (kind == 3) ? frdev::maxKeepNaN(r, h[(size_t)b * 5 + k]) : r + h[(size_t)b * 5 + k];
      if (kind == 2) r = std::sqrt(r);
      if (std::isnan(r)) return fail(FR_ENAN, "NaN Encountered in Solution Residual!");
      out[k] = r;
    }
    return 0;
  }
  if (kind == 10) {
    if (c->cfg.equation != 1) return fail(FR_EINVAL, "kinetic energy needs the Euler / Navier-Stokes equations");
    if ((rc = needPart(c, (size_t)nb * 2))) return rc;
    if (c->D == 3)
      hipLaunchKernelGGL(frdev::k_ke<3>, dim3(nb), dim3(256), 0, c->stream, c->U, c->detJ, c->wts, c->NS, c->nE,
                         c->part);
    else
      hipLaunchKernelGGL(frdev::k_ke<2>, dim3(nb), dim3(256), 0, c->stream, c->U, c->detJ, c->wts, c->NS, c->nE,
                         c->part);
    HIP_TRY(hipGetLastError());
    std::vector<double> h((size_t)nb * 2);
    HIP_TRY(hipMemcpyAsync(h.data(), c->part, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double ke = 0, vol = 0;
    for (int b = 0; b < nb; b++) {
      ke += h[(size_t)b * 2];
      vol += h[(size_t)b * 2 + 1];
    }
    out[0] = ke;
    out[1] = vol;
    return 0;
  }
  return fail(FR_EINVAL, "unknown reduction kind");
}

int fr_set_dt(fr_ctx *c, double dt) {
  if (!c) return fail(FR_EINVAL, "null argument");
  if (!(dt > 0)) return fail(FR_EINVAL, "dt must be positive");
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->cfg.dt = dt;
  c->dtType = 0;
  c->dtFirst = c->dtReady = false;
  c->gen++;
  HIP_TRY(hipMemcpy(c->dtG, &dt, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int fr_calc_dt(fr_ctx *c, int dtType, double CFL, double *dtMin) {
  if (!c) return fail(FR_EINVAL, "null argument");
  if (dtType != 1 && dtType != 2) return fail(FR_EINVAL, "dtType must be 1 (global) or 2 (local)");
  if (c->P < 0 || c->P > 5) return fail(FR_EINVAL, "CFL limit not available for this order!");
  if (!(CFL > 0)) return fail(FR_EINVAL, "CFL must be positive");
  int rc;
  if ((rc = cflBuffers(c))) return rc;
  c->cfl = CFL;
  c->dtType = dtType;
  c->dtFirst = true;
  c->dtReady = false;
  c->gen++;
  if ((rc = dtPrepare(c, c->stream))) return rc;
  double m = 0;
  HIP_TRY(hipMemcpyAsync(&m, c->dtG, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (!(m > 0) || std::isinf(m)) return fail(FR_ENAN, "CFL time step is not positive and finite");
  c->cfg.dt = m;
  if (dtMin) *dtMin = m;
  return 0;
}

int fr_get_dt(fr_ctx *c, double *dtEle) {
  if (!c || !dtEle) return fail(FR_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  // params->dt of the last update (or of fr_calc_dt / fr_set_dt)
  HIP_TRY(hipMemcpy(&c->cfg.dt, c->dtG, sizeof(double), hipMemcpyDeviceToHost));
  if (c->dtType != 2) {
    for (int e = 0; e < c->nE; e++) dtEle[e] = c->cfg.dt;
    return 0;
  }
  std::vector<double> h(c->nE);
  HIP_TRY(hipMemcpy(h.data(), c->dtE, h.size() * 8, hipMemcpyDeviceToHost));
  for (int e = 0; e < c->nE; e++) dtEle[e] = h[c->iperm[e]];
  return 0;
}

int fr_vjp_residual(fr_ctx *c, const double *lam, double *out) {
  if (!c || !lam || !out) return fail(FR_EINVAL, "null argument");
  int rc = adjCheck(c);
  if (rc) return rc;
  const size_t n = c->nState();
  std::vector<double> h(n);
  refToDev(c, lam, h.data(), c->NF, c->NS);
  HIP_TRY(hipMemcpyAsync(c->adjLam, h.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  if ((rc = vjpResidualDev(c, c->U, c->adjLam, c->adjOut))) return rc;
  HIP_TRY(hipMemcpyAsync(h.data(), c->adjOut, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  devToRef(c, h.data(), out, c->NF, c->NS);
  return 0;
}

int fr_vjp_step(fr_ctx *c, const double *lam1, double *lam0) {
  if (!c || !lam1 || !lam0) return fail(FR_EINVAL, "null argument");
  int rc = adjCheck(c);
  if (rc) return rc;
  // the step's time step is the one the next update takes (a constant of
  // the step); forming it here leaves it for that update
  if ((rc = dtPrepare(c, c->stream))) return rc;
  const size_t n = c->nState();
  std::vector<double> h(n);
  refToDev(c, lam1, h.data(), c->NF, c->NS);
  HIP_TRY(hipMemcpyAsync(c->adjLam, h.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  if ((rc = vjpStepDev(c, c->adjLam, c->adjOut))) return rc;
  HIP_TRY(hipMemcpyAsync(h.data(), c->adjOut, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  devToRef(c, h.data(), lam0, c->NF, c->NS);
  return 0;
}

int fr_vjp_steps(fr_ctx *c, int nSteps, int every, const double *lamN, double *lam0) {
  if (!c || !lamN || !lam0) return fail(FR_EINVAL, "null argument");
  if (nSteps < 1 || every < 1) return fail(FR_EINVAL, "nSteps and the checkpoint interval must be positive");
  int rc = adjCheck(c);
  if (rc) return rc;
  const size_t n = c->nState();
  const int nCk = (nSteps + every - 1) / every;
  // CFL time stepping: each step's time step is recorded in the forward sweep
  // and given back to its recomputation and its adjoint
  const bool cfl = c->dtType != 0;
  const size_t nDt = (c->dtType == 2) ? (size_t)c->nE : 1;
  double *dtArr = (c->dtType == 2) ? c->dtE : c->dtG;
  std::vector<double *> ck(nCk, nullptr), seg(every, nullptr);
  double *uN = nullptr, *dtHist = nullptr, *wsKeep = nullptr;
  auto cleanup = [&]() {
    for (double *p : ck) if (p) (void)hipFree(p);
    for (double *p : seg) if (p) (void)hipFree(p);
    if (uN) (void)hipFree(uN);
    if (dtHist) (void)hipFree(dtHist);
    if (wsKeep) (void)hipFree(wsKeep);
    c->dtFrozen = false;
  };
  auto copy = [&](double *dst, const double *src) {
    return hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToDevice, c->stream);
  };
  auto copyDt = [&](double *dst, const double *src) {
    return hipMemcpyAsync(dst, src, nDt * 8, hipMemcpyDeviceToDevice, c->stream);
  };
  for (int i = 0; i < nCk; i++)
    if ((rc = dmalloc(&ck[i], n))) { cleanup(); return rc; }
  for (int i = 0; i < every; i++)
    if ((rc = dmalloc(&seg[i], n))) { cleanup(); return rc; }
  if ((rc = dmalloc(&uN, n))) { cleanup(); return rc; }
  if (cfl && ((rc = dmalloc(&dtHist, nDt * nSteps)) || (!c->tri && (rc = dmalloc(&wsKeep, (size_t)c->nE * c->NFP))))) {
    cleanup();
    return rc;
  }
  // forward sweep, keeping the state at every `every`-th step
  for (int i = 0; i < nSteps; i++) {
    if (i % every == 0 && copy(ck[i / every], c->U) != hipSuccess) { cleanup(); return fail(FR_EHIP, "copy"); }
    if ((rc = updateOnce(c))) { cleanup(); return rc; }
    if (cfl && copyDt(dtHist + (size_t)i * nDt, dtArr) != hipSuccess) { cleanup(); return fail(FR_EHIP, "copy"); }
  }
  if (copy(uN, c->U) != hipSuccess) { cleanup(); return fail(FR_EHIP, "copy"); }
  if (wsKeep && hipMemcpyAsync(wsKeep, c->waveSp, (size_t)c->nE * c->NFP * 8, hipMemcpyDeviceToDevice,
                               c->stream) != hipSuccess) {
    cleanup();
    return fail(FR_EHIP, "copy");
  }
  // reverse sweep, recomputing each segment's step states from its checkpoint
  // with the recorded time steps
  c->dtFrozen = cfl;
  std::vector<double> h(n);
  refToDev(c, lamN, h.data(), c->NF, c->NS);
  HIP_TRY(hipMemcpyAsync(c->adjLam, h.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  double *lamA = c->adjLam, *lamB = c->adjOut;
  for (int k = nCk - 1; k >= 0; k--) {
    const int start = k * every, end = std::min(start + every, nSteps);
    if (copy(c->U, ck[k]) != hipSuccess) { cleanup(); return fail(FR_EHIP, "copy"); }
    c->traceOf = -1;
    for (int i = start; i < end; i++) {
      if (copy(seg[i - start], c->U) != hipSuccess) { cleanup(); return fail(FR_EHIP, "copy"); }
      if (i + 1 < end) {
        if (cfl && copyDt(dtArr, dtHist + (size_t)i * nDt) != hipSuccess) { cleanup(); return fail(FR_EHIP, "copy"); }
        if ((rc = updateOnce(c))) { cleanup(); return rc; }
      }
    }
    for (int i = end - 1; i >= start; i--) {
      if (copy(c->U, seg[i - start]) != hipSuccess) { cleanup(); return fail(FR_EHIP, "copy"); }
      if (cfl && copyDt(dtArr, dtHist + (size_t)i * nDt) != hipSuccess) { cleanup(); return fail(FR_EHIP, "copy"); }
      if ((rc = vjpStepDev(c, lamA, lamB))) { cleanup(); return rc; }
      std::swap(lamA, lamB);
    }
  }
  HIP_TRY(hipMemcpyAsync(h.data(), lamA, n * 8, hipMemcpyDeviceToHost, c->stream));
  if (copy(c->U, uN) != hipSuccess) { cleanup(); return fail(FR_EHIP, "copy"); }
  // the context as the forward sweep left it: the last step's time step and
  // the wave speeds of its last residual
  if (cfl) {
    if (copyDt(dtArr, dtHist + (size_t)(nSteps - 1) * nDt) != hipSuccess ||
        (wsKeep && hipMemcpyAsync(c->waveSp, wsKeep, (size_t)c->nE * c->NFP * 8, hipMemcpyDeviceToDevice,
                                  c->stream) != hipSuccess)) {
      cleanup();
      return fail(FR_EHIP, "copy");
    }
    c->dtReady = false;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->traceOf = -1;
  cleanup();
  devToRef(c, h.data(), lam0, c->NF, c->NS);
  return syncU32(c);
}

int fr_reduce_gradient(fr_ctx *c, int kind, double *out) {
  if (!c || !out) return fail(FR_EINVAL, "null argument");
  if (kind != 10) return fail(FR_EINVAL, "only the kinetic-energy reduction (10) has a gradient");
  if (c->cfg.equation != 1) return fail(FR_EINVAL, "kinetic energy needs the Euler / Navier-Stokes equations");
  const size_t n = c->nState();
  double *g;
  int rc = dmalloc(&g, n);
  if (rc) return rc;
  const long nPts = (long)c->nE * c->NS;
  if (c->D == 3)
    hipLaunchKernelGGL(frdev::k_ke_grad<3>, dim3(nblk(nPts)), dim3(256), 0, c->stream, c->U, c->detJ, c->wts, c->NS,
                       nPts, g);
  else
    hipLaunchKernelGGL(frdev::k_ke_grad<2>, dim3(nblk(nPts)), dim3(256), 0, c->stream, c->U, c->detJ, c->wts, c->NS,
                       nPts, g);
  std::vector<double> h(n);
  hipError_t e1 = hipGetLastError();
  hipError_t e2 = hipMemcpyAsync(h.data(), g, n * 8, hipMemcpyDeviceToHost, c->stream);
  hipError_t e3 = hipStreamSynchronize(c->stream);
  (void)hipFree(g);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(FR_EHIP, "kinetic-energy gradient");
  devToRef(c, h.data(), out, c->NF, c->NS);
  return 0;
}

#ifndef FR_BATCH_STREAMS
#define FR_BATCH_STREAMS 8  // graph branches of a batched step (twice the hardware queues a process has by default)
#endif

namespace {
std::atomic<int> gForceTile{-1};  // fr_batch_force_tile
}

int fr_batch_force_tile(int layout) {
  if (layout < -1 || layout > 1) return fail(FR_EINVAL, "tile layout is -1 (tuned), 0 or 1");
  gForceTile.store(layout);
  return 0;
}

int fr_batch_create(fr_ctx *const *ctxs, int n, fr_batch **out) {
  if (!ctxs || !out || n < 1) return fail(FR_EINVAL, "bad argument");
  *out = nullptr;
  fr_batch *b = new fr_batch();
  auto bad = [&](int code, const std::string &msg) {
    batchFree(b);
    delete b;
    return fail(code, msg);
  };
  const fr_ctx *c0 = ctxs[0];
  if (!c0) return bad(FR_EINVAL, "null context");
  b->nRK = c0->nRK;
  b->prec = c0->prec;
  for (int i = 0; i < n; i++) {
    fr_ctx *c = ctxs[i];
    if (!c) return bad(FR_EINVAL, "null context");
    if (c->cfg.equation != c0->cfg.equation || c->prec != b->prec || c->cfg.timeType != c0->cfg.timeType)
      return bad(FR_EINVAL, "batched meshes must share equation, precision and timeType");
    if (c->cfg.viscous) return bad(FR_EINVAL, "the batched kernel is inviscid");
    if (c->P < 1 || c->P > 5) return bad(FR_EINVAL, "batched meshes need P in 1..5");
    for (fr_ctx *o : b->ctx)
      if (o == c) return bad(FR_EINVAL, "a context appears twice in the batch");
    if (int rc0 = batchMemberCheck(c)) return bad(rc0, g_err);
    b->ctx.push_back(c);
  }
  // work items grouped by kernel (one launch per element type and order),
  // each group in mesh order, for both tile layouts
  for (int l = 0; l < 2; l++) {
    std::vector<frdev::BatchItem> items;
    for (int grp = 1; grp <= 15; grp++) {
      b->itemStart[l][grp] = (int)items.size();
      for (int i = 0; i < n; i++)
        if (batchGroup(b->ctx[i]) == grp) {
          b->groupNF[grp] = b->ctx[i]->NF;
          const int epb = batchEpb(b, b->ctx[i], l);
          const int nb = (b->ctx[i]->nE + epb - 1) / epb;
          for (int k = 0; k < nb; k++) items.push_back({i, k});
        }
      b->itemCount[l][grp] = (int)items.size() - b->itemStart[l][grp];
    }
    int rc;
    if ((rc = dmalloc(&b->items[l], items.size()))) return bad(rc, g_err);
    HIP_TRY(hipMemcpy(b->items[l], items.data(), items.size() * sizeof(frdev::BatchItem), hipMemcpyHostToDevice));
  }
  int rc = (b->prec == FR_PREC_FP32) ? batchArgs<float>(b) : batchArgs<double>(b);
  if (rc) return bad(rc, g_err);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bad(FR_EHIP, "stream");
  // The tuner: per group, the tile layout (threads per element, elements per
  // block) and the persistent grid (1..32 workgroups per CU) with the
  // shortest residual-only launch (it writes only divF[0]) of the stage
  // computation of the update, timed on the group alone.  The members'
  // traces are made complete first, so that every layout reads valid ones.
  for (fr_ctx *c : b->ctx) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->cur = 0;
    c->traceOf = -1;
    if ((rc = (c->prec == FR_PREC_FP32) ? ensureTraces<float>(c, 0) : ensureTraces<double>(c, 0)))
      return bad(rc, g_err);
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  int cu = 256;
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cu = prop.multiProcessorCount;
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return bad(FR_EHIP, "event");
  auto timeLaunch = [&](int grp, int lay, int grid, double &ms) -> int {
    int r;
    if ((r = batchLaunchGroup(b, -1, 0, b->stream, grp, lay, grid))) return r;  // warm-up
    (void)hipEventRecord(e0, b->stream);
    for (int k = 0; k < 3; k++)
      if ((r = batchLaunchGroup(b, -1, 0, b->stream, grp, lay, grid))) return r;
    (void)hipEventRecord(e1, b->stream);
    if (hipEventSynchronize(e1) != hipSuccess) return fail(FR_EHIP, "tuning run");
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    ms = t / 3.0;
    return 0;
  };
  b->nItems = 0;
  b->tunedMs = 0;
  for (int grp = 1; grp <= 15 && !rc; grp++) {
    if (!b->itemCount[0][grp]) continue;
    double best = 1e300;
    int gD, P;
    bool gTri;
    batchGroupKind(grp, gD, P, gTri);
    // layout 1 is a different tile only where it changes threads per element
    const bool l1 = gTri ? frdev::triTPEL(P, 1) != frdev::triTPEL(P, 0)
                         : frdev::tensorTPE(gD, P, 1) != frdev::tensorTPE(gD, P, 0);
    const int force = gForceTile.load();
    const int lay0 = force == 1 && l1 ? 1 : 0, lay1 = force == 0 || !l1 ? 0 : 1;
    for (int lay = lay0; lay <= lay1 && !rc; lay++)
      for (int mult : {1, 2, 4, 8, 16, 32}) {
        const int cnt = b->itemCount[lay][grp], grid = std::max(1, std::min(cu * mult, cnt));
        double ms = 0;
        if ((rc = timeLaunch(grp, lay, grid, ms))) break;
        if (ms < best) {
          best = ms;
          b->lay[grp] = lay;
          b->grid[grp] = grid;
        }
        if (grid == cnt) break;
      }
    b->groupMs[grp] = best;
    b->tunedMs += best;
    b->nItems += b->itemCount[b->lay[grp]][grp];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return bad(rc, g_err);
  // One step's launches, captured per trace parity.  Meshes of different
  // element types and orders are independent, so the groups' stages form
  // FR_BATCH_STREAMS branches of the graph (forked from and joined to the
  // batch stream) that run concurrently (the runtime spreads the branches'
  // streams over its hardware queues).  The groups go to the branches longest
  // first, each to the branch with the least tuned time so far, so that the
  // branches end together; a branch runs its groups one after the other.
  std::vector<int> order;
  for (int grp = 1; grp <= 15; grp++)
    if (b->itemCount[0][grp]) order.push_back(grp);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->groupMs[x] > b->groupMs[y]; });
  const int nBr = std::max(1, std::min<int>(FR_BATCH_STREAMS, (int)order.size()));
  std::vector<std::vector<int>> branch(nBr);
  std::vector<double> load(nBr, 0.0);
  for (int grp : order) {
    const int br = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    branch[br].push_back(grp);
    load[br] += b->groupMs[grp];
  }
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) return bad(FR_EHIP, "event");
  for (int br = 0; br < nBr; br++)
    if (hipStreamCreateWithFlags(&b->ps[br], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->join[br], hipEventDisableTiming) != hipSuccess)
      return bad(FR_EHIP, "stream");
  for (int p = 0; p < ((b->nRK & 1) ? 2 : 1); p++) {
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeGlobal) != hipSuccess) return bad(FR_EHIP, "capture");
    bool ok = hipEventRecord(b->fork, b->stream) == hipSuccess;
    for (int br = 0; br < nBr && ok && !rc; br++) {
      ok = hipStreamWaitEvent(b->ps[br], b->fork, 0) == hipSuccess;
      for (int grp : branch[br])
        for (int s = 0; s < b->nRK && ok && !rc; s++)
          rc = batchLaunchGroup(b, p, s, b->ps[br], grp, b->lay[grp], b->grid[grp]);
      ok = ok && hipEventRecord(b->join[br], b->ps[br]) == hipSuccess &&
           hipStreamWaitEvent(b->stream, b->join[br], 0) == hipSuccess;
    }
    if (!ok || rc) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(b->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      return bad(rc ? rc : FR_EHIP, rc ? g_err : std::string("graph capture"));
    }
    if (hipStreamEndCapture(b->stream, &b->graph[p]) != hipSuccess) return bad(FR_EHIP, "end capture");
    if (hipGraphInstantiate(&b->exec[p], b->graph[p], nullptr, nullptr, 0) != hipSuccess)
      return bad(FR_EHIP, "graph instantiate");
  }
  // the tuning launches wrote only divF[0]; the states and traces are as before
  // (the wave speeds of CFL members are rewritten by the next step's stages)
  for (fr_ctx *c : b->ctx) c->batches.push_back(b);
  b->alive.assign(b->ctx.size(), true);
  *out = b;
  return 0;
}

int fr_batch_update(fr_batch *b, int nSteps) {
  if (!b) return fail(FR_EINVAL, "null argument");
  if (b->broken) return fail(FR_EINVAL, "a member context of this batch was destroyed");
  int rc;
  bool stale = false;
  for (size_t i = 0; i < b->ctx.size(); i++) {
    if ((rc = batchMemberCheck(b->ctx[i]))) return rc;
    stale = stale || b->ctx[i]->gen != b->gen[i];
  }
  // a member's time step, geometry layout or FP32 energy offset changed since
  // the arguments were built: rebuild them in place
  if (stale) {
    for (fr_ctx *c : b->ctx) HIP_TRY(hipStreamSynchronize(c->stream));
    rc = (b->prec == FR_PREC_FP32) ? batchArgs<float>(b) : batchArgs<double>(b);
    if (rc) return rc;
  }
  // an even number of stages returns the traces to the buffer they started
  // in, so one graph (parity 0) serves every step; odd counts alternate
  int parity = (b->nRK & 1) ? (b->ctx[0]->cur & 1) : 0;
  if ((rc = batchPrepare(b, parity))) return rc;
  for (int i = 0; i < nSteps; i++) {
    // members with a CFL time step form it first, on the batch stream
    for (fr_ctx *c : b->ctx)
      if (c->dtType != 0 && (rc = dtPrepare(c, b->stream))) return rc;
    HIP_TRY(hipGraphLaunch(b->exec[parity], b->stream));
    for (fr_ctx *c : b->ctx)
      if (c->dtType != 0) c->dtFirst = c->dtReady = false;
    parity ^= (b->nRK & 1);
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (fr_ctx *c : b->ctx) {
    c->cur = parity;
    c->traceOf = 0;
    c->traceEpb = c->tri ? 0 : batchEpb(b, c, b->lay[batchGroup(c)]);
    c->finalDropped = !c->keepFinal && c->nRK > 1;
  }
  return 0;
}

int fr_batch_tuning(fr_batch *b, int *grid, int *nItems, double *msStage0) {
  if (!b) return fail(FR_EINVAL, "null argument");
  if (grid) *grid = *std::max_element(b->grid, b->grid + 16);
  if (nItems) *nItems = b->nItems;
  if (msStage0) *msStage0 = b->tunedMs;
  return 0;
}

int fr_batch_group_tuning(fr_batch *b, int group, int *threadsPerElement, int *elementsPerBlock, int *grid,
                          int *nItems, double *ms) {
  if (!b) return fail(FR_EINVAL, "null argument");
  if (group < 1 || group > 15)
    return fail(FR_EINVAL, "group is an order 1..5, plus 5 for triangles, plus 10 for hexes");
  int gD, P;
  bool gTri;
  batchGroupKind(group, gD, P, gTri);
  const int lay = b->lay[group];
  const int tpe = gTri ? frdev::triTPEL(P, lay) : frdev::tensorTPE(gD, P, lay);
  if (threadsPerElement) *threadsPerElement = tpe;
  if (elementsPerBlock) *elementsPerBlock = tpe >= 256 ? 1 : 256 / tpe;
  if (grid) *grid = b->itemCount[lay][group] ? b->grid[group] : 0;
  if (nItems) *nItems = b->itemCount[lay][group];
  if (ms) *ms = b->itemCount[lay][group] ? b->groupMs[group] : 0.0;
  return 0;
}

int fr_batch_destroy(fr_batch *b) {
  if (!b) return 0;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batchFree(b);
  delete b;
  return 0;
}

int fr_set_stream(fr_ctx *c, void *stream) {
  if (!c) return fail(FR_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->stream = stream ? (hipStream_t)stream : c->own;
  return 0;
}

int fr_synchronize(fr_ctx *c) {
  if (!c) return fail(FR_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int fr_geometry_mode(fr_ctx *c, int request, int *active) {
  if (!c || request < -1 || request > 3) return fail(FR_EINVAL, "bad argument");
  if (request >= 0 && c->geoReq != request) {
    c->geoReq = request;
    c->gen++;
  }
  if (active) *active = c->geoActive();
  return 0;
}

int fr_keep_final_residual(fr_ctx *c, int keep) {
  if (!c || keep < 0 || keep > 1) return fail(FR_EINVAL, "bad argument");
  if (c->keepFinal != keep) {
    c->keepFinal = keep;
    c->gen++;
  }
  return 0;
}

int fr_set_timing(fr_ctx *c, int on) {
  if (!c) return fail(FR_EINVAL, "null argument");
  c->timing = on != 0;
  return 0;
}

int fr_kernel_times(fr_ctx *c, int kernel, double *ms, long *count) {
  if (!c || kernel < 0 || kernel > 3) return fail(FR_EINVAL, "bad argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 4; k++) {
    for (auto &pr : c->ev[k]) {
      float t = 0;
      HIP_TRY(hipEventElapsedTime(&t, pr.first, pr.second));
      c->ms[k] += t;
      c->cnt[k] += 1;
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    c->ev[k].clear();
  }
  if (ms) *ms = c->ms[kernel];
  if (count) *count = c->cnt[kernel];
  c->ms[kernel] = 0;
  c->cnt[kernel] = 0;
  return 0;
}

int fr_residual_bytes(fr_ctx *c, double *bytes) {
  if (!c || !bytes) return fail(FR_EINVAL, "null argument");
  const double s = (double)c->tsize(), F = (double)c->conn.faces.size();
  *bytes = (double)c->nE * c->NS * (2.0 * c->NF * s + (double)c->D * c->D * s) +
           F * c->NFPF * (c->D + 1) * s + 20.0 * F;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// fr_reduce_noalloc: fr_reduce with its block partial sums in `scratch`, a
// fixed device array, instead of the context's buffer (which fr_reduce
// allocates on first use).  `scratch` is a host-side stand-in in host code:
// the device address of each GPU's copy comes from hipGetSymbolAddress,
// taken once per device with that device current, and only that address is
// passed to kernels.  Every call on a GPU shares that GPU's copy; calls that
// overlap on one GPU (on different streams or threads) race.  Each call
// writes every partial it reads (one entry per block and field), so nothing
// needs clearing.
// ---------------------------------------------------------------------------
namespace {

__device__ __align__(256) double scratch[FR_NOALLOC_SCRATCH_DOUBLES];

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
double *gScratchPtr[kMaxDevices] = {};

// Device address of device dev's copy of `scratch` (dev is current).
int scratchAddress(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(FR_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (!gScratchPtr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FR_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    gScratchPtr[dev] = (double *)p;
  }
  *out = gScratchPtr[dev];
  return 0;
}

// Makes device `dev` current for the lifetime of the guard.
struct DeviceGuard {
  int prev = -1;
  hipError_t err = hipSuccess;
  explicit DeviceGuard(int dev) {
    int cur = -1;
    err = hipGetDevice(&cur);
    if (err == hipSuccess && cur != dev) {
      err = hipSetDevice(dev);
      if (err == hipSuccess) prev = cur;
    }
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// out[k] = the block partial sums part[b * stride + k], b = 0 .. nb-1, added
// (or, max != 0, combined by the NaN-keeping maximum) in block order, as the
// host finishes them in fr_reduce: one thread per k.
__global__ void k_finish_partials(const double *__restrict__ part, int nb, int stride, int nk, int max,
                                  double *__restrict__ out) {
  const int k = threadIdx.x;
  if (k >= nk) return;
  double r = 0.0;
  for (int b = 0; b < nb; b++) {
    const double v = part[(size_t)b * stride + k];
    r = max ? frdev::maxKeepNaN(r, v) : r + v;
  }
  out[k] = r;
}

}  // namespace

extern "C" {

int fr_reduce_noalloc(fr_ctx *c, int kind, double *out) {
  if (!c || !out) return fail(FR_EINVAL, "null argument");
  if (!(kind >= 1 && kind <= 3) && kind != 10) return fail(FR_EINVAL, "unknown reduction kind");
  if (kind == 10 && c->cfg.equation != 1)
    return fail(FR_EINVAL, "kinetic energy needs the Euler / Navier-Stokes equations");
  const int nb = (int)nblk(c->nE);
  const int stride = kind == 10 ? 2 : 5, nk = kind == 10 ? 2 : c->NF;
  // the block partials, then the finished values
  const size_t need = (size_t)nb * stride + 8;
  if (need > (size_t)FR_NOALLOC_SCRATCH_DOUBLES)
    return fail(FR_EINVAL, "fr_reduce_noalloc: the mesh needs " + std::to_string(need) +
                               " doubles of scratch, more than FR_NOALLOC_SCRATCH_DOUBLES (" +
                               std::to_string(FR_NOALLOC_SCRATCH_DOUBLES) + "): at most " +
                               std::to_string((FR_NOALLOC_SCRATCH_DOUBLES - 8) / 5 * 256) + " elements");
  DeviceGuard dg(c->dev);
  if (dg.err != hipSuccess) return fail(FR_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(dg.err));
  double *sp = nullptr;
  int rc = scratchAddress(c->dev, &sp);
  if (rc) return rc;
  double *part = sp, *fin = sp + (size_t)nb * stride;
  if (kind == 10) {
    if (c->D == 3)
      hipLaunchKernelGGL(frdev::k_ke<3>, dim3(nb), dim3(256), 0, c->stream, c->U, c->detJ, c->wts, c->NS, c->nE,
                         part);
    else
      hipLaunchKernelGGL(frdev::k_ke<2>, dim3(nb), dim3(256), 0, c->stream, c->U, c->detJ, c->wts, c->NS, c->nE,
                         part);
  } else if (c->prec == FR_PREC_FP32) {
    hipLaunchKernelGGL(frdev::k_norm<float>, dim3(nb), dim3(256), 0, c->stream, (const float *)c->divF[0], c->detJ,
                       c->wts, kind, c->NF, c->NS, c->nE, part);
  } else {
    hipLaunchKernelGGL(frdev::k_norm<double>, dim3(nb), dim3(256), 0, c->stream, (const double *)c->divF[0],
                       c->detJ, c->wts, kind, c->NF, c->NS, c->nE, part);
  }
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(64), 0, c->stream, part, nb, stride, nk, kind == 3 ? 1 : 0,
                     fin);
  HIP_TRY(hipGetLastError());
  double h[8];
  HIP_TRY(hipMemcpyAsync(h, fin, (size_t)nk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (kind == 10) {
    out[0] = h[0];
    out[1] = h[1];
    return 0;
  }
  for (int k = 0; k < c->NF; k++) {
    double r = h[k];
    if (kind == 2) r = std::sqrt(r);
    if (std::isnan(r)) return fail(FR_ENAN, "NaN Encountered in Solution Residual!");
    out[k] = r;
  }
  return 0;
}

}  // extern "C"
