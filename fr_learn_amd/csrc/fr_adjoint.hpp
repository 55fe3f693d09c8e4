// fr_adjoint.hpp -- hand-written adjoint (vector-Jacobian product) of the
// inviscid FR residual, FP64.
//
// For a cotangent lam of divF (layout [ele][field][spt]) the kernels return
// lam^T d divF / dU in the state's layout.  They transpose the residual
// kernel phase by phase, with the same index maps:
//   correction  divF[t] += sum_f corr(t, q_f) J[q_f]  ->  lamJ[q] = sum over
//               the solution points on q's line of corr * lam  (a g-line
//               contraction, like an extrapolation with g_L / g_R weights);
//   face flux   J[q] = Fn(U_L, U_R, n_L) * scale - disFn[q]: the cotangent of
//               the common flux at a face point is g = lamJ * scale of this
//               element plus that of the neighbour's paired point (both
//               elements evaluate the same Fn); the derivative with respect
//               to this element's trace comes from evaluating the common
//               flux (and a boundary's ghost state) in Dual<NF>;
//   volume      divF[t] = sum_d Dm F~_d and -disFn = -extrapolation(F~_d):
//               transposed along the lines, then F~ = JGinv F transposed,
//               then the flux Jacobian from inviscidFlux in Dual<NF>;
//   traces      u_o[q] = extrapolation of U: transposed with corrEntries.
// Element-parallel and gather-only: a first kernel writes every element's
// scaled face cotangents (the adjoint of the trace exchange), the second
// reads its neighbours' ones, so no atomics are needed.
#pragma once

#include <hip/hip_runtime.h>

#include "fr_dual.hpp"
#include "fr_kernels.hpp"

namespace frdev {

// lamJ at flux point q: sum_m g_side[m] * lam[line_q(m)]
template <int D, int P, int NF>
__device__ __forceinline__ void lamJAt(const double *s1, const double *lamE, int q, double *out) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  constexpr int NS = G::NS, P1 = G::P1;
  int base, stride, side, dn;
  fptLine<D, P>(q, base, stride, side, dn);
  const double *g = s1 + (side ? O::GR : O::GL);
#pragma unroll
  for (int k = 0; k < NF; k++) {
    double acc = 0;
#pragma unroll
    for (int m = 0; m < P1; m++) acc += g[m] * lamE[k * NS + base + m * stride];
    out[k] = acc;
  }
}

// (a) scaled face cotangents A[ele][field][fpt] = lamJ * (+dA left / boundary,
// -dA right), one thread per (element, flux point)
template <int D, int P, int NF>
__global__ __launch_bounds__(256) void k_vjp_faces(const double *__restrict__ lam, const double *__restrict__ fgeo,
                                                   const int8_t *__restrict__ finfo,
                                                   const double *__restrict__ oned, int nEles,
                                                   double *__restrict__ A) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, NFACE = G::NFACE;
  __shared__ double s1[O::SIZE];
  for (int i = threadIdx.x; i < O::SIZE; i += blockDim.x) s1[i] = oned[i];
  __syncthreads();
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nEles * NFP) return;
  const int e = (int)(idx / NFP), q = (int)(idx - (long)e * NFP);
  double lj[NF];
  lamJAt<D, P, NF>(s1, lam + (size_t)e * NF * NS, q, lj);
  const int fi = finfo[(size_t)e * NFACE + q / NFPF];
  const double dA = fgeo[((size_t)e * (D + 1) + D) * NFP + q];
  const bool right = fi < FINFO_BOUNDARY && ((fi >> 5) & 1);
  const double scale = right ? -dA : dA;
#pragma unroll
  for (int k = 0; k < NF; k++) A[((size_t)e * NF + k) * NFP + q] = lj[k] * scale;
}

// (b) the state cotangent, one TPE-thread group per element
template <int D, int P, int NF, bool BND>
__global__ __launch_bounds__((ElemGeo<D, P>::NTH)) void k_vjp_state(
    const double *__restrict__ U, const double *__restrict__ lam, const double *__restrict__ Uf,
    const double *__restrict__ A, const double *__restrict__ JG, const double *__restrict__ fgeo,
    const int32_t *__restrict__ fnbr, const int8_t *__restrict__ finfo, const double *__restrict__ oned, Phys ph,
    int nEles, double *__restrict__ out) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  using DN = Dual<NF>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, P1 = G::P1, TPE = G::TPE, EPB = G::EPB;
  constexpr int NFACE = G::NFACE;
  __shared__ double sU[EPB][NF][NS];
  __shared__ double sL[EPB][NF][NS];
  __shared__ double sJ[EPB][NF][NFP];   // lamJ
  __shared__ double sUo[EPB][NF][NFP];  // cotangent of this element's traces
  __shared__ double s1[O::SIZE];
  const int slot = threadIdx.x / TPE, t = threadIdx.x - slot * TPE;
  const int e = blockIdx.x * EPB + slot;
  const bool act = e < nEles;
  for (int i = threadIdx.x; i < O::SIZE; i += blockDim.x) s1[i] = oned[i];
  if (act)
    for (int i = t; i < NF * NS; i += TPE) {
      (&sU[slot][0][0])[i] = U[(size_t)e * NF * NS + i];
      (&sL[slot][0][0])[i] = lam[(size_t)e * NF * NS + i];
    }
  __syncthreads();
  // face points: lamJ, the flux cotangent g and its pull-back to this trace
  if (act) {
    for (int q = t; q < NFP; q += TPE) {
      double lj[NF];
      lamJAt<D, P, NF>(s1, &sL[slot][0][0], q, lj);
      int base, stride, side, dn;
      fptLine<D, P>(q, base, stride, side, dn);
      const double *l1 = s1 + (side ? O::LR : O::LL);
      DN uo[NF], un[NF], n[D], Fn[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        double acc = 0;
#pragma unroll
        for (int m = 0; m < P1; m++) acc += l1[m] * sU[slot][k][base + m * stride];
        uo[k] = DN::var(acc, k);
        sJ[slot][k][q] = lj[k];
      }
#pragma unroll
      for (int c = 0; c < D; c++) n[c] = DN(fgeo[((size_t)e * (D + 1) + c) * NFP + q]);
      const double dA = fgeo[((size_t)e * (D + 1) + D) * NFP + q];
      const int fi = finfo[(size_t)e * NFACE + q / NFPF];
      const bool bnd = BND && fi >= FINFO_BOUNDARY;
      double g[NF];
      if (bnd) {
        boundaryState<D, NF, DN>(ph, fi - FINFO_BOUNDARY, uo, n, un);
        commonFlux<D, NF, DN>(ph, true, uo, un, n, Fn);
#pragma unroll
        for (int k = 0; k < NF; k++) g[k] = lj[k] * dA;
      } else {
        const int en = fnbr[(size_t)e * NFACE + q / NFPF];
        const int qn = neighbourFpt<D, P>(q, fi);
        const bool right = (fi >> 5) & 1;
#pragma unroll
        for (int k = 0; k < NF; k++) {
          un[k] = DN(Uf[((size_t)en * NF + k) * NFP + qn]);
          g[k] = lj[k] * (right ? -dA : dA) + A[((size_t)en * NF + k) * NFP + qn];
        }
        // the left element's state first, chosen per lane (one common flux
        // per wavefront instead of one per side)
        DN sL[NF], sR[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
          sL[k] = right ? un[k] : uo[k];
          sR[k] = right ? uo[k] : un[k];
        }
        commonFlux<D, NF, DN>(ph, false, sL, sR, n, Fn);
      }
#pragma unroll
      for (int j = 0; j < NF; j++) {
        double acc = 0;
#pragma unroll
        for (int k = 0; k < NF; k++) acc += g[k] * Fn[k].d[j];
        sUo[slot][j][q] = acc;
      }
    }
  }
  __syncthreads();
  if (!act || t >= NS) return;
  int si, sj, sk;
  if (D == 3) {
    sk = t / (P1 * P1);
    sj = (t - P1 * P1 * sk) / P1;
    si = t - P1 * sj - P1 * P1 * sk;
  } else {
    si = t % P1;
    sj = t / P1;
    sk = 0;
  }
  int cq[NFACE], cside[NFACE], cax[NFACE];
  corrEntries<D, P>(t, cq, cside, cax);
  // cotangent of the contravariant fluxes F~_d at this point
  double lFt[D][NF];
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int id = (d == 0) ? si : (d == 1) ? sj : sk;
#pragma unroll
    for (int k = 0; k < NF; k++) {
      double acc = 0;
#pragma unroll
      for (int m = 0; m < P1; m++) {
        const int sp = (d == 0) ? m + P1 * (sj + P1 * sk) : (d == 1) ? si + P1 * (m + P1 * sk) : si + P1 * (sj + P1 * m);
        acc += s1[O::DM + m * P1 + id] * sL[slot][k][sp];
      }
#pragma unroll
      for (int f = 0; f < NFACE; f++)
        if (faceAxis<D>(f) == d)
          acc -= (cside[f] ? 1.0 : -1.0) * s1[(cside[f] ? O::LR : O::LL) + cax[f]] * sJ[slot][k][cq[f]];
      lFt[d][k] = acc;
    }
  }
  // physical fluxes: F~_d = sum_d2 JG(d, d2) F_d2
  double lF[D][NF];
#pragma unroll
  for (int d2 = 0; d2 < D; d2++)
#pragma unroll
    for (int k = 0; k < NF; k++) {
      double acc = 0;
#pragma unroll
      for (int d = 0; d < D; d++) acc += JG[((size_t)e * D * D + d * D + d2) * NS + t] * lFt[d][k];
      lF[d2][k] = acc;
    }
  DN u[NF], F[D][NF];
#pragma unroll
  for (int k = 0; k < NF; k++) u[k] = DN::var(sU[slot][k][t], k);
  inviscidFlux<D, NF, DN>(ph, u, F);
#pragma unroll
  for (int j = 0; j < NF; j++) {
    double acc = 0;
#pragma unroll
    for (int d2 = 0; d2 < D; d2++)
#pragma unroll
      for (int k = 0; k < NF; k++) acc += lF[d2][k] * F[d2][k].d[j];
    // trace extrapolation, transposed
#pragma unroll
    for (int f = 0; f < NFACE; f++) acc += s1[(cside[f] ? O::LR : O::LL) + cax[f]] * sUo[slot][j][cq[f]];
    out[((size_t)e * NF + j) * NS + t] = acc;
  }
}

// ---------------------------------------------------------------------------
// Viscous (LDG) residual.  Besides the inviscid chain, divF depends on U
// through the corrected gradient Q = JGinv^T (Dm U + lift(u^ - u)) / detJ at
// the solution points (point viscous flux) and its traces Qf = E Q (face
// viscous flux), and u^ couples neighbours.  Two more exchanges follow the
// face-cotangent one: k_vjp_face_vis pulls the common fluxes back to each flux
// point's traces (one thread per point), k_vjp_state_vis (one element per
// workgroup of TPE threads) the point fluxes and the gradient back to the
// state and writes the cotangent of each flux point's jump u^ - u (B), and
// k_vjp_ldg gathers the neighbour's one to pull the common solution u^ back
// to this element's traces.
// ---------------------------------------------------------------------------
// Face part of the viscous state VJP, one thread per (element, flux point):
// the common flux (inviscid + LDG viscous) pulled back to this element's
// state trace and gradient traces, Uo[ele][field][fpt] and
// Qo[ele][dim][field][fpt], contracted with the point's face cotangent
// (lamJ * dA plus the paired point's A); and the forward LDG jump u^ - u,
// Cj[ele][field][fpt].  Uf are the state traces, Qf the gradient traces.
template <int D, int P, int NF, bool BND>
__global__ __launch_bounds__(256) void k_vjp_face_vis(
    const double *__restrict__ lam, const double *__restrict__ Uf, const double *__restrict__ Qf,
    const double *__restrict__ A, const double *__restrict__ fgeo, const int32_t *__restrict__ fnbr,
    const int8_t *__restrict__ finfo, const double *__restrict__ oned, Phys ph, int nEles,
    double *__restrict__ Uo, double *__restrict__ Qo, double *__restrict__ Cj) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, P1 = G::P1, NFACE = G::NFACE;
  using DM = Dual<1>;  // one direction per pass (see k_vjp_state_vis)
  __shared__ double s1[O::SIZE];
  for (int i = threadIdx.x; i < O::SIZE; i += blockDim.x) s1[i] = oned[i];
  __syncthreads();
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nEles * NFP) return;
  const int e = (int)(idx / NFP), q = (int)(idx - (long)e * NFP);
  {
    double lj[NF];
    {
      // lamJ at this point from the element's cotangent along the point's line
      int base, stride, side, dn;
      fptLine<D, P>(q, base, stride, side, dn);
      const double *g = s1 + (side ? O::GR : O::GL);
      const double *le = lam + (size_t)e * NF * NS;
#pragma unroll
      for (int k = 0; k < NF; k++) {
        double acc = 0;
#pragma unroll
        for (int m = 0; m < P1; m++) acc += g[m] * le[k * NS + base + m * stride];
        lj[k] = acc;
      }
    }
    double uov[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) uov[k] = Uf[((size_t)e * NF + k) * NFP + q];
    double nv[D];
#pragma unroll
    for (int c = 0; c < D; c++) nv[c] = fgeo[((size_t)e * (D + 1) + c) * NFP + q];
    const double dA = fgeo[((size_t)e * (D + 1) + D) * NFP + q];
    const int fi = finfo[(size_t)e * NFACE + q / NFPF];
    const bool bnd = BND && fi >= FINFO_BOUNDARY;
    const bool right = !bnd && ((fi >> 5) & 1);
    const int en = bnd ? 0 : fnbr[(size_t)e * NFACE + q / NFPF];
    const int qn = bnd ? 0 : neighbourFpt<D, P>(q, fi);
    double g[NF], unv[NF];
    // state directions: dual passes through the inviscid and viscous fluxes
#pragma unroll 1
    for (int sj = 0; sj < NF; sj++) {
      DM uo[NF], un[NF], n[D], Fn[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) uo[k] = (k == sj) ? DM::var(uov[k], 0) : DM(uov[k]);
#pragma unroll
      for (int c = 0; c < D; c++) n[c] = DM(nv[c]);
      // the gradient traces do not depend on this pass's direction: plain
      // doubles (half the registers of constant duals), except at boundaries,
      // whose ghost gradient may depend on the state
      double gov[D][NF];
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int k = 0; k < NF; k++) gov[d][k] = Qf[(((size_t)e * D + d) * NF + k) * NFP + q];
      if (bnd) {
        const int bc = fi - FINFO_BOUNDARY;
        DM go[D][NF], gn[D][NF];
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int k = 0; k < NF; k++) go[d][k] = DM(gov[d][k]);
        boundaryState<D, NF, DM>(ph, bc, uo, n, un);
        if constexpr (NF > 1) {
          boundaryGradient<D, NF, DM>(ph, bc, uo, un, go, gn);
        } else {
#pragma unroll
          for (int d = 0; d < D; d++) gn[d][0] = DM(0.0);
        }
        commonFlux<D, NF, DM>(ph, true, uo, un, n, Fn);
        ldgViscousFlux<D, NF, DM>(ph, (bc == 10) ? 2 : 1, uo, un, go, gn, n, Fn);
#pragma unroll
        for (int k = 0; k < NF; k++) {
          g[k] = lj[k] * dA;
          unv[k] = un[k].v;
        }
      } else {
#pragma unroll
        for (int k = 0; k < NF; k++) {
          unv[k] = Uf[((size_t)en * NF + k) * NFP + qn];
          un[k] = DM(unv[k]);
          g[k] = lj[k] * (right ? -dA : dA) + A[((size_t)en * NF + k) * NFP + qn];
        }
        double gnv[D][NF];
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int k = 0; k < NF; k++) gnv[d][k] = Qf[(((size_t)en * D + d) * NF + k) * NFP + qn];
        if (right) {
          commonFlux<D, NF, DM>(ph, false, un, uo, n, Fn);
          ldgViscousFlux<D, NF, DM, double>(ph, 0, un, uo, gnv, gov, n, Fn);
        } else {
          commonFlux<D, NF, DM>(ph, false, uo, un, n, Fn);
          ldgViscousFlux<D, NF, DM, double>(ph, 0, uo, un, gov, gnv, n, Fn);
        }
      }
      double acc = 0;
#pragma unroll
      for (int k = 0; k < NF; k++) acc += g[k] * Fn[k].d[0];
      Uo[((size_t)e * NF + sj) * NFP + q] = acc;
    }
    // gradient directions: the LDG viscous flux is affine in this point's
    // gradient trace (viscous fluxes are linear in the gradient, and the
    // boundary gradients are linear maps of it), so its derivative along
    // e_dk is the flux at go = e_dk less the flux at go = 0, with the
    // neighbour's gradient held at zero; the inviscid flux does not depend
    // on it
    {
      double nd[D], uo[NF], un[NF];
#pragma unroll
      for (int c = 0; c < D; c++) nd[c] = nv[c];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        uo[k] = uov[k];
        un[k] = unv[k];
      }
      const int bc = bnd ? fi - FINFO_BOUNDARY : 0;
      auto viscF = [&](const double (&go)[D][NF], double *F) {
        double gn[D][NF];
#pragma unroll
        for (int k = 0; k < NF; k++) F[k] = 0.0;
        if (bnd) {
          if constexpr (NF > 1) {
            boundaryGradient<D, NF, double>(ph, bc, uo, un, go, gn);
          } else {
#pragma unroll
            for (int d = 0; d < D; d++) gn[d][0] = 0.0;
          }
          ldgViscousFlux<D, NF, double>(ph, (bc == 10) ? 2 : 1, uo, un, go, gn, nd, F);
        } else {
#pragma unroll
          for (int d = 0; d < D; d++)
#pragma unroll
            for (int k = 0; k < NF; k++) gn[d][k] = 0.0;
          if (right) ldgViscousFlux<D, NF, double>(ph, 0, un, uo, gn, go, nd, F);
          else ldgViscousFlux<D, NF, double>(ph, 0, uo, un, go, gn, nd, F);
        }
      };
      if (!bnd) {
        // interior: Fn depends on this point's gradient only through this
        // side's viscous flux F = viscousFlux(uo, go), linearly, with
        // dFn_k / dF[d][k] = w[d] the same for every field: 1/2 n_d from the
        // average, plus (left side) pen |n|^2 times the coefficient of F_d
        // in n.(F_L - F_R) -- with face.cpp:245's substitution of F_L,y for
        // F_L,z when viscFix = 0 -- or (right side) minus pen |n|^2 n_d.
        // viscousFlux is linear and homogeneous in the gradient, so the
        // derivative along e_dk is the flux of the unit gradient itself.
        double w[D];
        if constexpr (NF == 1) {
#pragma unroll
          for (int d = 0; d < D; d++) w[d] = 0.5 * nd[d];
        } else {
          const double pen = ldgPenalty<D, double>(ph, nd);
          double nn = 0;
#pragma unroll
          for (int d = 0; d < D; d++) nn += nd[d] * nd[d];
          const double pn = pen * nn;
#pragma unroll
          for (int d = 0; d < D; d++) w[d] = 0.5 * nd[d] + (right ? -pn * nd[d] : pn * nd[d]);
          if constexpr (D == 3)
            if (!right && !ph.viscFix) {  // F_L,y stands in for F_L,z in n.(F_L - F_R)
              w[1] += pn * nd[2];
              w[2] -= pn * nd[2];
            }
        }
#pragma unroll
        for (int dk = 0; dk < D * NF; dk++) {
          const int d = dk / NF, sj = dk - d * NF;
          double ge[D][NF], F[D][NF];
#pragma unroll
          for (int d2 = 0; d2 < D; d2++)
#pragma unroll
            for (int k = 0; k < NF; k++) ge[d2][k] = (d2 == d && k == sj) ? 1.0 : 0.0;
          viscousFlux<D, NF, double>(ph, uo, ge, F);
          double acc = 0;
#pragma unroll
          for (int k = 0; k < NF; k++) {
            double fk = 0;
#pragma unroll
            for (int d2 = 0; d2 < D; d2++) fk += w[d2] * F[d2][k];
            acc += g[k] * fk;
          }
          Qo[(((size_t)e * D + d) * NF + sj) * NFP + q] = acc;
        }
      } else {
        double go[D][NF], F0[NF], F1[NF];
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int k = 0; k < NF; k++) go[d][k] = 0.0;
        viscF(go, F0);
#pragma unroll 1
        for (int dk = 0; dk < D * NF; dk++) {
          const int d = dk / NF, sj = dk - d * NF;
#pragma unroll
          for (int d2 = 0; d2 < D; d2++)
#pragma unroll
            for (int k = 0; k < NF; k++) go[d2][k] = (d2 == d && k == sj) ? 1.0 : 0.0;
          viscF(go, F1);
          double acc = 0;
#pragma unroll
          for (int k = 0; k < NF; k++) acc += g[k] * (F1[k] - F0[k]);
          Qo[(((size_t)e * D + d) * NF + sj) * NFP + q] = acc;
        }
      }
    }
    // forward LDG jump of this point (as k_grad forms it)
    double jmp[NF];
    ldgJump<D, NF, double>(ph, bnd, right, uov, unv, nv, jmp);
#pragma unroll
    for (int k = 0; k < NF; k++) Cj[((size_t)e * NF + k) * NFP + q] = jmp[k];
  }
}

template <int D, int P, int NF, bool BND>
__global__ __launch_bounds__((ElemGeo<D, P>::TPE)) void k_vjp_state_vis(
    const double *__restrict__ U, const double *__restrict__ lam, const double *__restrict__ Uo,
    const double *__restrict__ Qo, const double *__restrict__ Cj, const double *__restrict__ JG,
    const double *__restrict__ detJ, const int8_t *__restrict__ finfo, const double *__restrict__ oned, Phys ph,
    int nEles, double *__restrict__ out, double *__restrict__ B) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, P1 = G::P1, TPE = G::TPE, NFACE = G::NFACE;
  // Derivatives one direction at a time: pass p = grp * NF + j seeds field j
  // of the state (grp 0) or of the gradient's dimension grp - 1.  Every value
  // of a pass, seeded or not, carries as many partials as the dual type has,
  // so wider duals (NF or NF (D + 1) directions per pass) hold several times
  // the registers for the many unseeded values and spill heavily.
  using DM = Dual<1>;
  __shared__ double sU[NF][NS], sL[NF][NS];
  __shared__ double sJ[NF][NFP];       // lamJ
  __shared__ double sC[NF][NFP];       // forward LDG jumps u^ - u
  __shared__ double sUo[NF][NFP];      // cotangent of this element's state traces
  __shared__ double sQo[D][NF][NFP];   // cotangent of this element's gradient traces
  __shared__ double sLQ[D][NF][NS];    // cotangent of the reference gradient
  __shared__ double s1[O::SIZE];
  const int e = blockIdx.x, t = threadIdx.x;
  for (int i = t; i < O::SIZE; i += TPE) s1[i] = oned[i];
  for (int i = t; i < NF * NS; i += TPE) {
    (&sU[0][0])[i] = U[(size_t)e * NF * NS + i];
    (&sL[0][0])[i] = lam[(size_t)e * NF * NS + i];
  }
  __syncthreads();
  for (int q = t; q < NFP; q += TPE) {
    double lj[NF];
    lamJAt<D, P, NF>(s1, &sL[0][0], q, lj);
#pragma unroll
    for (int k = 0; k < NF; k++) {
      sJ[k][q] = lj[k];
      sUo[k][q] = Uo[((size_t)e * NF + k) * NFP + q];
      sC[k][q] = Cj[((size_t)e * NF + k) * NFP + q];
#pragma unroll
      for (int d = 0; d < D; d++) sQo[d][k][q] = Qo[(((size_t)e * D + d) * NF + k) * NFP + q];
    }
  }
  __syncthreads();
  double lU[NF];
  if (t < NS) {
    int si, sj, sk;
    if (D == 3) {
      sk = t / (P1 * P1);
      sj = (t - P1 * P1 * sk) / P1;
      si = t - P1 * sj - P1 * P1 * sk;
    } else {
      si = t % P1;
      sj = t / P1;
      sk = 0;
    }
    int cq[NFACE], cside[NFACE], cax[NFACE];
    corrEntries<D, P>(t, cq, cside, cax);
    double pJ[D * D];
#pragma unroll
    for (int c = 0; c < D * D; c++) pJ[c] = JG[((size_t)e * D * D + c) * NS + t];
    const double pdet = detJ[(size_t)e * NS + t];
    double Q[D][NF];
    // hex P3: the derivatives on the matrix cores (the gradient's, and the
    // transposed one of the flux cotangent below)
    constexpr bool MV = FR_MFMA && D == 3 && P == 3;
    if constexpr (MV) {
      double own[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) own[k] = sU[k][t];
      mfmaGradient<NF, NS, double, false>(s1, &sU[0][0], own, &sC[0][0], t, pJ, pdet, Q);
    } else {
      correctedGradient<D, P, NF, double>(s1, &sU[0][0], &sC[0][0], t, si, sj, sk, pJ, pdet, Q);
    }
    // cotangent of the total point flux
    double lFt[D][NF], lF[D][NF];
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int id = (d == 0) ? si : (d == 1) ? sj : sk;
#pragma unroll
      for (int k = 0; k < NF; k++) {
        double acc = 0;
        if constexpr (MV) {
          const MfmaLane<double> L(s1, t);
          const double b = d == 2 ? sL[k][t] : sL[k][d == 0 ? L.sx : L.sy];
          const double der = MfmaVec<double>::mma(L.aDT, b, 0.0);
          acc = d < 2 ? lanePull(d == 0 ? L.ix : L.iy, der) : der;
        } else {
#pragma unroll
          for (int m = 0; m < P1; m++) {
            const int sp = (d == 0) ? m + P1 * (sj + P1 * sk) : (d == 1) ? si + P1 * (m + P1 * sk) : si + P1 * (sj + P1 * m);
            acc += s1[O::DM + m * P1 + id] * sL[k][sp];
          }
        }
#pragma unroll
        for (int f = 0; f < NFACE; f++)
          if (faceAxis<D>(f) == d)
            acc -= (cside[f] ? 1.0 : -1.0) * s1[(cside[f] ? O::LR : O::LL) + cax[f]] * sJ[k][cq[f]];
        lFt[d][k] = acc;
      }
    }
#pragma unroll
    for (int d2 = 0; d2 < D; d2++)
#pragma unroll
      for (int k = 0; k < NF; k++) {
        double acc = 0;
#pragma unroll
        for (int d = 0; d < D; d++) acc += pJ[d * D + d2] * lFt[d][k];
        lF[d2][k] = acc;
      }
    double lQ[D][NF];
    // state directions: dual passes
#pragma unroll 1
    for (int sj = 0; sj < NF; sj++) {
      DM u[NF], Fi[D][NF], Fv[D][NF];
#pragma unroll
      for (int k = 0; k < NF; k++) u[k] = (k == sj) ? DM::var(sU[k][t], 0) : DM(sU[k][t]);
      // the gradient held fixed: plain doubles
      viscousFlux<D, NF, DM, double>(ph, u, Q, Fv);
      inviscidFlux<D, NF, DM>(ph, u, Fi);
      double acc = 0;
#pragma unroll
      for (int d2 = 0; d2 < D; d2++)
#pragma unroll
        for (int k = 0; k < NF; k++) acc += lF[d2][k] * (Fi[d2][k].d[0] + Fv[d2][k].d[0]);
      lU[sj] = acc;
    }
    // gradient directions: the point viscous flux is linear in the gradient
    // (zero at zero gradient), so its derivative along e_dk is its value there
    {
      double u[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) u[k] = sU[k][t];
      // (unrolled: each unit gradient's flux folds to a few terms)
#pragma unroll
      for (int dk = 0; dk < D * NF; dk++) {
        const int d = dk / NF, sj = dk - d * NF;
        double qd[D][NF], Fv[D][NF];
#pragma unroll
        for (int d2 = 0; d2 < D; d2++)
#pragma unroll
          for (int k = 0; k < NF; k++) qd[d2][k] = (d2 == d && k == sj) ? 1.0 : 0.0;
        viscousFlux<D, NF, double>(ph, u, qd, Fv);
        double acc = 0;
#pragma unroll
        for (int d2 = 0; d2 < D; d2++)
#pragma unroll
          for (int k = 0; k < NF; k++) acc += lF[d2][k] * Fv[d2][k];
        lQ[d][sj] = acc;
      }
    }
    // gradient traces and state traces, transposed
#pragma unroll
    for (int f = 0; f < NFACE; f++) {
      const double c = s1[(cside[f] ? O::LR : O::LL) + cax[f]];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        lU[k] += c * sUo[k][cq[f]];
#pragma unroll
        for (int d = 0; d < D; d++) lQ[d][k] += c * sQo[d][k][cq[f]];
      }
    }
    // Q = JGinv^T Qref / detJ, transposed
    const double invDet = 1.0 / pdet;
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int k = 0; k < NF; k++) {
        double acc = 0;
#pragma unroll
        for (int d1 = 0; d1 < D; d1++) acc += pJ[d * D + d1] * lQ[d1][k];
        sLQ[d][k][t] = invDet * acc;
      }
  }
  __syncthreads();
  if (t < NS) {
    int si, sj, sk;
    if (D == 3) {
      sk = t / (P1 * P1);
      sj = (t - P1 * P1 * sk) / P1;
      si = t - P1 * sj - P1 * P1 * sk;
    } else {
      si = t % P1;
      sj = t / P1;
      sk = 0;
    }
    // Qref = Dm U + ..., transposed along the lines
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int id = (d == 0) ? si : (d == 1) ? sj : sk;
#pragma unroll
      for (int k = 0; k < NF; k++) {
        double acc = 0;
#pragma unroll
        for (int m = 0; m < P1; m++) {
          const int sp = (d == 0) ? m + P1 * (sj + P1 * sk) : (d == 1) ? si + P1 * (m + P1 * sk) : si + P1 * (sj + P1 * m);
          acc += s1[O::DM + m * P1 + id] * sLQ[d][k][sp];
        }
        lU[k] += acc;
      }
    }
#pragma unroll
    for (int k = 0; k < NF; k++) out[((size_t)e * NF + k) * NS + t] = lU[k];
  }
  // cotangent of the jumps: the lift along each flux point's line
  for (int q = t; q < NFP; q += TPE) {
    int base, stride, side, dn;
    fptLine<D, P>(q, base, stride, side, dn);
    const double *gq = s1 + (side ? O::GR : O::GL);
    const double sg = side ? 1.0 : -1.0;
#pragma unroll
    for (int k = 0; k < NF; k++) {
      double acc = 0;
#pragma unroll
      for (int m = 0; m < P1; m++) acc += gq[m] * sLQ[dn][k][base + m * stride];
      B[((size_t)e * NF + k) * NFP + q] = sg * acc;
    }
  }
}

// The common solution's pull-back: u^ = (uL+uR)/2 - pen (uL-uR) inside
// (either side's jump u^ - u uses it), (u + ghost(u))/2 on boundaries.
template <int D, int P, int NF, bool BND>
__global__ __launch_bounds__((ElemGeo<D, P>::TPE)) void k_vjp_ldg(
    const double *__restrict__ U, const double *__restrict__ B, const double *__restrict__ fgeo,
    const int32_t *__restrict__ fnbr, const int8_t *__restrict__ finfo, const double *__restrict__ oned,
    Phys ph, int nEles, double *__restrict__ out) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  using DN = Dual<NF>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, P1 = G::P1, TPE = G::TPE, NFACE = G::NFACE;
  __shared__ double sU[NF][NS];
  __shared__ double sUo[NF][NFP];
  __shared__ double s1[O::SIZE];
  const int e = blockIdx.x, t = threadIdx.x;
  for (int i = t; i < O::SIZE; i += TPE) s1[i] = oned[i];
  for (int i = t; i < NF * NS; i += TPE) (&sU[0][0])[i] = U[(size_t)e * NF * NS + i];
  __syncthreads();
  for (int q = t; q < NFP; q += TPE) {
    const int fi = finfo[(size_t)e * NFACE + q / NFPF];
    const bool bnd = BND && fi >= FINFO_BOUNDARY;
    double bo[NF], lUC[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) bo[k] = B[((size_t)e * NF + k) * NFP + q];
    if (bnd) {
      int base, stride, side, dn;
      fptLine<D, P>(q, base, stride, side, dn);
      const double *l1 = s1 + (side ? O::LR : O::LL);
      DN uo[NF], ug[NF], n[D];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        double acc = 0;
#pragma unroll
        for (int m = 0; m < P1; m++) acc += l1[m] * sU[k][base + m * stride];
        uo[k] = DN::var(acc, k);
      }
#pragma unroll
      for (int c = 0; c < D; c++) n[c] = DN(fgeo[((size_t)e * (D + 1) + c) * NFP + q]);
      boundaryState<D, NF, DN>(ph, fi - FINFO_BOUNDARY, uo, n, ug);
#pragma unroll
      for (int j = 0; j < NF; j++) {
        double acc = 0;
#pragma unroll
        for (int k = 0; k < NF; k++) acc += bo[k] * 0.5 * ((j == k ? 1.0 : 0.0) + ug[k].d[j]);
        sUo[j][q] = acc - bo[j];
      }
    } else {
      const int en = fnbr[(size_t)e * NFACE + q / NFPF];
      const int qn = neighbourFpt<D, P>(q, fi);
      double nv[D];
#pragma unroll
      for (int c = 0; c < D; c++) nv[c] = fgeo[((size_t)e * (D + 1) + c) * NFP + q];
      const double pen = ldgPenalty<D, double>(ph, nv);
      const bool right = (fi >> 5) & 1;
      const double coef = right ? 0.5 + pen : 0.5 - pen;
#pragma unroll
      for (int k = 0; k < NF; k++) {
        lUC[k] = bo[k] + B[((size_t)en * NF + k) * NFP + qn];
        sUo[k][q] = coef * lUC[k] - bo[k];
      }
    }
  }
  __syncthreads();
  if (t < NS) {
    int cq[NFACE], cside[NFACE], cax[NFACE];
    corrEntries<D, P>(t, cq, cside, cax);
#pragma unroll
    for (int k = 0; k < NF; k++) {
      double acc = 0;
#pragma unroll
      for (int f = 0; f < NFACE; f++) acc += s1[(cside[f] ? O::LR : O::LL) + cax[f]] * sUo[k][cq[f]];
      out[((size_t)e * NF + k) * NS + t] += acc;
    }
  }
}

// out = alpha * (dt / detJ) * x + beta * (dt / detJ) * y  (y may be null)
static __global__ void k_dtj_combine(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ detJ, const double *__restrict__ dtE, int dtStride,
                              double alpha, double beta, int NF, int NS, long n, double *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long e = i / ((long)NF * NS);
  const int s = (int)(i % NS);
  const double dtJ = dtE[e * dtStride] / detJ[e * NS + s];
  double v = alpha * x[i] * dtJ;
  if (y) v += beta * y[i] * dtJ;
  out[i] = v;
}

// acc += x
static __global__ void k_accumulate(double *__restrict__ acc, const double *__restrict__ x, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) acc[i] += x[i];
}

// gradient of the kinetic energy sum_e sum_s w_s detJ |m|^2 / (2 rho) with
// respect to U: (-|m|^2 / (2 rho^2), m / rho, 0) * w_s detJ
template <int D>
__global__ void k_ke_grad(const double *__restrict__ U, const double *__restrict__ detJ,
                          const double *__restrict__ wts, int NS, long nPts, double *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nPts) return;
  const long e = i / NS;
  const int s = (int)(i % NS);
  const double *Ue = U + (size_t)e * (D + 2) * NS;
  double *Oe = out + (size_t)e * (D + 2) * NS;
  const double w = wts[s] * detJ[i];
  const double rho = Ue[s];
  double m2 = 0;
  for (int d = 0; d < D; d++) m2 += Ue[(d + 1) * NS + s] * Ue[(d + 1) * NS + s];
  Oe[s] = -0.5 * m2 / (rho * rho) * w;
  for (int d = 0; d < D; d++) Oe[(d + 1) * NS + s] = Ue[(d + 1) * NS + s] / rho * w;
  Oe[(D + 1) * NS + s] = 0.0;
}

}  // namespace frdev
