// fr_tri.hpp -- the stage kernel of the linear triangle (2D), dense operators.
//
// The triangle's FR operators (fr_tri_ops.cpp) are dense: (P+1)(P+2)/2 solution
// points with no tensor lines.  A workgroup of 256 threads holds 256 / TPE
// elements of TPE = triTPE(P) threads each (one solution point and one flux
// point per thread: 8 threads at P = 1, 16 at P = 2..4, 32 at P = 5), several
// elements per wavefront; the batched kernels may instead use TriGeo's
// layout 1 (TPE = max(points), elements straddling wavefronts).  The table
// Dr | Ds | E | L (at most 1638 values) is staged in LDS once per workgroup.
// The phases follow the quad kernel's (k_stage): physical and reference flux
// at the solution points and the common flux at the flux points; the jump
// Fn dA - disFn; divergence plus correction; the stage update; the traces of
// the new state.  The stage arguments and modes are those of k_stage; every
// face's traces go through HBM (no in-block exchange).
#pragma once

#include "fr_batch.hpp"
#include "fr_kernels.hpp"

namespace frdev {

// LAY: tile layout, 0 = triTPE(P) threads per element (every per-mesh kernel),
// 1 = max(solution points, flux points) (a choice of the batched kernels'
// tuner); NTH = EPB * TPE <= 256 threads per workgroup.
template <int P, int LAY = 0>
struct TriGeo {
  static constexpr int P1 = P + 1, NS = P1 * (P + 2) / 2, NFPF = P1, NFACE = 3, NFP = 3 * P1;
  static constexpr int TPE = triTPEL(P, LAY), EPB = 256 / TPE, NTH = EPB * TPE;
  static constexpr int DR = 0, DS = NS * NS, E = 2 * NS * NS, L = 2 * NS * NS + NFP * NS;
  static constexpr int TAB = 2 * NS * NS + 2 * NFP * NS;
  static_assert(NS <= TPE && NFP <= TPE, "one point per thread");
};

template <int P, int NF, typename T, int LAY = 0>
struct TriSmem {
  using G = TriGeo<P, LAY>;
  T tab[G::TAB];
  T sU[G::EPB][NF][G::NS];     // stage state
  T sF[G::EPB][2][NF][G::NS];  // reference flux F~_r, F~_s
  T sJ[G::EPB][NF][G::NFP];    // jump Fn dA - disFn
  T sN[G::EPB][NF][G::NS];     // updated state
};

// Reference unit normal of the edge of flux point q (edges v0v1, v1v2, v2v0).
template <typename T>
__device__ __forceinline__ void triNormal(int f, T &nx, T &ny) {
  const T h = (T)0.70710678118654752440;
  nx = (f == 0) ? (T)0 : (f == 1) ? h : (T)-1;
  ny = (f == 0) ? (T)-1 : (f == 1) ? h : (T)0;
}

// One RK stage of element block b (EPB elements) with the arguments a.
template <int P, int NF, typename T, bool BND, bool FIN, int LAY = 0>
__device__ __forceinline__ void triStageBody(const StageArgs<T> &a, const int b, TriSmem<P, NF, T, LAY> &sm,
                                             const int tid) {
  using G = TriGeo<P, LAY>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, TPE = G::TPE, EPB = G::EPB, NFACE = G::NFACE;
  const int slot = tid / TPE, t = tid - slot * TPE;
  const int e = b * EPB + slot;
  const bool act = e < a.nEles;
  const bool hasS = act && t < NS, hasQ = act && t < NFP;
  const size_t ec = (size_t)(act ? e : a.nEles - 1);
  const int tc = t < NS ? t : NS - 1, qc = t < NFP ? t : NFP - 1;

  for (int i = tid; i < G::TAB; i += G::NTH) sm.tab[i] = a.oned[i];
  if (act)
    for (int i = t; i < NF * NS; i += TPE) (&sm.sU[slot][0][0])[i] = a.Uin[ec * NF * NS + i];
  // this thread's solution point: metrics, detJ, dt, base state
  T pJ[4];
#pragma unroll
  for (int c = 0; c < 4; c++) pJ[c] = a.JG[(ec * 4 + c) * NS + tc];
  const double pdet = a.detJ[ec * NS + tc];
  const double pdt = a.dtE[ec * a.dtStride];
  double pb[NF];
#pragma unroll
  for (int k = 0; k < NF; k++)
    pb[k] = FIN ? a.Ubase[(ec * NF + k) * NS + tc] : (double)a.UbaseT[(ec * NF + k) * NS + tc];
  // this thread's flux point: face record, normal, dA, neighbour trace
  const int fi = a.finfo[ec * NFACE + qc / NFPF];
  const int nbe = a.fnbr[ec * NFACE + qc / NFPF];
  const bool bnd = BND && fi >= FINFO_BOUNDARY;
  T n[2], un[NF];
  n[0] = a.fgeo[(ec * 3 + 0) * NFP + qc];
  n[1] = a.fgeo[(ec * 3 + 1) * NFP + qc];
  const T dA = a.fgeo[(ec * 3 + 2) * NFP + qc];
  if (!bnd && nbe >= 0) {
    const T *Ur = a.UfIn + (size_t)nbe * NF * NFP + neighbourFpt<2, P>(qc, fi);
#pragma unroll
    for (int k = 0; k < NF; k++) un[k] = Ur[k * NFP];
  } else {
#pragma unroll
    for (int k = 0; k < NF; k++) un[k] = (T)0;
  }
  __syncthreads();

  const T *Dr = sm.tab + G::DR, *Ds = sm.tab + G::DS, *Em = sm.tab + G::E, *Lm = sm.tab + G::L;
  // (1) flux at the solution point, common flux at the flux point
  if (hasS) {
    T u[NF], F[2][NF];
#pragma unroll
    for (int k = 0; k < NF; k++) u[k] = sm.sU[slot][k][t];
    inviscidFlux<2, NF, T>(a.ph, u, F);
#pragma unroll
    for (int d = 0; d < 2; d++)
#pragma unroll
      for (int k = 0; k < NF; k++) sm.sF[slot][d][k][t] = pJ[d * 2 + 0] * F[0][k] + pJ[d * 2 + 1] * F[1][k];
  }
  T Fn[NF];
  if (hasQ) {
    T uo[NF], ub[NF], Fc[NF];
    // operator entries read once for all fields (the LDS reads bound these loops)
#pragma unroll
    for (int k = 0; k < NF; k++) uo[k] = (T)0;
#pragma unroll 3
    for (int s = 0; s < NS; s++) {
      const T w = Em[t * NS + s];
#pragma unroll
      for (int k = 0; k < NF; k++) uo[k] += w * sm.sU[slot][k][s];
    }
    const bool right = !bnd && ((fi >> 5) & 1);
    if (bnd) {
      boundaryState<2, NF, T>(a.ph, fi - FINFO_BOUNDARY, uo, n, ub);
    } else {
#pragma unroll
      for (int k = 0; k < NF; k++) ub[k] = un[k];
    }
    // the left element's state first, chosen per lane (one common flux per
    // wavefront instead of one per side; right implies !bnd)
    T sL[NF], sR[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) {
      sL[k] = right ? ub[k] : uo[k];
      sR[k] = right ? uo[k] : ub[k];
    }
    commonFlux<2, NF, T>(a.ph, bnd, sL, sR, n, Fc);
    const T scale = right ? -dA : dA;
#pragma unroll
    for (int k = 0; k < NF; k++) Fn[k] = Fc[k] * scale;
  }
  elementSync<TPE>();
  // (2) jump at the flux point: Fn dA - (E F~) . tNorm
  if (hasQ) {
    T tx, ty;
    triNormal<T>(t / NFPF, tx, ty);
    T fr[NF], fs[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) fr[k] = fs[k] = (T)0;
#pragma unroll 3
    for (int s = 0; s < NS; s++) {
      const T w = Em[t * NS + s];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        fr[k] += w * sm.sF[slot][0][k][s];
        fs[k] += w * sm.sF[slot][1][k][s];
      }
    }
#pragma unroll
    for (int k = 0; k < NF; k++) sm.sJ[slot][k][t] = Fn[k] - (tx * fr[k] + ty * fs[k]);
  }
  elementSync<TPE>();
  // (3) divergence + correction, stage update (as k_stage phase 4)
  if (hasS) {
    const size_t eo = ec * NF * NS;
    T dvs[NF], cor[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) dvs[k] = cor[k] = (T)0;
#pragma unroll 3
    for (int s = 0; s < NS; s++) {
      const T wr = Dr[t * NS + s], ws = Ds[t * NS + s];
#pragma unroll
      for (int k = 0; k < NF; k++) dvs[k] += wr * sm.sF[slot][0][k][s] + ws * sm.sF[slot][1][k][s];
    }
#pragma unroll 3
    for (int q = 0; q < NFP; q++) {
      const T wl = Lm[t * NFP + q];
#pragma unroll
      for (int k = 0; k < NF; k++) cor[k] += wl * sm.sJ[slot][k][q];
    }
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const T dv = dvs[k] + cor[k];
      const size_t gi = eo + (size_t)k * NS + t;
      if (a.divF) a.divF[gi] = dv;
      if (a.mode == 0) continue;
      const double dtJ = pdt / pdet;
      double u = pb[k];
      if (a.mode == 2) {
        const T r0 = (a.nprev > 0) ? a.dprev0[gi] : (T)0;
        const T r1 = (a.nprev > 1) ? a.dprev1[gi] : (T)0;
        const T r2 = (a.nprev > 2) ? a.dprev2[gi] : (T)0;
        if (a.nprev > 0) u -= a.c0 * (double)r0 * dtJ;
        if (a.nprev > 1) u -= a.c1 * (double)r1 * dtJ;
        if (a.nprev > 2) u -= a.c2 * (double)r2 * dtJ;
        const double cl = (a.nprev == 0) ? a.c0 : (a.nprev == 1) ? a.c1 : (a.nprev == 2) ? a.c2 : a.c3;
        u -= cl * (double)dv * dtJ;
      } else {
        u = u - a.rk * (double)dv * dtJ;
      }
      if (a.mode == 1) {
        a.Unext[gi] = (T)u;
      } else {
        a.Uout[gi] = u;
        if (a.Uout32) a.Uout32[gi] = (float)u;
      }
      sm.sN[slot][k][t] = (T)u;
    }
  }
  if (a.mode == 0) return;
  elementSync<TPE>();
  // (4) traces of the updated state
  if (hasQ) {
    T *Fe = a.UfOut + ec * NF * NFP;
    T acc[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) acc[k] = (T)0;
#pragma unroll 3
    for (int s = 0; s < NS; s++) {
      const T w = Em[t * NS + s];
#pragma unroll
      for (int k = 0; k < NF; k++) acc[k] += w * sm.sN[slot][k][s];
    }
#pragma unroll
    for (int k = 0; k < NF; k++) Fe[k * NFP + t] = acc[k];
  }
}

template <int P, int NF, typename T, bool BND, bool FIN>
__global__ __launch_bounds__(256) void k_tri_stage(StageArgs<T> a) {
  __shared__ TriSmem<P, NF, T> sm;
  const int b = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);  // XCD-contiguous blocks
  triStageBody<P, NF, T, BND, FIN>(a, b, sm, (int)threadIdx.x);
}

// Batched: a persistent workgroup of TriGeo<P, LAY>::NTH threads walks the
// triangle meshes' element blocks of order P (see k_batch_stage), with each
// mesh's arguments.
template <int P, int NF, typename T, bool FIN, int LAY>
__global__ __launch_bounds__(256) void k_batch_tri_stage(const StageArgs<T> *__restrict__ args,
                                                         const BatchItem *__restrict__ items, int nItems) {
  __shared__ TriSmem<P, NF, T, LAY> sm;
  for (int it = blockIdx.x; it < nItems; it += gridDim.x) {
    const BatchItem w = items[it];
    triStageBody<P, NF, T, true, FIN, LAY>(args[w.inst], w.block, sm, itemThreadIdx());
    __syncthreads();  // the next item reuses the LDS
  }
}

// CFL time step of each triangle, as k_dt for quads (ele::calcWaveSpFpts,
// ele.cpp:1031-1084; ele::calcDt, ele.cpp:1086-1105): the traces E U at the
// flux points, the wave speed |v.n| + c (Euler) or |a.n| (advection) with the
// element's own normals, and dt = factor * vol / sum(waveSp * wSurf * dA * h),
// where vol = sum(detJ * w) with the nodal basis' quadrature weights (the
// element's area: detJ is constant on a linear triangle).  dA is relative to
// the reference edge's unit normal; h (sqrt 2 on the hypotenuse v1v2, else 1)
// makes sum(wSurf * dA * h) over an edge its physical length, as it is on
// quads.  One thread per element.
template <int P, int NF>
__global__ __launch_bounds__(256) void k_tri_dt(const double *__restrict__ U, const double *__restrict__ ownN,
                                                const double *__restrict__ fgeo, const double *__restrict__ detJ,
                                                const double *__restrict__ wts, const double *__restrict__ wsurf,
                                                const double *__restrict__ tab, double factor, Phys ph, int nEles,
                                                double *__restrict__ dtE) {
  using G = TriGeo<P>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEles) return;
  double vol = 0;
  for (int s = 0; s < NS; s++) vol += detJ[(size_t)e * NS + s] * wts[s];
  const double *Ue = U + (size_t)e * NF * NS;
  double intWave = 0;
  for (int q = 0; q < NFP; q++) {
    const double nx = ownN[((size_t)e * 2 + 0) * NFP + q], ny = ownN[((size_t)e * 2 + 1) * NFP + q];
    double ws;
    if constexpr (NF == 1) {
      ws = fabs(ph.advect[0] * nx + ph.advect[1] * ny);
    } else {
      const double *Eq = tab + G::E + q * NS;
      double u[NF];
      for (int k = 0; k < NF; k++) {
        double acc = 0;
        for (int s = 0; s < NS; s++) acc += Eq[s] * Ue[k * NS + s];
        u[k] = acc;
      }
      const double rho = u[0], vx = u[1] / rho, vy = u[2] / rho;
      const double pr = (ph.gamma - 1) * (u[3] - 0.5 * rho * (vx * vx + vy * vy));
      const double csq = fmax(ph.gamma * pr / rho, 0.0);
      ws = fabs(vx * nx + vy * ny) + sqrt(csq);
    }
    const double h = (q / NFPF == 1) ? 1.41421356237309504880 : 1.0;
    intWave += ws * wsurf[q % NFPF] * fgeo[((size_t)e * 3 + 2) * NFP + q] * h;
  }
  dtE[e] = factor * vol / intWave;
}

// Face traces U_fpts = E U_spts, one thread per (element, flux point).
template <int P, int NF, typename T>
__global__ void k_tri_trace(const T *__restrict__ U, T *__restrict__ Uf, const T *__restrict__ tab, int nEles) {
  using G = TriGeo<P>;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nEles * G::NFP) return;
  const long e = idx / G::NFP;
  const int q = (int)(idx - e * G::NFP);
  const T *Em = tab + G::E + q * G::NS;
#pragma unroll
  for (int k = 0; k < NF; k++) {
    T acc = (T)0;
    for (int s = 0; s < G::NS; s++) acc += Em[s] * U[(e * NF + k) * G::NS + s];
    Uf[(e * NF + k) * G::NFP + q] = acc;
  }
}

}  // namespace frdev
