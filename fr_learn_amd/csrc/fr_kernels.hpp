// fr_kernels.hpp -- HIP kernels of the FR residual + RK path (gfx950).
//
// Device layout (element-major, solution/flux point fastest; one element's
// data is one contiguous block so a wavefront reads it coalesced):
//   U, U0, divF   [ele][field][spt]
//   JG            [ele][dimRef*D + dimPhys][spt]   adjugate metrics det(J) J^-1
//   detJ          [ele][spt]
//   Uf            [ele][field][fpt]                extrapolated face traces
//   fgeo          [ele][c][fpt]   c < D: unit normal of the face's LEFT element,
//                                 c == D: this element's own dA
//   nbr           [ele][fpt]      neighbour trace base (ele'*NF*NFP + fpt') or -1
//   fkind         [ele][face]     0 left side, 1 right side, 2+bc boundary
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fr_layout.hpp"
#include "fr_physics.hpp"

namespace frdev {

template <typename T>
struct ResArgs {
  const T *__restrict__ U;
  const T *__restrict__ Uf;
  const T *__restrict__ JG;
  const T *__restrict__ fgeo;
  const int32_t *__restrict__ nbr;
  const int8_t *__restrict__ fkind;
  const T *__restrict__ oned;
  T *__restrict__ divF;
  int nEles;
  Phys ph;
};

// Common normal flux at one face point, (UL, UR) in the left element's frame.
template <int D, int NF, typename T>
__device__ __forceinline__ void commonFlux(const Phys &ph, bool boundary, const T *UL, const T *UR, const T *n,
                                           T *Fn) {
  if constexpr (NF == 1) {
    laxFriedrichs<D, T>(ph, UL, UR, n, Fn);
  } else if constexpr (D == 2) {
    if (!boundary && ph.riemannType == 1) roe2d<T>(ph, UL, UR, n, Fn);
    else rusanov<D, T>(ph, UL, UR, n, Fn);
  } else {
    rusanov<D, T>(ph, UL, UR, n, Fn);
  }
}

// ---------------------------------------------------------------------------
// Face traces U_fpts = opp_spts_to_fpts * U_spts (solver::extrapolateU,
// reference src/solver.cpp:401-412), one thread per (element, flux point).
// ---------------------------------------------------------------------------
template <int D, int P, int NF, typename T>
__global__ __launch_bounds__(256) void k_trace(const T *__restrict__ U, T *__restrict__ Uf,
                                               const T *__restrict__ oned, int nEles) {
  using G = ElemGeo<D, P>;
  constexpr int NS = G::NS, NFP = G::NFP, P1 = G::P1;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nEles * NFP) return;
  const int e = (int)(idx / NFP), q = (int)(idx - (long)e * NFP);
  int base, stride, side, dn;
  fptLine<D, P>(q, base, stride, side, dn);
  const T *l1 = oned + (side ? OneD<P>::LR : OneD<P>::LL);
  const T *Ue = U + (size_t)e * NF * NS;
  T *Fe = Uf + (size_t)e * NF * NFP;
#pragma unroll
  for (int k = 0; k < NF; k++) {
    T acc = (T)0;
#pragma unroll
    for (int m = 0; m < P1; m++) acc += l1[m] * Ue[k * NS + base + m * stride];
    Fe[k * NFP + q] = acc;
  }
}

// ---------------------------------------------------------------------------
// Fused inviscid residual of one element per TPE-thread group:
//   calcInviscidFlux_spts (solver.cpp:514-539), calcInviscidFlux_faces
//   (face.cpp:132-175 + intFace/boundFace hooks), extrapolateNormalFlux
//   (solver.cpp:764-786), calcDivF_spts (solver.cpp:707-727) and
//   correctDivFlux (solver.cpp:789-806).
// Output divF = reference-space residual divF_spts[step].
// ---------------------------------------------------------------------------
template <int D, int P, int NF, typename T>
__global__ __launch_bounds__((ElemGeo<D, P>::NTH)) void k_residual(ResArgs<T> a) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, P1 = G::P1, TPE = G::TPE, EPB = G::EPB;
  constexpr int NFACE = G::NFACE;
  __shared__ T sU[EPB][NF][NS];
  __shared__ T sF[EPB][D][NF][NS];
  __shared__ T sJ[EPB][NF][NFP];
  __shared__ T s1[O::SIZE];
  const int slot = threadIdx.x / TPE, t = threadIdx.x - slot * TPE;
  const int e = blockIdx.x * EPB + slot;
  const bool act = e < a.nEles;
  for (int i = threadIdx.x; i < O::SIZE; i += blockDim.x) s1[i] = a.oned[i];
  if (act) {
    const T *Ue = a.U + (size_t)e * NF * NS;
    for (int i = t; i < NF * NS; i += TPE) (&sU[slot][0][0])[i] = Ue[i];
  }
  __syncthreads();

  // (1) point fluxes, transformed to reference space F~_d = sum JGinv(d,d2) F_d2
  if (act) {
    for (int s = t; s < NS; s += TPE) {
      T u[NF], F[D][NF], J[D * D];
#pragma unroll
      for (int k = 0; k < NF; k++) u[k] = sU[slot][k][s];
      inviscidFlux<D, NF, T>(a.ph, u, F);
#pragma unroll
      for (int c = 0; c < D * D; c++) J[c] = a.JG[((size_t)e * D * D + c) * NS + s];
#pragma unroll
      for (int d1 = 0; d1 < D; d1++)
#pragma unroll
        for (int k = 0; k < NF; k++) {
          T acc = (T)0;
#pragma unroll
          for (int d2 = 0; d2 < D; d2++) acc += J[d1 * D + d2] * F[d2][k];
          sF[slot][d1][k][s] = acc;
        }
    }
  }
  __syncthreads();

  // (2) face points: own trace, discontinuous normal flux, common flux, jump
  if (act) {
    for (int q = t; q < NFP; q += TPE) {
      int base, stride, side, dn;
      fptLine<D, P>(q, base, stride, side, dn);
      const T *l1 = s1 + (side ? O::LR : O::LL);
      const T sg = side ? (T)1 : (T)-1;
      T uo[NF], dis[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        T acc = (T)0, acc2 = (T)0;
#pragma unroll
        for (int m = 0; m < P1; m++) {
          const int s = base + m * stride;
          acc += l1[m] * sU[slot][k][s];
          acc2 += (l1[m] * sg) * sF[slot][dn][k][s];
        }
        uo[k] = acc;
        dis[k] = acc2;
      }
      const int fk = a.fkind[(size_t)e * NFACE + q / NFPF];
      T n[D];
#pragma unroll
      for (int c = 0; c < D; c++) n[c] = a.fgeo[((size_t)e * (D + 1) + c) * NFP + q];
      const T dA = a.fgeo[((size_t)e * (D + 1) + D) * NFP + q];
      T un[NF], Fn[NF];
      if (fk >= 2) {
        boundaryState<D, NF, T>(a.ph, fk - 2, uo, n, un);
      } else {
        const int32_t g = a.nbr[(size_t)e * NFP + q];
#pragma unroll
        for (int k = 0; k < NF; k++) un[k] = a.Uf[(size_t)g + k * NFP];
      }
      if (fk == 1) commonFlux<D, NF, T>(a.ph, false, un, uo, n, Fn);
      else commonFlux<D, NF, T>(a.ph, fk >= 2, uo, un, n, Fn);
      const T scale = (fk == 1) ? -dA : dA;
#pragma unroll
      for (int k = 0; k < NF; k++) sJ[slot][k][q] = Fn[k] * scale - dis[k];
    }
  }
  __syncthreads();

  // (3) divergence of F~ plus correction of the normal-flux jump
  if (act) {
    T *out = a.divF + (size_t)e * NF * NS;
    for (int s = t; s < NS; s += TPE) {
      int i, j, kk;
      if (D == 3) {
        kk = s / (P1 * P1);
        j = (s - P1 * P1 * kk) / P1;
        i = s - P1 * j - P1 * P1 * kk;
      } else {
        i = s % P1;
        j = s / P1;
        kk = 0;
      }
      int cq[NFACE], cside[NFACE], cax[NFACE];
      corrEntries<D, P>(s, cq, cside, cax);
      T cval[NFACE];
#pragma unroll
      for (int f = 0; f < NFACE; f++) cval[f] = s1[(cside[f] ? O::GR : O::GL) + cax[f]];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        T div, acc = (T)0;
#pragma unroll
        for (int m = 0; m < P1; m++) acc += s1[O::DM + i * P1 + m] * sF[slot][0][k][m + P1 * (j + P1 * kk)];
        div = acc;
        acc = (T)0;
#pragma unroll
        for (int m = 0; m < P1; m++) acc += s1[O::DM + j * P1 + m] * sF[slot][1][k][i + P1 * (m + P1 * kk)];
        div = acc + div;
        if constexpr (D == 3) {
          acc = (T)0;
#pragma unroll
          for (int m = 0; m < P1; m++) acc += s1[O::DM + kk * P1 + m] * sF[slot][2][k][i + P1 * (j + P1 * m)];
          div = acc + div;
        }
        T corr = (T)0;
#pragma unroll
        for (int f = 0; f < NFACE; f++) corr += cval[f] * sJ[slot][k][cq[f]];
        out[k * NS + s] = corr + div;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Metrics (solver::setPosSptsFpts + calcTransforms, solver.cpp:904-951,
// 1227-1356) from the linear element's nodes X[ele][node][3].  One thread per
// (element, point); points 0..NS-1 are spts, the rest fpts.
// ---------------------------------------------------------------------------
template <int D>
__global__ void k_metrics(const double *__restrict__ X, const double *__restrict__ dshS,
                          const double *__restrict__ dshF, const double *__restrict__ tN, int NS, int NFP, int nN,
                          int nEles, double *__restrict__ JG, double *__restrict__ detJ, double *__restrict__ ownN,
                          double *__restrict__ fgeo, int *__restrict__ negFlag) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int NP = NS + NFP;
  if (idx >= (long)nEles * NP) return;
  const int e = (int)(idx / NP), p = (int)(idx - (long)e * NP);
  const bool isS = p < NS;
  const int pt = isS ? p : p - NS, npt = isS ? NS : NFP;
  const double *dsh = isS ? dshS : dshF;
  double J[3][3];  // J[dimRef][dimPhys]
  for (int d = 0; d < D; d++)
    for (int d2 = 0; d2 < D; d2++) {
      double acc = 0.0;
      for (int n = 0; n < nN; n++) acc += dsh[((size_t)d * npt + pt) * nN + n] * X[((size_t)e * nN + n) * 3 + d2];
      J[d][d2] = acc;
    }
  double G[3][3], det;
  if (D == 2) {
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    G[0][0] = J[1][1];
    G[0][1] = -J[1][0];
    G[1][0] = -J[0][1];
    G[1][1] = J[0][0];
  } else {
    const double xr = J[0][0], xs = J[1][0], xt = J[2][0];
    const double yr = J[0][1], ys = J[1][1], yt = J[2][1];
    const double zr = J[0][2], zs = J[1][2], zt = J[2][2];
    det = xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
    G[0][0] = ys * zt - yt * zs; G[0][1] = xt * zs - xs * zt; G[0][2] = xs * yt - xt * ys;
    G[1][0] = yt * zr - yr * zt; G[1][1] = xr * zt - xt * zr; G[1][2] = xt * yr - xr * yt;
    G[2][0] = yr * zs - ys * zr; G[2][1] = xs * zr - xr * zs; G[2][2] = xr * ys - xs * yr;
  }
  if (isS) {
    for (int d1 = 0; d1 < D; d1++)
      for (int d2 = 0; d2 < D; d2++) JG[((size_t)e * D * D + d1 * D + d2) * NS + pt] = G[d1][d2];
    detJ[(size_t)e * NS + pt] = det;
    if (det < 0) atomicOr(negFlag, 1);
  } else {
    double nrm[3];
    for (int d1 = 0; d1 < D; d1++) {
      nrm[d1] = 0.;
      for (int d2 = 0; d2 < D; d2++) nrm[d1] += G[d2][d1] * tN[pt * 3 + d2];
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
    for (int d = 0; d < D; d++) ownN[((size_t)e * D + d) * NFP + pt] = nrm[d];
    fgeo[((size_t)e * (D + 1) + D) * NFP + pt] = dA;
  }
}

// Face-normal assembly: a right-side flux point uses the left element's
// normal at the paired point (intFace computes the common flux with normL,
// face.cpp:289-356, and scales by -dAR, intFace.cpp:116-120).
template <int D>
__global__ void k_fgeo(const double *__restrict__ ownN, const int32_t *__restrict__ pairEF,
                       const int8_t *__restrict__ fkind, int NFP, int NFPF, int nEles, double *__restrict__ fgeo) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nEles * NFP) return;
  const int e = (int)(idx / NFP), q = (int)(idx - (long)e * NFP);
  const int fk = fkind[(size_t)e * 2 * D + q / NFPF];
  int se = e, sq = q;
  if (fk == 1) {
    const int32_t g = pairEF[(size_t)e * NFP + q];
    se = g / NFP;
    sq = g - se * NFP;
  }
  for (int c = 0; c < D; c++) fgeo[((size_t)e * (D + 1) + c) * NFP + q] = ownN[((size_t)se * D + c) * NFP + sq];
}

// ---------------------------------------------------------------------------
// Runge-Kutta stage updates (solver::timeStepA / timeStepB,
// solver.cpp:307-385), elementwise over [ele][field][spt].
// ---------------------------------------------------------------------------
// U = U0 - rk * divF / detJ * dt; with copy != 0 first sets U0 = U
// (copyUspts_U0, solver.cpp:387-392).  U32 (optional) receives a float copy.
template <typename TD>
__global__ void k_rk_a(double *__restrict__ U, double *__restrict__ U0, const TD *__restrict__ divF,
                       const double *__restrict__ detJ, float *__restrict__ U32, double rk, double dt, int copy,
                       int NF, int NS, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long e = i / ((long)NF * NS);
  const int s = (int)(i % NS);
  double u0;
  if (copy) {
    u0 = U[i];
    U0[i] = u0;
  } else {
    u0 = U0[i];
  }
  const double v = u0 - rk * (double)divF[i] / detJ[e * NS + s] * dt;
  U[i] = v;
  if (U32) U32[i] = (float)v;
}

// Final combination: U = U0 (copyU0_Uspts, solver.cpp:394-399, skipped for a
// one-stage scheme), then U -= RKb[s] * divF[s] / detJ * dt for s = 0..nRK-1
// in order -- the same operations as nRK separate timeStepB passes.
template <typename TD>
__global__ void k_rk_b(double *__restrict__ U, const double *__restrict__ U0, const TD *__restrict__ d0,
                       const TD *__restrict__ d1, const TD *__restrict__ d2, const TD *__restrict__ d3,
                       const double *__restrict__ detJ, float *__restrict__ U32, double b0, double b1, double b2,
                       double b3, double dt, int nRK, int NF, int NS, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long e = i / ((long)NF * NS);
  const int s = (int)(i % NS);
  const double J = detJ[e * NS + s];
  double u = (nRK > 1) ? U0[i] : U[i];
  u -= b0 * (double)d0[i] / J * dt;
  if (nRK > 1) u -= b1 * (double)d1[i] / J * dt;
  if (nRK > 2) u -= b2 * (double)d2[i] / J * dt;
  if (nRK > 3) u -= b3 * (double)d3[i] / J * dt;
  U[i] = u;
  if (U32) U32[i] = (float)u;
}

// Single timeStepB pass (solver.cpp:347-385) for the ABI entry point.
template <typename TD>
__global__ void k_rk_b1(double *__restrict__ U, const TD *__restrict__ dF, const double *__restrict__ detJ,
                        float *__restrict__ U32, double b, double dt, int NF, int NS, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long e = i / ((long)NF * NS);
  const int s = (int)(i % NS);
  const double u = U[i] - b * (double)dF[i] / detJ[e * NS + s] * dt;
  U[i] = u;
  if (U32) U32[i] = (float)u;
}

__global__ void k_to_float(const double *__restrict__ a, float *__restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------
// Reductions (FP64).  One thread per element accumulates in the reference's
// loop order, then a block tree reduction; partial results per block go to
// `part` and are summed on the host in block order (deterministic).
// ---------------------------------------------------------------------------
// residual norm of divF_spts[0] (ele::getNormResidual, ele.cpp:1882-1903)
template <typename TD>
__global__ __launch_bounds__(256) void k_norm(const TD *__restrict__ divF, const double *__restrict__ detJ,
                                              const double *__restrict__ wts, int normType, int NF, int NS,
                                              int nEles, double *__restrict__ part) {
  __shared__ double red[256][5];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double r[5] = {0, 0, 0, 0, 0};
  if (e < nEles) {
    for (int s = 0; s < NS; s++) {
      const double J = detJ[(size_t)e * NS + s];
      for (int k = 0; k < NF; k++) {
        const double v = (double)divF[((size_t)e * NF + k) * NS + s];
        if (normType == 1) r[k] += fabs(v) * wts[s];
        else if (normType == 2) r[k] += v * v / J * wts[s];
        else r[k] = fmax(fabs(v) / J, r[k]);
      }
    }
  }
  for (int k = 0; k < 5; k++) red[threadIdx.x][k] = r[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 5; k++) {
        const double o = red[threadIdx.x + w][k];
        red[threadIdx.x][k] = (normType == 3) ? fmax(red[threadIdx.x][k], o) : red[threadIdx.x][k] + o;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; k++) part[(size_t)blockIdx.x * 5 + k] = red[0][k];
}

// kinetic energy 0.5*rho*|u|^2 and volume integrated with the spt quadrature
template <int D>
__global__ __launch_bounds__(256) void k_ke(const double *__restrict__ U, const double *__restrict__ detJ,
                                            const double *__restrict__ wts, int NS, int nEles,
                                            double *__restrict__ part) {
  __shared__ double red[256][2];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double ke = 0, vol = 0;
  if (e < nEles) {
    const double *Ue = U + (size_t)e * (D + 2) * NS;
    for (int s = 0; s < NS; s++) {
      const double w = wts[s] * detJ[(size_t)e * NS + s];
      const double rho = Ue[s];
      double m2 = 0;
      for (int d = 0; d < D; d++) m2 += Ue[(d + 1) * NS + s] * Ue[(d + 1) * NS + s];
      ke += w * 0.5 * m2 / rho;
      vol += w;
    }
  }
  red[threadIdx.x][0] = ke;
  red[threadIdx.x][1] = vol;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x][0] += red[threadIdx.x + w][0];
      red[threadIdx.x][1] += red[threadIdx.x + w][1];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[(size_t)blockIdx.x * 2 + 0] = red[0][0];
    part[(size_t)blockIdx.x * 2 + 1] = red[0][1];
  }
}

}  // namespace frdev
