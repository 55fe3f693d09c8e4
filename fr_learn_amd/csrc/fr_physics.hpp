// fr_physics.hpp -- pointwise physics shared by the residual kernels:
// Euler / advection fluxes, common (Riemann) fluxes and boundary states.
// Every routine restates one reference function; arithmetic is written in the
// same order so FP64 results agree with the reference to rounding.
#pragma once

#include <hip/hip_runtime.h>

namespace frdev {

struct Phys {
  double gamma, lambda, diffD;
  double advect[3];
  double rhoBound, uBound, vBound, wBound, pBound;
  double TWall, RGas, uWall, vWall, wWall;
  double penFact, tau, mu_inf, prandtl, rt_inf, c_sth;
  int equation, riemannType, fixVis, viscFix;
};

// Euler flux F[d][k] (inviscidFlux, reference src/flux.cpp:37-88) or linear
// advection flux.
template <int D, int NF, typename T>
__device__ __forceinline__ void inviscidFlux(const Phys &ph, const T *U, T F[D][NF]) {
  if constexpr (NF == 1) {
#pragma unroll
    for (int d = 0; d < D; d++) F[d][0] = (T)ph.advect[d] * U[0];
  } else if constexpr (D == 2) {
    const T g = (T)ph.gamma;
    T rho = U[0], u = U[1] / rho, v = U[2] / rho, E = U[3];
    T p = (g - (T)1.0) * (E - (T)0.5 * rho * (u * u + v * v));
    F[0][0] = rho * u;         F[1][0] = rho * v;
    F[0][1] = rho * u * u + p; F[1][1] = rho * u * v;
    F[0][2] = rho * v * u;     F[1][2] = rho * v * v + p;
    F[0][3] = (E + p) * u;     F[1][3] = (E + p) * v;
  } else {
    const T g = (T)ph.gamma;
    T rho = U[0], u = U[1] / rho, v = U[2] / rho, w = U[3] / rho, E = U[4];
    T p = (g - (T)1.0) * (E - (T)0.5 * rho * (u * u + v * v + w * w));
    F[0][0] = rho * u;         F[1][0] = rho * v;         F[2][0] = rho * w;
    F[0][1] = rho * u * u + p; F[1][1] = rho * u * v;     F[2][1] = rho * u * w;
    F[0][2] = rho * v * u;     F[1][2] = rho * v * v + p; F[2][2] = rho * v * w;
    F[0][3] = rho * w * u;     F[1][3] = rho * w * v;     F[2][3] = rho * w * w + p;
    F[0][4] = (E + p) * u;     F[1][4] = (E + p) * v;     F[2][4] = (E + p) * w;
  }
}

// Rusanov / local Lax-Friedrichs flux for Euler (face::rusanovFlux,
// src/face.cpp:289-356, static mesh).
template <int D, typename T>
__device__ __forceinline__ void rusanov(const Phys &ph, const T *UL, const T *UR, const T *n, T *Fn) {
  constexpr int NF = D + 2;
  T FL[D][NF], FR[D][NF];
  inviscidFlux<D, NF, T>(ph, UL, FL);
  inviscidFlux<D, NF, T>(ph, UR, FR);
  const T g = (T)ph.gamma;
  T FnL[NF], FnR[NF];
#pragma unroll
  for (int i = 0; i < NF; i++) FnL[i] = FnR[i] = (T)0;
  T rhoL = UL[0], rhoR = UR[0];
  T uL = UL[1] / rhoL, uR = UR[1] / rhoR, vL = UL[2] / rhoL, vR = UR[2] / rhoR;
  T pL, pR;
  if constexpr (D == 2) {
    pL = (g - (T)1.0) * (UL[3] - (T)0.5 * rhoL * (uL * uL + vL * vL));
    pR = (g - (T)1.0) * (UR[3] - (T)0.5 * rhoR * (uR * uR + vR * vR));
  } else {
    T wL = UL[3] / rhoL, wR = UR[3] / rhoR;
    pL = (g - (T)1.0) * (UL[4] - (T)0.5 * rhoL * (uL * uL + vL * vL + wL * wL));
    pR = (g - (T)1.0) * (UR[4] - (T)0.5 * rhoR * (uR * uR + vR * vR + wR * wR));
  }
  T vnL = (T)0, vnR = (T)0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    vnL += n[d] * UL[d + 1] / rhoL;
    vnR += n[d] * UR[d + 1] / rhoR;
#pragma unroll
    for (int i = 0; i < NF; i++) {
      FnL[i] += n[d] * FL[d][i];
      FnR[i] += n[d] * FR[d][i];
    }
  }
  T csqL = fmax(g * pL / rhoL, (T)0.0), csqR = fmax(g * pR / rhoR, (T)0.0);
  T eigL = fabs(vnL) + sqrt(csqL), eigR = fabs(vnR) + sqrt(csqR);
  T eig = fmax(eigL, eigR);
#pragma unroll
  for (int i = 0; i < NF; i++) Fn[i] = (T)0.5 * (FnL[i] + FnR[i] - eig * (UR[i] - UL[i]));
}

// Roe flux with Harten-type entropy fix, 2D (face::roeFlux,
// src/face.cpp:358-484).
template <typename T>
__device__ __forceinline__ void roe2d(const Phys &ph, const T *UL, const T *UR, const T *n, T *Fn) {
  const T g = (T)ph.gamma;
  T vL[2], vR[2], um[2], du[4];
  for (int i = 0; i < 2; i++) {
    vL[i] = UL[i + 1] / UL[0];
    vR[i] = UR[i + 1] / UR[0];
  }
  T pL = (g - (T)1.0) * (UL[3] - ((T)0.5 * UL[0] * (vL[0] * vL[0] + vL[1] * vL[1])));
  T pR = (g - (T)1.0) * (UR[3] - ((T)0.5 * UR[0] * (vR[0] * vR[0] + vR[1] * vR[1])));
  T hL = (UL[3] + pL) / UL[0], hR = (UR[3] + pR) / UR[0];
  T sq_rho = sqrt(UR[0] / UL[0]);
  T rrho = (T)1. / (sq_rho + (T)1.);
  T usq = 0, unm = 0;
  for (int i = 0; i < 2; i++) {
    um[i] = rrho * (vL[i] + sq_rho * vR[i]);
    usq += (T)0.5 * um[i] * um[i];
    unm += um[i] * n[i];
  }
  T hm = rrho * (hL + sq_rho * hR);
  T am_sq = (g - (T)1.) * (hm - usq);
  T am = sqrt(am_sq);
  T rhoUnL = 0, rhoUnR = 0;
  for (int i = 0; i < 2; i++) {
    rhoUnL += UL[i + 1] * n[i];
    rhoUnR += UR[i + 1] * n[i];
  }
  Fn[0] = rhoUnL + rhoUnR;
  Fn[1] = rhoUnL * vL[0] + rhoUnR * vR[0] + (pL + pR) * n[0];
  Fn[2] = rhoUnL * vL[1] + rhoUnR * vR[1] + (pL + pR) * n[1];
  Fn[3] = rhoUnL * hL + rhoUnR * hR;
  for (int i = 0; i < 4; i++) du[i] = UR[i] - UL[i];
  T lambda0 = fabs(unm), lambdaP = fabs(unm + am), lambdaM = fabs(unm - am);
  T eps = (T)0.5 * (fabs(rhoUnL / UL[0] - rhoUnR / UR[0]) + fabs(sqrt(g * pL / UL[0]) - sqrt(g * pR / UR[0])));
  if (lambda0 < (T)2. * eps) lambda0 = (T)0.25 * lambda0 * lambda0 / eps + eps;
  if (lambdaP < (T)2. * eps) lambdaP = (T)0.25 * lambdaP * lambdaP / eps + eps;
  if (lambdaM < (T)2. * eps) lambdaM = (T)0.25 * lambdaM * lambdaM / eps + eps;
  T a2 = (T)0.5 * (lambdaP + lambdaM) - lambda0;
  T a3 = (T)0.5 * (lambdaP - lambdaM) / am;
  T a1 = a2 * (g - (T)1.) / am_sq;
  T a4 = a3 * (g - (T)1.);
  T a5 = usq * du[0] - um[0] * du[1] - um[1] * du[2] + du[3];
  T a6 = unm * du[0] - n[0] * du[1] - n[1] * du[2];
  T aL1 = a1 * a5 - a3 * a6, bL1 = a4 * a5 - a2 * a6;
  Fn[0] -= lambda0 * du[0] + aL1;
  Fn[1] -= lambda0 * du[1] + aL1 * um[0] + bL1 * n[0];
  Fn[2] -= lambda0 * du[2] + aL1 * um[1] + bL1 * n[1];
  Fn[3] -= lambda0 * du[3] + aL1 * hm + bL1 * unm;
  for (int i = 0; i < 4; i++) Fn[i] *= (T)0.5;
}

// Lax-Friedrichs flux for linear advection (face::laxFriedrichsFlux,
// src/face.cpp:486-514).
template <int D, typename T>
__device__ __forceinline__ void laxFriedrichs(const Phys &ph, const T *UL, const T *UR, const T *n, T *Fn) {
  T uAvg = (T)0.5 * (UL[0] + UR[0]);
  T uDiff = UL[0] - UR[0];
  T vNorm = (T)ph.advect[0] * n[0] + (T)ph.advect[1] * n[1];
  if constexpr (D == 3) vNorm += (T)ph.advect[2] * n[2];
  Fn[0] = vNorm * uAvg + (T)0.5 * (T)ph.lambda * fabs(vNorm) * uDiff;
}

// Ghost (right) state of a boundary flux point (boundFace::applyBCs,
// src/boundFace.cpp:33-277; static grid, so the grid velocity is zero).
template <int D, int NF, typename T>
__device__ __forceinline__ void boundaryState(const Phys &ph, int bc, const T *UL, const T *n, T *UR) {
  if constexpr (NF == 1) {
    UR[0] = (T)0;  // the reference leaves the advection ghost state at zero
  } else {
    const T g = (T)ph.gamma, gmo = g - (T)1.;
    T rhoR = 0, pR = 0, ER = 0;
    T vL[3] = {0, 0, 0}, vR[3] = {0, 0, 0};
    const T vB[3] = {(T)ph.uBound, (T)ph.vBound, (T)ph.wBound};
    T rhoL = UL[0], eL = UL[D + 1];
    for (int i = 0; i < D; i++) vL[i] = UL[i + 1] / UL[0];
    T vSq = 0;
    for (int i = 0; i < D; i++) vSq += vL[i] * vL[i];
    T pL = (g - (T)1.0) * (eL - (T)0.5 * rhoL * vSq);
    auto energy = [&](void) {
      T s = 0;
      for (int i = 0; i < D; i++) s += vR[i] * vR[i];
      return (pR / (g - (T)1.0)) + (T)0.5 * rhoR * s;
    };
    switch (bc) {
      case 4:  // SUB_IN
        rhoR = (T)ph.rhoBound;
        for (int i = 0; i < D; i++) vR[i] = vB[i];
        pR = pL;
        ER = energy();
        break;
      case 5:  // SUB_OUT
        rhoR = rhoL;
        for (int i = 0; i < D; i++) vR[i] = vL[i];
        pR = (T)ph.pBound;
        ER = energy();
        break;
      case 2:  // SUP_IN
        rhoR = (T)ph.rhoBound;
        vR[0] = vB[0]; vR[1] = vB[1]; vR[2] = vB[2];
        pR = (T)ph.pBound;
        ER = energy();
        break;
      case 3:  // SUP_OUT
        rhoR = rhoL;
        for (int i = 0; i < D; i++) vR[i] = vL[i];
        ER = eL;
        break;
      case 8:   // SLIP_WALL
      case 12: {  // SYMMETRY
        rhoR = rhoL;
        T vnL = 0;
        for (int i = 0; i < D; i++) vnL += vL[i] * n[i];
        for (int i = 0; i < D; i++) vR[i] = vL[i] - (T)2.0 * vnL * n[i];
        ER = eL;
        break;
      }
      case 13:  // ISOTHERMAL_NOSLIP_MOVING
        pR = pL;
        rhoR = pR / ((T)ph.RGas * (T)ph.TWall);
        vR[0] = (T)ph.uWall;
        vR[1] = (T)ph.vWall;
        if (D == 3) vR[2] = (T)ph.wWall;
        ER = energy();
        break;
      case 9:  // ISOTHERMAL_NOSLIP
        pR = pL;
        rhoR = pR / ((T)ph.RGas * (T)ph.TWall);
        ER = energy();
        break;
      case 10:  // ADIABATIC_NOSLIP
        rhoR = rhoL;
        pR = pL;
        ER = energy();
        break;
      case 1: {  // CHAR_INOUT (Riemann invariants)
        T vnL = 0, vnB = 0;
        for (int i = 0; i < D; i++) vnL += vL[i] * n[i];
        for (int i = 0; i < D; i++) vnB += vB[i] * n[i];
        T cL = sqrt(g * pL / rhoL), cB = sqrt(g * (T)ph.pBound / (T)ph.rhoBound);
        T R_L = (fabs(vnB) >= cB && vnL >= 0) ? vnB + (T)2. / gmo * cB : vnL + (T)2. / gmo * cL;
        T R_B = (fabs(vnB) >= cB && vnL < 0) ? vnL - (T)2. / gmo * cL : vnB - (T)2. / gmo * cB;
        T rhoFac = (T).25 * (g - (T)1.) * (R_L - R_B);
        T rhoBG = rhoFac * rhoFac / g;
        if (vnL < 0) rhoBG *= pow((T)ph.rhoBound, g) / (T)ph.pBound;
        else rhoBG *= pow(rhoL, g) / pL;
        rhoR = pow(rhoBG, (T)1. / gmo);
        pR = rhoFac * rhoFac * rhoR / g;
        T VB = (T)0.5 * (R_L + R_B);
        for (int i = 0; i < D; i++) vR[i] = (vnL < 0) ? vB[i] + (VB - vnB) * n[i] : vL[i] + (VB - vnL) * n[i];
        T vM2 = 0;
        for (int i = 0; i < D; i++) vM2 += vR[i] * vR[i];
        ER = pR / gmo + (T).5 * rhoR * vM2;
        break;
      }
      default:
        break;
    }
    UR[0] = rhoR;
    for (int i = 0; i < D; i++) UR[i + 1] = rhoR * vR[i];
    UR[D + 1] = ER;
  }
}

}  // namespace frdev
