// fr_physics.hpp -- pointwise physics shared by the residual kernels:
// Euler / advection fluxes, common (Riemann) fluxes and boundary states.
// Every routine restates one reference function; arithmetic is written in the
// same order so FP64 results agree with the reference to rounding.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>

namespace frdev {

struct Phys {
  double gamma, lambda, diffD;
  double advect[3];
  double rhoBound, uBound, vBound, wBound, pBound;
  double TWall, RGas, uWall, vWall, wWall;
  double penFact, tau, mu_inf, prandtl, rt_inf, c_sth;
  int equation, riemannType, fixVis, viscFix;
  // Energy offset of the FP32 states (inviscid Euler): they hold E - eOff,
  // where eOff = pOff / (gamma - 1) for a reference pressure pOff near the
  // flow's mean.  The FP32 flux routines then work with p - pOff, whose
  // rounding is relative to the pressure variation instead of the (low-Mach)
  // background; hOff = gamma * eOff restores the total enthalpy of the energy
  // flux.  The constant pOff dropped from the momentum flux is a
  // divergence-free term of the discretisation (free-stream preservation).
  // Zero in FP64 contexts, whose routines ignore them.
  double eOff, pOff, hOff;
};

// FP32 states carry the energy offset (Phys::eOff); FP64 and dual-number
// states do not.
template <typename T>
constexpr bool hasEnergyOffset() {
  return sizeof(T) == 4;
}

// 1 / x for the flux routines: the hardware reciprocal estimate refined by
// Newton-Raphson steps, e = 1 - x r, r <- r + r e (each step squares the
// relative error), within an ulp of the correctly rounded value for
// densities of ordinary magnitude.  FP64: two steps from the 2^-22-accurate
// estimate, half the instructions of a full IEEE division; FP32: one step
// from the 1-ulp estimate, three instructions instead of ten.
template <typename T>
__device__ __forceinline__ T recip(T x) {
  if constexpr (sizeof(T) == 8) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
  } else if constexpr (std::is_same<T, float>::value) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
  } else {
    return (T)1 / x;  // dual numbers (the adjoint's derivatives)
  }
}

// sqrt(x) of a non-negative x of ordinary magnitude (sound speeds, face
// areas); rootAndInv also gives 1 / sqrt(x).  FP64: Goldschmidt's coupled
// iteration for g -> sqrt(x), h -> 1 / (2 sqrt(x)) from the hardware
// reciprocal square root estimate y (relative error below 2^-22): g = x y,
// h = y / 2, then per step r = 1/2 - g h, g <- g + g r, h <- h + h r, which
// squares the relative error; a Newton correction g <- g + (x - g^2) h
// after the step leaves g within an ulp of the correctly rounded root.  An x
// below the smallest normal double is taken as that value (root 1.5e-154
// instead of 0: the estimate of 0 is infinite).  FP32: the hardware root
// and reciprocal (1 ulp each; the IEEE-rounded library root takes about 16
// instructions for its scaling and corrections).
__device__ __forceinline__ void rootSteps(double x, double &g, double &h, int steps) {
  x = fmax(x, 2.2250738585072014e-308);
  const double y = __builtin_amdgcn_rsq(x);
  g = x * y;
  h = 0.5 * y;
  for (int i = 0; i < steps; i++) {
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
  }
  g = fma(fma(-g, g, x), h, g);
}
template <typename T>
__device__ __forceinline__ T rootPos(T x) {
  if constexpr (sizeof(T) == 8) {
    double g, h;
    rootSteps(x, g, h, 1);
    return g;
  } else if constexpr (std::is_same<T, float>::value) {
    return __builtin_amdgcn_sqrtf(x);
  } else {
    return sqrt(x);  // dual numbers
  }
}
template <typename T>
__device__ __forceinline__ void rootAndInv(T x, T &root, T &inv) {
  if constexpr (sizeof(T) == 8) {
    double g, h;
    rootSteps(x, g, h, 2);
    root = g;
    inv = h + h;
  } else if constexpr (std::is_same<T, float>::value) {
    root = __builtin_amdgcn_sqrtf(x);
    inv = __builtin_amdgcn_rcpf(root);
  } else {
    root = sqrt(x);  // dual numbers
    inv = (T)1 / root;
  }
}

// Euler flux F[d][k] (inviscidFlux, reference src/flux.cpp:37-88) or linear
// advection flux.
template <int D, int NF, typename T>
__device__ __forceinline__ void inviscidFlux(const Phys &ph, const T *U, T F[D][NF]) {
  if constexpr (NF == 1) {
#pragma unroll
    for (int d = 0; d < D; d++) F[d][0] = (T)ph.advect[d] * U[0];
  } else if constexpr (D == 2) {
    // velocities use one reciprocal of rho instead of a division each (the
    // reference divides; the difference is at most an ulp per velocity)
    const T g = (T)ph.gamma;
    const T rho = U[0], ri = recip(rho);
    T u = U[1] * ri, v = U[2] * ri, E = U[3];
    T p = (g - (T)1.0) * (E - (T)0.5 * rho * (u * u + v * v));
    // total enthalpy rho*H = E + p (FP32: offset-free E and p plus hOff)
    T Ep = E + p;
    if constexpr (hasEnergyOffset<T>()) Ep = Ep + (T)ph.hOff;
    F[0][0] = rho * u;         F[1][0] = rho * v;
    F[0][1] = rho * u * u + p; F[1][1] = rho * u * v;
    F[0][2] = rho * v * u;     F[1][2] = rho * v * v + p;
    F[0][3] = Ep * u;          F[1][3] = Ep * v;
  } else {
    const T g = (T)ph.gamma;
    const T rho = U[0], ri = recip(rho);
    T u = U[1] * ri, v = U[2] * ri, w = U[3] * ri, E = U[4];
    T p = (g - (T)1.0) * (E - (T)0.5 * rho * (u * u + v * v + w * w));
    T Ep = E + p;
    if constexpr (hasEnergyOffset<T>()) Ep = Ep + (T)ph.hOff;
    F[0][0] = rho * u;         F[1][0] = rho * v;         F[2][0] = rho * w;
    F[0][1] = rho * u * u + p; F[1][1] = rho * u * v;     F[2][1] = rho * u * w;
    F[0][2] = rho * v * u;     F[1][2] = rho * v * v + p; F[2][2] = rho * v * w;
    F[0][3] = rho * w * u;     F[1][3] = rho * w * v;     F[2][3] = rho * w * w + p;
    F[0][4] = Ep * u;          F[1][4] = Ep * v;          F[2][4] = Ep * w;
  }
}

// Normal Euler flux F(U).n of one side plus its normal velocity and squared
// sound speed: the sum over d of n[d] F[d][i] that face::rusanovFlux forms
// from inviscidFlux (flux.cpp:37-88), factored through the normal velocity
// vn = v.n -- mass rho vn, momentum rho vn v_c + p n_c, energy (E + p) vn --
// without holding the D x NF flux table (the same terms, summed in another
// order: an ulp-level difference).  One reciprocal of rho (recip) replaces
// the reference's divisions (also an ulp-level difference).
template <int D, typename T>
__device__ __forceinline__ void normalFlux(const Phys &ph, const T g, const T *U, const T *n, T *Fn, T &vn, T &csq) {
  constexpr int NF = D + 2;
  const T rho = U[0], E = U[D + 1], ri = recip(rho);
  T v[D];
#pragma unroll
  for (int d = 0; d < D; d++) v[d] = U[d + 1] * ri;
  T q = v[0] * v[0] + v[1] * v[1];
  if (D == 3) q = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const T p = (g - (T)1.0) * (E - (T)0.5 * rho * q);
  T Ep = E + p, pAbs = p;  // FP32: offset-free E and p (Phys::eOff)
  if constexpr (hasEnergyOffset<T>()) {
    Ep = Ep + (T)ph.hOff;
    pAbs = p + (T)ph.pOff;
  }
  vn = n[0] * v[0];
#pragma unroll
  for (int d = 1; d < D; d++) vn += n[d] * v[d];
  const T mn = rho * vn;
  Fn[0] = mn;
#pragma unroll
  for (int c = 0; c < D; c++) Fn[1 + c] = mn * v[c] + p * n[c];
  Fn[D + 1] = Ep * vn;
  csq = fmax(g * pAbs * ri, (T)0.0);
}

// Rusanov / local Lax-Friedrichs flux for Euler (face::rusanovFlux,
// src/face.cpp:289-356, static mesh).  eigOut (may be null) receives
// max(eigL, eigR), the wave speed the reference stores for the CFL time step
// (face.cpp:354).  swap = true takes UA as the right state and UB as the
// left one: the flux is symmetric in the two sides but for the sign of the
// dissipation term eig (UR - UL), and sums, fmax and a negated product are
// exact under the exchange, so the result is bit for bit the flux of the
// exchanged arguments -- a lane picks its side by a sign, not by selecting
// every field.
template <int D, typename T>
__device__ __forceinline__ void rusanovN(const Phys &ph, const T *UA, const T *UB, const T *n, T *Fn,
                                         T *eigOut = nullptr, bool swap = false) {
  constexpr int NF = D + 2;
  const T g = (T)ph.gamma;
  T FnA[NF], FnB[NF], vnA, vnB, csqA, csqB;
  normalFlux<D, T>(ph, g, UA, n, FnA, vnA, csqA);
  normalFlux<D, T>(ph, g, UB, n, FnB, vnB, csqB);
  const T eig = fmax(fabs(vnA) + rootPos(csqA), fabs(vnB) + rootPos(csqB));
  if (eigOut) *eigOut = eig;
  const T eigS = swap ? -eig : eig;
#pragma unroll
  for (int i = 0; i < NF; i++) Fn[i] = (T)0.5 * (FnA[i] + FnB[i] - eigS * (UB[i] - UA[i]));
}

// Roe flux with Harten-type entropy fix, 2D (face::roeFlux,
// src/face.cpp:358-484).  The reference's divisions by the densities, by eps
// and by the Roe sound speed are multiplications by reciprocals (recip), an
// ulp-level difference.
template <typename T>
__device__ __forceinline__ void roe2d(const Phys &ph, const T *UL, const T *UR, const T *n, T *Fn) {
  const T g = (T)ph.gamma;
  const T riL = recip(UL[0]), riR = recip(UR[0]);
  T vL[2], vR[2], um[2], du[4];
  for (int i = 0; i < 2; i++) {
    vL[i] = UL[i + 1] * riL;
    vR[i] = UR[i + 1] * riR;
  }
  T pL = (g - (T)1.0) * (UL[3] - ((T)0.5 * UL[0] * (vL[0] * vL[0] + vL[1] * vL[1])));
  T pR = (g - (T)1.0) * (UR[3] - ((T)0.5 * UR[0] * (vR[0] * vR[0] + vR[1] * vR[1])));
  // FP32 (Phys::eOff): pL, pR are offset-free; the enthalpies and sound
  // speeds take the absolute values
  T EpL = UL[3] + pL, EpR = UR[3] + pR, pAL = pL, pAR = pR;
  if constexpr (hasEnergyOffset<T>()) {
    EpL = EpL + (T)ph.hOff;
    EpR = EpR + (T)ph.hOff;
    pAL = pL + (T)ph.pOff;
    pAR = pR + (T)ph.pOff;
  }
  T hL = EpL * riL, hR = EpR * riR;
  T sq_rho = sqrt(UR[0] * riL);
  T rrho = recip(sq_rho + (T)1.);
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
  T eps = (T)0.5 * (fabs(rhoUnL * riL - rhoUnR * riR) + fabs(sqrt(g * pAL * riL) - sqrt(g * pAR * riR)));
  const T reps = recip(eps);
  if (lambda0 < (T)2. * eps) lambda0 = (T)0.25 * lambda0 * lambda0 * reps + eps;
  if (lambdaP < (T)2. * eps) lambdaP = (T)0.25 * lambdaP * lambdaP * reps + eps;
  if (lambdaM < (T)2. * eps) lambdaM = (T)0.25 * lambdaM * lambdaM * reps + eps;
  T a2 = (T)0.5 * (lambdaP + lambdaM) - lambda0;
  T a3 = (T)0.5 * (lambdaP - lambdaM) * recip(am);
  T a1 = a2 * (g - (T)1.) * recip(am_sq);
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
    // FP32 states hold E - eOff (Phys::eOff): pL and the ghost energies below
    // are then offset-free (p - pOff, E - eOff), the absolute pressure pLa is
    // used where the formulas need it.  FP64: pLa = pL, pOff = 0.
    constexpr bool OFF = hasEnergyOffset<T>();
    T rhoR = 0, pR = 0, ER = 0;
    T vL[3] = {0, 0, 0}, vR[3] = {0, 0, 0};
    const T vB[3] = {(T)ph.uBound, (T)ph.vBound, (T)ph.wBound};
    T rhoL = UL[0], eL = UL[D + 1];
    for (int i = 0; i < D; i++) vL[i] = UL[i + 1] / UL[0];
    T vSq = 0;
    for (int i = 0; i < D; i++) vSq += vL[i] * vL[i];
    T pL = (g - (T)1.0) * (eL - (T)0.5 * rhoL * vSq);
    const T pLa = OFF ? pL + (T)ph.pOff : pL;
    // the ghost pressure pR is offset-free like pL
    const T pBnd = OFF ? (T)(ph.pBound - ph.pOff) : (T)ph.pBound;
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
        pR = pBnd;
        ER = energy();
        break;
      case 2:  // SUP_IN
        rhoR = (T)ph.rhoBound;
        vR[0] = vB[0]; vR[1] = vB[1]; vR[2] = vB[2];
        pR = pBnd;
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
        rhoR = pLa / ((T)ph.RGas * (T)ph.TWall);
        vR[0] = (T)ph.uWall;
        vR[1] = (T)ph.vWall;
        if (D == 3) vR[2] = (T)ph.wWall;
        ER = energy();
        break;
      case 9:  // ISOTHERMAL_NOSLIP
        pR = pL;
        rhoR = pLa / ((T)ph.RGas * (T)ph.TWall);
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
        T cL = sqrt(g * pLa / rhoL), cB = sqrt(g * (T)ph.pBound / (T)ph.rhoBound);
        T R_L = (fabs(vnB) >= cB && vnL >= 0) ? vnB + (T)2. / gmo * cB : vnL + (T)2. / gmo * cL;
        T R_B = (fabs(vnB) >= cB && vnL < 0) ? vnL - (T)2. / gmo * cL : vnB - (T)2. / gmo * cB;
        T rhoFac = (T).25 * (g - (T)1.) * (R_L - R_B);
        T rhoBG = rhoFac * rhoFac / g;
        if (vnL < 0) rhoBG *= pow((T)ph.rhoBound, g) / (T)ph.pBound;
        else rhoBG *= pow(rhoL, g) / pLa;
        rhoR = pow(rhoBG, (T)1. / gmo);
        pR = rhoFac * rhoFac * rhoR / g;
        if constexpr (OFF) pR = pR - (T)ph.pOff;
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

// ---------------------------------------------------------------------------
// viscous terms
// ---------------------------------------------------------------------------
// Navier-Stokes viscous flux Fv[d][k] from U and the physical gradient
// g[d][k] (viscousFlux, reference src/flux.cpp:90-226).  viscFix = 0 keeps the
// reference's index defects: tauxz = mu(du/dz + dv/dx), tauyz = mu(du/dz +
// dv/dy) (flux.cpp:197-198) and Fv[0][3] = -tauzz (flux.cpp:216); viscFix = 1
// uses the physical tauxz = mu(du/dz + dw/dx), tauyz = mu(dv/dz + dw/dy),
// Fv[0][3] = -tauxz.  Advection-diffusion: Fv = -diffD grad u
// (viscousFluxAD, flux.cpp:361-366).
// TG, the type of the gradient, may differ from T: the adjoint's state
// passes differentiate through U only and keep the gradient in plain doubles.
template <int D, int NF, typename T, typename TG = T>
__device__ __forceinline__ void viscousFlux(const Phys &ph, const T *U, const TG g[D][NF], T Fv[D][NF]) {
  if constexpr (NF == 1) {
#pragma unroll
    for (int d = 0; d < D; d++) Fv[d][0] = -(T)ph.diffD * g[d][0];
  } else {
    // FP64: one reciprocal of rho (recip) for the reference's divisions by
    // rho; other types (FP32, the adjoint's dual numbers) divide, which is
    // cheaper for them
    const T rho = U[0];
    const T ri = (sizeof(T) == 8) ? recip(rho) : (T)0;
    auto over = [&](const T &x) -> T {
      if constexpr (sizeof(T) == 8) return x * ri;
      else return x / rho;
    };
    const T u = over(U[1]), v = over(U[2]);
    T e = over(U[D + 1]) - (T)0.5 * (u * u + v * v);
    T w = (T)0;
    if constexpr (D == 3) {
      w = over(U[3]);
      e -= (T)0.5 * (w * w);
    }
    const TG dRho_dx = g[0][0], dRhoU_dx = g[0][1], dRhoV_dx = g[0][2], dE_dx = g[0][D + 1];
    const TG dRho_dy = g[1][0], dRhoU_dy = g[1][1], dRhoV_dy = g[1][2], dE_dy = g[1][D + 1];
    TG dRho_dz = 0, dRhoU_dz = 0, dRhoV_dz = 0, dRhoW_dx = 0, dRhoW_dy = 0, dRhoW_dz = 0, dE_dz = 0;
    if constexpr (D == 3) {
      dRho_dz = g[2][0];
      dRhoU_dz = g[2][1];
      dRhoV_dz = g[2][2];
      dRhoW_dx = g[0][3];
      dRhoW_dy = g[1][3];
      dRhoW_dz = g[2][3];
      dE_dz = g[2][4];
    }
    T mu = (T)ph.mu_inf;
    if (!ph.fixVis) {  // Sutherland's law
      const T rt = ((T)ph.gamma - (T)1.0) * e / (T)ph.rt_inf;
      mu *= pow(rt, (T)1.5) * ((T)1. + (T)ph.c_sth) / (rt + (T)ph.c_sth);
    }
    const T du_dx = over(dRhoU_dx - dRho_dx * u), du_dy = over(dRhoU_dy - dRho_dy * u);
    const T dv_dx = over(dRhoV_dx - dRho_dx * v), dv_dy = over(dRhoV_dy - dRho_dy * v);
    T du_dz = 0, dv_dz = 0, dw_dx = 0, dw_dy = 0, dw_dz = 0;
    if constexpr (D == 3) {
      du_dz = over(dRhoU_dz - dRho_dz * u);
      dv_dz = over(dRhoV_dz - dRho_dz * v);
      dw_dx = over(dRhoW_dx - dRho_dx * w);
      dw_dy = over(dRhoW_dy - dRho_dy * w);
      dw_dz = over(dRhoW_dz - dRho_dz * w);
    }
    T dK_dx, dK_dy, dK_dz = 0;
    if constexpr (D == 2) {
      dK_dx = (T)0.5 * (u * u + v * v) * dRho_dx + rho * (u * du_dx + v * dv_dx);
      dK_dy = (T)0.5 * (u * u + v * v) * dRho_dy + rho * (u * du_dy + v * dv_dy);
    } else {
      dK_dx = (T)0.5 * (u * u + v * v + w * w) * dRho_dx + rho * (u * du_dx + v * dv_dx + w * dw_dx);
      dK_dy = (T)0.5 * (u * u + v * v + w * w) * dRho_dy + rho * (u * du_dy + v * dv_dy + w * dw_dy);
      dK_dz = (T)0.5 * (u * u + v * v + w * w) * dRho_dz + rho * (u * du_dz + v * dv_dz + w * dw_dz);
    }
    const T de_dx = over(dE_dx - dK_dx - dRho_dx * e);
    const T de_dy = over(dE_dy - dK_dy - dRho_dy * e);
    T de_dz = 0;
    if constexpr (D == 3) de_dz = over(dE_dz - dK_dz - dRho_dz * e);
    const T diag = (du_dx + dv_dy + dw_dz) / (T)3.0;
    const T tauxx = (T)2.0 * mu * (du_dx - diag), tauyy = (T)2.0 * mu * (dv_dy - diag);
    const T tauxy = mu * (du_dy + dv_dx);
    T tauxz = 0, tauyz = 0, tauzz = 0;
    if constexpr (D == 3) {
      if (ph.viscFix) {
        tauxz = mu * (du_dz + dw_dx);
        tauyz = mu * (dv_dz + dw_dy);
      } else {
        tauxz = mu * (du_dz + dv_dx);
        tauyz = mu * (du_dz + dv_dy);
      }
      tauzz = (T)2.0 * mu * (dw_dz - diag);
    }
    const T kq = (mu / (T)ph.prandtl) * (T)ph.gamma;
    Fv[0][0] = (T)0.0;
    Fv[0][1] = -tauxx;
    Fv[0][2] = -tauxy;
    Fv[0][D + 1] = -(u * tauxx + v * tauxy + w * tauxz + kq * de_dx);
    Fv[1][0] = (T)0.0;
    Fv[1][1] = -tauxy;
    Fv[1][2] = -tauyy;
    Fv[1][D + 1] = -(u * tauxy + v * tauyy + w * tauyz + kq * de_dy);
    if constexpr (D == 3) {
      Fv[0][3] = ph.viscFix ? -tauxz : -tauzz;
      Fv[1][3] = -tauyz;
      Fv[2][0] = (T)0.0;
      Fv[2][1] = -tauxz;
      Fv[2][2] = -tauyz;
      Fv[2][3] = -tauzz;
      Fv[2][4] = -(u * tauxz + v * tauyz + w * tauzz + kq * de_dz);
    }
  }
}

// Adiabatic no-slip wall gradient state (boundFace::applyViscousBCs,
// boundFace.cpp:279-320): gR = gL, then the energy gradient is rebuilt with
// zero temperature gradient.  Other boundaries keep gR = gL.
template <int D, int NF, typename T>
__device__ __forceinline__ void boundaryGradient(const Phys &ph, int bc, const T *UL, const T *UR, const T gL[D][NF],
                                                 T gR[D][NF]) {
#pragma unroll
  for (int d = 0; d < D; d++)
#pragma unroll
    for (int k = 0; k < NF; k++) gR[d][k] = gL[d][k];
  if constexpr (NF > 1) {
    if (bc != 10) return;
    T rhovSq = 0;
    for (int d = 0; d < D; d++) rhovSq += UL[d + 1] * UL[d + 1];
    const T pL = ((T)ph.gamma - (T)1.) * (UL[D + 1] - (T)0.5 * rhovSq / UL[0]);
    const T e = pL / (((T)ph.gamma - (T)1.) * UR[0]);
    T gv[D][D];
    for (int d1 = 0; d1 < D; d1++)
      for (int d2 = 0; d2 < D; d2++) gv[d1][d2] = (gR[d1][d2 + 1] - gR[d1][0] * UR[d2 + 1] / UR[0]) / UR[0];
    T vSq = 0;
    for (int d = 0; d < D; d++) vSq += UR[d + 1] * UR[d + 1] / (UR[0] * UR[0]);
    for (int d1 = 0; d1 < D; d1++) {
      gR[d1][D + 1] = (e + (T)0.5 * vSq) * gR[d1][0];
      for (int d2 = 0; d2 < D; d2++) gR[d1][D + 1] += UR[d2 + 1] * gv[d2][d1];
    }
  }
}

// LDG penalty direction switch (face.cpp:587-597 / 206-215): -penFact when
// n.(1, 1, sqrt 2) < 0 (n.(1,1) in 2D).
template <int D, typename T>
__device__ __forceinline__ T ldgPenalty(const Phys &ph, const T *n) {
  T pen = (T)ph.penFact;
  if (D == 2) {
    if (n[0] + n[1] < 0) pen = -(T)ph.penFact;
  } else if (n[0] + n[1] + (T)1.4142135623730951 * n[2] < 0) {
    pen = -(T)ph.penFact;
  }
  return pen;
}

// Adds the common LDG viscous normal flux sum_d Fc[d] n[d] to the inviscid
// common flux Fn (face::calcViscousFlux, face.cpp:177-287), in the left
// element's frame.  bnd: 0 interior, 1 Dirichlet-type boundary, 2 adiabatic
// wall.
template <int D, int NF, typename T, typename TG = T>
__device__ __forceinline__ void ldgViscousFlux(const Phys &ph, int bnd, const T *UL, const T *UR, const TG gL[D][NF],
                                               const TG gR[D][NF], const T *n, T *Fn) {
  T Fc[D][NF];
  if constexpr (NF == 1) {
    // advection-diffusion: central flux of -diffD grad u (face.cpp:270-279)
    T FL[D][NF], FR[D][NF];
    viscousFlux<D, NF, T, TG>(ph, UL, gL, FL);
    viscousFlux<D, NF, T, TG>(ph, UR, gR, FR);
    T t = 0;
#pragma unroll
    for (int d = 0; d < D; d++) t += (T)0.5 * (FL[d][0] + FR[d][0]) * n[d];
    Fn[0] += t;
    return;
  }
  if (bnd) {
    T Fb[D][NF];
    if (bnd > 1) viscousFlux<D, NF, T, TG>(ph, UR, gR, Fb);
    else viscousFlux<D, NF, T, TG>(ph, UL, gL, Fb);
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int k = 0; k < NF; k++) Fc[d][k] = Fb[d][k] + (T)ph.tau * n[d] * (UL[k] - UR[k]);
  } else {
    T FL[D][NF], FR[D][NF];
    viscousFlux<D, NF, T, TG>(ph, UL, gL, FL);
    viscousFlux<D, NF, T, TG>(ph, UR, gR, FR);
    const T pen = ldgPenalty<D, T>(ph, n);
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const T dF0 = FL[0][k] - FR[0][k], dF1 = FL[1][k] - FR[1][k];
      const T dU = UL[k] - UR[k];
      T nd;
      if constexpr (D == 2) {
        nd = n[0] * dF0 + n[1] * dF1;
      } else {
        // face.cpp:245 takes tempFL[1] here; viscFix = 1 takes tempFL[2]
        const T dF2 = (ph.viscFix ? FL[2][k] : FL[1][k]) - FR[2][k];
        nd = n[0] * dF0 + n[1] * dF1 + n[2] * dF2;
      }
#pragma unroll
      for (int d = 0; d < D; d++)
        Fc[d][k] = (T)0.5 * (FL[d][k] + FR[d][k]) + pen * n[d] * nd + (T)ph.tau * n[d] * dU;
    }
  }
#pragma unroll
  for (int k = 0; k < NF; k++)
#pragma unroll
    for (int d = 0; d < D; d++) Fn[k] += Fc[d][k] * n[d];
}

}  // namespace frdev
