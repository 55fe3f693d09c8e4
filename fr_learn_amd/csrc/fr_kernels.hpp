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
//   fnbr          [ele][face]     neighbour element (interior faces)
//   finfo         [ele][face]     packed face record (fr_layout.hpp packFinfo)
//   dUf           [ele][dim][field][fpt]  gradient traces (viscous only)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fr_layout.hpp"
#include "fr_physics.hpp"

#ifndef FR_DEFER_NBR
#define FR_DEFER_NBR 0  // intermediate inviscid stages load the neighbour traces in the face phase
#endif
#ifndef FR_MFMA
#define FR_MFMA 1  // hex P3 inviscid stages: the line contractions of F~ on the matrix cores
#endif
#ifndef FR_FOLD
#define FR_FOLD 1  // stage kernels: the jump correction folded into the derivative
#endif
#ifndef FR_INBLOCK
#define FR_INBLOCK 1  // neighbours in the same workgroup: traces extrapolated from their LDS state (VALU variants)
#endif
#ifndef FR_MF_WAVES
#define FR_MF_WAVES 3  // waves per SIMD asked of the hex P3 matrix-core stage kernels
#endif
#ifndef FR_MF_FIN_WAVES
#define FR_MF_FIN_WAVES 3  // the same for their final-stage variant
#endif
#ifndef FR_STAGE_MINW
#define FR_STAGE_MINW 1  // waves per SIMD asked of the intermediate inviscid stage kernel
#endif

namespace frdev {

// Common normal flux at one face point, (UL, UR) in the left element's frame.
// ws (may be null) receives the wave speed the reference stores at the face's
// left flux point for the CFL time step: max(eigL, eigR) of the Rusanov flux
// (face.cpp:354; boundary faces always use Rusanov, face.cpp:140-150) or
// |a.n| of Lax-Friedrichs (face.cpp:498-508).  Returns false where the
// reference stores none (the Roe flux, face.cpp:358-484).
template <int D, int NF, typename T>
__device__ __forceinline__ bool commonFlux(const Phys &ph, bool boundary, const T *UL, const T *UR, const T *n,
                                           T *Fn, T *ws = nullptr) {
  if constexpr (NF == 1) {
    laxFriedrichs<D, T>(ph, UL, UR, n, Fn);
    if (ws) {
      T vN = (T)ph.advect[0] * n[0] + (T)ph.advect[1] * n[1];
      if constexpr (D == 3) vN += (T)ph.advect[2] * n[2];
      *ws = fabs(vN);
    }
    return true;
  } else if constexpr (D == 2) {
    if (!boundary && ph.riemannType == 1) {
      roe2d<T>(ph, UL, UR, n, Fn);
      return false;
    }
    rusanovN<D, T>(ph, UL, UR, n, Fn, ws);
    return true;
  } else {
    rusanovN<D, T>(ph, UL, UR, n, Fn, ws);
    return true;
  }
}

// commonFlux of the left / right pair (right ? (UB, UA) : (UA, UB)): the
// Rusanov flux takes the order as a sign (rusanovN), the others get the
// states selected per field.
template <int D, int NF, typename T>
__device__ __forceinline__ bool commonFluxSided(const Phys &ph, bool boundary, bool right, const T *UA, const T *UB,
                                                const T *n, T *Fn, T *ws = nullptr) {
  if constexpr (NF > 1 && D == 3) {
    rusanovN<D, T>(ph, UA, UB, n, Fn, ws, right);
    return true;
  } else {
    T sL[NF], sR[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) {
      sL[k] = right ? UB[k] : UA[k];
      sR[k] = right ? UA[k] : UB[k];
    }
    return commonFlux<D, NF, T>(ph, boundary, sL, sR, n, Fn, ws);
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
// One RK stage of one element per TPE-thread group:
//   calcInviscidFlux_spts (solver.cpp:514-539), calcInviscidFlux_faces
//   (face.cpp:132-175 + intFace/boundFace hooks), extrapolateNormalFlux
//   (solver.cpp:764-786), calcDivF_spts (solver.cpp:707-727),
//   correctDivFlux (solver.cpp:789-806), then the stage update
//   (timeStepA / timeStepB, solver.cpp:307-385) and the face traces of the
//   updated state for the next stage (extrapolateU, solver.cpp:401-412).
// mode 0: residual only (divF out).
// mode 1: stage update Unext = UbaseT - rk * divF / detJ * dt (the base in
//         the residual precision: the FP32 path's stage states are FP32);
//         Unext may be null when the next stage forms its state itself.
// With rprev set, the stage state is not read: it is formed from the base
// and the previous stage's residual, Us = base - rkPrev * rprev / detJ * dt
// (the mode-1 update of the previous stage), so that an RK step stores no
// intermediate stage state.
// mode 2: final update U = Ubase - sum_r RKb[r] divF_r / detJ * dt, the
//         subtractions in stage order as the timeStepB passes do.
// mode 3: final Jameson-type update U = Ubase - rk * divF / detJ * dt.
// The state an element reads (Uin / Ubase) is read only by that element, so
// Unext / Uout may alias them; traces are double-buffered (UfIn != UfOut).
// ---------------------------------------------------------------------------
// Geometry addressing.  Per point: JG[ele][c][spt], detJ[ele][spt],
// fgeo[ele][c][fpt].  Affine meshes (every element's metrics and detJ the
// same at all its solution points, every face's normal and dA the same at all
// its flux points): JG[ele][c], detJ[ele], fgeo[ele][c][face] -- one value
// per element or face (34 instead of 1024 values per hex element at P = 3).
template <int D, int P>
struct GeoAddr {
  int gS, gP, fS;  // (ele, c) stride and point offset of JG/detJ; (ele, c) stride of fgeo
  __device__ __forceinline__ GeoAddr(int affine, int spt) {
    gS = affine ? 1 : ElemGeo<D, P>::NS;
    gP = affine ? 0 : spt;
    fS = affine ? ElemGeo<D, P>::NFACE : ElemGeo<D, P>::NFP;
    aff = affine;
  }
  __device__ __forceinline__ size_t jg(size_t e, int c) const { return (e * D * D + c) * gS + gP; }
  __device__ __forceinline__ size_t det(size_t e) const { return e * gS + gP; }
  __device__ __forceinline__ size_t fg(size_t e, int c, int q) const {
    return (e * (D + 1) + c) * fS + (aff ? q / ElemGeo<D, P>::NFPF : q);
  }
  int aff;
};

// Nodal geometry of linear elements (8-node hexes, 4-node quads): the map
// from reference to physical coordinates written in monomials of the
// reference coordinates r,
//   hex:  x(r) = a + b r0 + c r1 + d r2 + e r0 r1 + f r0 r2 + g r1 r2 + h r0 r1 r2
//   quad: x(r) = a + b r0 + c r1 + e r0 r1
// (the trilinear / bilinear shape functions of funcs.cpp:178-260 expanded).
// geoN[ele][NC] holds b c d e f g h (hex) or b c e (quad), D components each.
// They are combinations of node differences, so the only absolute position,
// a, is not needed and the coefficients keep the precision of the element's
// size.  The stage kernel forms the metrics at its solution points and the
// normals at its flux points from them instead of reading per-point arrays.
template <int D>
struct LinearMap {
  static constexpr int NC = (D == 3) ? 21 : 6;
};

// Jacobian J[dimRef][dimPhys] = dx_dimPhys / dr_dimRef of a linear element at
// reference point r from its map coefficients cf (LinearMap).
template <int D, typename T>
__host__ __device__ __forceinline__ void linearJacobian(const T *cf, const T *r, T J[D][D]) {
  if constexpr (D == 3) {
    const T r01 = r[0] * r[1], r02 = r[0] * r[2], r12 = r[1] * r[2];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      J[0][c] = cf[c] + cf[9 + c] * r[1] + cf[12 + c] * r[2] + cf[18 + c] * r12;
      J[1][c] = cf[3 + c] + cf[9 + c] * r[0] + cf[15 + c] * r[2] + cf[18 + c] * r02;
      J[2][c] = cf[6 + c] + cf[12 + c] * r[0] + cf[15 + c] * r[1] + cf[18 + c] * r01;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      J[0][c] = cf[c] + cf[4 + c] * r[1];
      J[1][c] = cf[2 + c] + cf[4 + c] * r[0];
    }
  }
}

// Adjugate G = det(J) J^-1, G[dimRef][dimPhys] (the reference's JGinv,
// solver.cpp:1273-1281), and det(J) as the return value.
template <int D, typename T>
__host__ __device__ __forceinline__ T adjugate(const T J[D][D], T G[D][D]) {
  if constexpr (D == 2) {
    G[0][0] = J[1][1];
    G[0][1] = -J[1][0];
    G[1][0] = -J[0][1];
    G[1][1] = J[0][0];
    return J[0][0] * J[1][1] - J[0][1] * J[1][0];
  } else {
    const T xr = J[0][0], xs = J[1][0], xt = J[2][0];
    const T yr = J[0][1], ys = J[1][1], yt = J[2][1];
    const T zr = J[0][2], zs = J[1][2], zt = J[2][2];
    G[0][0] = ys * zt - yt * zs; G[0][1] = xt * zs - xs * zt; G[0][2] = xs * yt - xt * ys;
    G[1][0] = yt * zr - yr * zt; G[1][1] = xr * zt - xt * zr; G[1][2] = xt * yr - xr * yt;
    G[2][0] = yr * zs - ys * zr; G[2][1] = xs * zr - xr * zs; G[2][2] = xr * ys - xs * yr;
    return xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
  }
}

template <typename T>
struct StageArgs {
  const T *Uin;
  const T *__restrict__ UfIn;
  T *__restrict__ UfOut;
  const T *__restrict__ JG;
  const T *__restrict__ fgeo;
  const int32_t *__restrict__ fnbr;  // [ele][face] neighbour element
  const int8_t *__restrict__ finfo;  // [ele][face] packed face record
  const T *__restrict__ oned;
  T *divF;
  const double *Ubase;  // FP64 step base state (final stages)
  const T *UbaseT;      // the same in the residual precision (intermediate stages)
  const double *__restrict__ detJ;
  T *Unext;
  double *Uout;
  float *Uout32;
  const T *dprev0, *dprev1, *dprev2;
  const T *rprev;  // previous stage's residual, or null (stage state in Uin)
  double rkPrev;   // its coefficient in the stage state
  const T *__restrict__ dUf;  // viscous: gradient traces [ele][dim][field][fpt]
  const T *__restrict__ Fv;   // viscous: common viscous normal flux [ele][field][fpt] (k_vflux)
  const double *dtE;          // time step: dtE[ele * dtStride] (stride 0: one global dt)
  int dtStride;
  double *waveSp;             // CFL wave speeds [ele][fpt] written at left / boundary points, or null
  double c0, c1, c2, c3, rk;
  int mode, nprev, nEles;
  Phys ph;
  int affine;  // geometry layout (GeoAddr)
  int nodal;   // 1: the stage kernel forms its metrics and normals from geoN (GM = 1)
  const T *__restrict__ geoN;  // linear map coefficients [ele][LinearMap<D>::NC]
  // debug outputs of a residual evaluation (k_stage_dbg only): the
  // transformed flux F_spts [ele][dim][field][spt], the common normal flux
  // times dA Fn_fpts [ele][field][fpt] and the corrected jump
  // disFn_fpts = Fn_fpts - disFn [ele][field][fpt]
  T *dbgF, *dbgFn, *dbgJ;
  int rev;  // walk each XCD's range of element blocks backwards (k_stage)
};

// Face f of the two faces normal to reference axis d (hex: x 2,3; y 4,5;
// z 0,1 -- quad: x 3,1; y 0,2).
template <int D>
__device__ __forceinline__ int dirFace(int d, int which) {
  if (D == 3) return (d == 0) ? 2 + which : (d == 1) ? 4 + which : which;
  return (d == 0) ? (which ? 1 : 3) : (which ? 2 : 0);
}

// Synchronisation of the threads of one element between LDS phases: a
// wavefront-level barrier with wavefront-scope fences when each element's
// threads lie inside one wavefront (TPE divides 64), else the workgroup
// barrier.  Only the phases that read another element's LDS data need the
// workgroup.
template <int TPE>
__device__ __forceinline__ void elementSync() {
  if constexpr (TPE <= 64 && 64 % TPE == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Reference axis normal to face f.
template <int D>
__device__ __forceinline__ int faceAxis(int f) {
  if (D == 3) return (f < 2) ? 2 : (f < 4) ? 0 : 1;
  return (f & 1) ? 0 : 1;
}

// LDG common solution at one flux point (face::ldgSolution, face.cpp:582-607)
// returned as the jump dUc = UC - own trace (face.cpp:159-168,
// intFace.cpp:122-126).  uo / un are this element's and the neighbour's (or
// ghost) traces; right: this element is the face's right element.
template <int D, int NF, typename T>
__device__ __forceinline__ void ldgJump(const Phys &ph, bool bnd, bool right, const T *uo, const T *un, const T *n,
                                        T *dUc) {
  const T pen = bnd ? (T)0 : ldgPenalty<D, T>(ph, n);
#pragma unroll
  for (int k = 0; k < NF; k++) {
    const T UL = right ? un[k] : uo[k], UR = right ? uo[k] : un[k];
    const T UC = bnd ? (T)0.5 * (UL + UR) : (T)0.5 * (UL + UR) - pen * (UL - UR);
    dUc[k] = UC - uo[k];
  }
}

// Hex P3 line operations on the matrix cores.  One element is one wavefront,
// lane t = solution point i + 4j + 16k.  Along each reference axis the
// element has 16 lines of 4 points, and every line operation of the stage is
// a product A B of a small operator A (rows: the derivative matrix Dm, the
// extrapolations l_L and l_R to the line's two faces) and the 4 x 16 matrix
// B of the lines' values:
//  - FP64: v_mfma_f64_4x4x4_4b, four independent 4x4x4 products; block b
//    holds lines 4b .. 4b+3, with A_b[m][k] at lane m + 4b + 16k,
//    B_b[k][n] at lane n + 4b + 16k and C_b[m][n] at lane n + 4b + 16m.
//    The derivative (A = Dm) and the extrapolations (A rows l_L, l_R) are
//    separate products.
//  - FP32: v_mfma_f32_16x16x4f32, A[m][k] at lane m + 16k, B[k][n] at lane
//    n + 16k, C[m][n] at lane n + 16 (m / 4), register m % 4.  Rows 4g of A
//    hold Dm row g (derivative at point g in register 0) and rows 1 + 4s of
//    A hold the extrapolations (register 1), so one product gives both.
// In both, line n = n0 + 4 n1 of the axis takes its point k from lane
// n + 16k and returns its derivative at point g in lane n + 16g: for the z
// axis (n = i + 4j) that is the lane's own point, so B is the lanes' own
// values and the derivative lands in place; for x (n = j + 4k, point i) and
// y (n = i + 4k, point j) the operand comes from the permuted lane (a lane
// permute, or LDS for state already there) and the derivative goes back
// through the inverse permutation.  The extrapolations land in the lanes of
// flux points: z lines put their left face value in lane n and right face
// value in lane 16 + n, x lines in lanes 32 + n and 48 + n (rows 2, 3 of A
// in FP64, rows 9, 13 in FP32) and y lines in lanes n and 16 + n, so that
// one accumulator holds the z and x face values (the 64 flux points of faces
// 0-3) and another the y values (faces 4, 5) -- lane t handles flux point
// mfmaFpt(t, r) in round r of the face phases.  The operator entries are
// those of the VALU line sums (calcDivF_spts, extrapolateNormalFlux,
// extrapolateU); the matrix cores sum the four products in their own order.
template <typename T>
struct MfmaVec;
template <>
struct MfmaVec<double> {
  static __device__ __forceinline__ double mma(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
  }
  // derivative (aD) and face values added to acc (aE) of one axis
  static __device__ __forceinline__ void axis(double aD, double aE, double b, double &acc, double &der) {
    der = mma(aD, b, 0.0);
    acc = mma(aE, b, acc);
  }
  static __device__ __forceinline__ double ext(double aE, double b, double acc) { return mma(aE, b, acc); }
  // the derivative alone
  static __device__ __forceinline__ double der(double aD, double aE, double b) {
    (void)aE;
    return mma(aD, b, 0.0);
  }
};
template <>
struct MfmaVec<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ type mma(float a, float b, float c1) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, type{0.f, c1, 0.f, 0.f}, 0, 0, 0);
  }
  static __device__ __forceinline__ void axis(float aD, float aE, float b, float &acc, float &der) {
    (void)aD;
    const type c = mma(aE, b, acc);
    der = c[0];
    acc = c[1];
  }
  static __device__ __forceinline__ float ext(float aE, float b, float acc) { return mma(aE, b, acc)[1]; }
  static __device__ __forceinline__ float der(float aD, float aE, float b) {
    (void)aD;
    return mma(aE, b, 0.f)[0];
  }
};

// Flux point of lane t in round r (0, 1) of the face phases of a hex P3
// element: round 0 holds faces 0-3 (lanes 16f ..), round 1 faces 4 and 5;
// within a face, the lane of line n has the point fptLine numbers n or, on
// faces 1, 3 and 4 (whose point order runs against the line order), n ^ 3.
__device__ __forceinline__ int mfmaFpt(int t, int r) {
  const int flip = r == 0 ? (t >> 4) & 1 : (t >> 4) == 0;
  return 64 * r + (t ^ (3 * flip));
}

// The lane's part of the products of a hex P3 element (lane t): its A
// entries (the derivative aD; the face values of the z and y axes aEzy and
// of the x axis aEx), the solution point it stands for in the x and y
// operands as an LDS index (px, py) and a lane byte address (bx, by), and
// the lane holding the x / y derivative of its own point (ix, iy).
template <typename T>
struct MfmaLane {
  T aD, aDT, aEzy, aEx;  // aDT: the transposed derivative matrix (FP64 layout)
  int sx, sy, px, py, bx, by, ix, iy;
  // FD: the derivative entries (aD, and the derivative rows of the FP32
  // operands) from the folded derivative OneD::DMF
  __device__ __forceinline__ MfmaLane(const T *s1, int t, bool FD = false) {
    using O = OneD<3>;
    const int k = t >> 4;
    const T lL = s1[O::LL + k], lR = s1[O::LR + k];
    const int dm = FD ? O::DMF : O::DM;
    if constexpr (sizeof(T) == 8) {
      const int m = t & 3;
      aD = s1[dm + m * 4 + k];
      aDT = s1[O::DM + k * 4 + m];
      aEzy = m == 0 ? lL : m == 1 ? lR : (T)0;
      aEx = m == 2 ? lL : m == 3 ? lR : (T)0;
    } else {
      const int m = t & 15;
      const T d = (m & 3) == 0 ? s1[dm + (m >> 2) * 4 + k] : (T)0;
      aD = d;
      aDT = (T)0;
      aEzy = m == 1 ? lL : m == 5 ? lR : d;
      aEx = m == 9 ? lL : m == 13 ? lR : d;
    }
    const int ln = t & 15, lp = t >> 4;
    sx = lp + 4 * (ln & 3) + 16 * (ln >> 2);
    sy = (ln & 3) + 4 * lp + 16 * (ln >> 2);
    px = padIdx<3>(sx);
    py = padIdx<3>(sy);
    bx = 4 * sx;
    by = 4 * sy;
    // point t = i + 4j + 16k is point i of x line j + 4k (lane j + 4k + 16i)
    // and point j of y line i + 4k (lane i + 4k + 16j)
    ix = 4 * ((t >> 2) + 16 * (t & 3));
    iy = 4 * ((t & 3) + 4 * (t >> 4) + 16 * ((t >> 2) & 3));
  }
};

// Value of v in lane addr / 4 of the wavefront (ds_bpermute: the LDS unit's
// lane crossbar, no LDS memory).  Every lane of the wave is active here.
__device__ __forceinline__ float lanePull(int addr, float v) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}
__device__ __forceinline__ double lanePull(int addr, double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Phase (2) on the matrix cores: per reference axis d the contravariant
// flux F~_d = sum_d2 JGinv(d,d2) F_d2, its derivative along d added to div,
// and its values at the two faces normal to d, stored negated as the start
// of the jump (the left face's outward flux is -F~_d): sJ at the round-0
// and round-1 flux points of the lane.  FD: the folded derivative
// (OneD::DMF) alone, no face values (the jump is the common flux alone).
template <int NF, typename T, bool FD = false, typename SM>
__device__ __forceinline__ void mfmaAxes(SM &sm, const int slot, const int t, const T *pJ, const T (&F)[3][NF],
                                         T (&div)[NF], T *dbgF) {
  using V = MfmaVec<T>;
  auto &sJ = sm.sJ;
  const MfmaLane<T> L(sm.s1, t, FD);
  T e0[NF], e1[NF];
#pragma unroll
  for (int k = 0; k < NF; k++) e0[k] = e1[k] = (T)0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
#pragma unroll
    for (int k = 0; k < NF; k++) {
      T ft = (T)0;
#pragma unroll
      for (int d2 = 0; d2 < 3; d2++) ft += pJ[d * 3 + d2] * F[d2][k];
      if (dbgF) dbgF[(d * NF + k) * 64 + t] = ft;
      const T b = d < 2 ? lanePull(d == 0 ? L.bx : L.by, ft) : ft;
      T der;
      if constexpr (FD) der = V::der(L.aD, d == 0 ? L.aEx : L.aEzy, b);
      else V::axis(L.aD, d == 0 ? L.aEx : L.aEzy, b, d == 1 ? e1[k] : e0[k], der);
      const T dv = d < 2 ? lanePull(d == 0 ? L.ix : L.iy, der) : der;
      div[k] = d == 0 ? dv : dv + div[k];
    }
  }
  if constexpr (!FD) {
    const T sg = ((t >> 4) & 1) ? (T)-1 : (T)1;
    const int q0 = mfmaFpt(t, 0), q1 = mfmaFpt(t, 1);
#pragma unroll
    for (int k = 0; k < NF; k++) {
      sJ[slot][k][q0] = sg * e0[k];
      if (t < 32) sJ[slot][k][q1] = sg * e1[k];
    }
  }
  elementSync<64>();
}

// Face values of a state S held in LDS (one element, fields RS apart; padded
// rows when PADV) at the lane's flux point of round r (mfmaFpt): round 0
// (faces 0-3) from the z lines -- the lane's own values own[] -- and the x
// lines, round 1 (faces 4, 5) from the y lines.  All lanes take part.
template <int NF, int RS, typename T, bool PADV = true>
__device__ __forceinline__ void mfmaFaceRound(const T *s1, const T *S, const int t, const T (&own)[NF], const int r,
                                              T (&v)[NF]) {
  using V = MfmaVec<T>;
  const MfmaLane<T> L(s1, t);
  const int ix = PADV ? L.px : L.sx, iy = PADV ? L.py : L.sy;
#pragma unroll
  for (int k = 0; k < NF; k++)
    v[k] = r == 0 ? V::ext(L.aEx, S[k * RS + ix], V::ext(L.aEzy, own[k], (T)0)) : V::ext(L.aEzy, S[k * RS + iy], (T)0);
}

// mfmaFaceRound of values held only in registers (own[], one per lane): the
// x and y operands come from the permuted lanes.
template <int NF, typename T>
__device__ __forceinline__ void mfmaFaceRoundReg(const T *s1, const int t, const T (&own)[NF], const int r,
                                                 T (&v)[NF]) {
  using V = MfmaVec<T>;
  const MfmaLane<T> L(s1, t);
#pragma unroll
  for (int k = 0; k < NF; k++)
    v[k] = r == 0 ? V::ext(L.aEx, lanePull(L.bx, own[k]), V::ext(L.aEzy, own[k], (T)0))
                  : V::ext(L.aEzy, lanePull(L.by, own[k]), (T)0);
}

// correctedGradient (below) for hex P3 with the reference-space derivative on
// the matrix cores: the state S in LDS (fields RS apart, padded rows when
// PADV) and the lane's own values own[], the jumps sCe [NF][NFP] at flux
// points numbered as fptLine numbers them.  Same terms; the matrix cores sum
// the derivative's four products in their own order.  All lanes take part.
template <int NF, int RS, typename T, bool PADV>
__device__ __forceinline__ void mfmaGradient(const T *s1, const T *S, const T (&own)[NF], const T *sCe, const int t,
                                             const T *J, T detJ, T g[3][NF]) {
  using V = MfmaVec<T>;
  using O = OneD<3>;
  constexpr int NFP = ElemGeo<3, 3>::NFP, NFACE = 6;
  const MfmaLane<T> L(s1, t);
  int cq[NFACE], cside[NFACE], cax[NFACE];
  corrEntries<3, 3>(t, cq, cside, cax);
  T gr[3][NF];
#pragma unroll
  for (int d = 0; d < 3; d++) {
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const T b = d == 2 ? own[k] : S[k * RS + (d == 0 ? (PADV ? L.px : L.sx) : (PADV ? L.py : L.sy))];
      const T der = V::der(L.aD, d == 0 ? L.aEx : L.aEzy, b);
      const T dv = d < 2 ? lanePull(d == 0 ? L.ix : L.iy, der) : der;
      T corr = (T)0;
#pragma unroll
      for (int f = 0; f < NFACE; f++)
        if (faceAxis<3>(f) == d)
          corr += (s1[(cside[f] ? O::GR : O::GL) + cax[f]] * (cside[f] ? (T)1 : (T)-1)) * sCe[k * NFP + cq[f]];
      gr[d][k] = corr + dv;
    }
  }
  const T invDet = (T)1. / detJ;
#pragma unroll
  for (int d1 = 0; d1 < 3; d1++)
#pragma unroll
    for (int k = 0; k < NF; k++) {
      T acc = gr[0][k] * J[0 * 3 + d1] + gr[1][k] * J[1 * 3 + d1];
      acc = acc + gr[2][k] * J[2 * 3 + d1];
      g[d1][k] = invDet * acc;
    }
}

// Corrected physical gradient at solution point t (calcGradU_spts,
// solver.cpp:808-821; correctGradU 823-879): the reference-space gradient
// sum_m Dm u along each axis, plus correctU = correction * tNorm applied to the
// flux-point jumps dUc (only the two faces normal to the axis are nonzero, in
// ascending flux-point order), then transformed with JGinv / detJ.
// sUe: [NF][NS] state, sCe: [NF][NFP] jumps, J: JGinv(dimRef, dimPhys).
// PAD: sUe uses the padded LDS rows (ElemGeo::NSP, padIdx).
template <int D, int P, int NF, typename T, bool PAD = false>
__device__ __forceinline__ void correctedGradient(const T *s1, const T *sUe, const T *sCe, int t, int si, int sj,
                                                  int sk, const T *J, T detJ, T g[D][NF]) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  constexpr int NFP = G::NFP, P1 = G::P1, NFACE = G::NFACE;
  constexpr int NS = PAD ? G::NSP : G::NS;
  int cq[NFACE], cside[NFACE], cax[NFACE];
  corrEntries<D, P>(t, cq, cside, cax);
  T gr[D][NF];
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int id = (d == 0) ? si : (d == 1) ? sj : sk;
#pragma unroll
    for (int k = 0; k < NF; k++) {
      T acc = (T)0;
#pragma unroll
      for (int m = 0; m < P1; m++) {
        const int sp = (d == 0) ? m + P1 * (sj + P1 * sk) : (d == 1) ? si + P1 * (m + P1 * sk) : si + P1 * (sj + P1 * m);
        acc += s1[O::DM + id * P1 + m] * sUe[k * NS + (PAD ? padIdx<P>(sp) : sp)];
      }
      T corr = (T)0;
#pragma unroll
      for (int f = 0; f < NFACE; f++)
        if (faceAxis<D>(f) == d)
          corr += (s1[(cside[f] ? O::GR : O::GL) + cax[f]] * (cside[f] ? (T)1 : (T)-1)) * sCe[k * NFP + cq[f]];
      gr[d][k] = corr + acc;
    }
  }
  const T invDet = (T)1. / detJ;
#pragma unroll
  for (int d1 = 0; d1 < D; d1++)
#pragma unroll
    for (int k = 0; k < NF; k++) {
      T acc = gr[0][k] * J[0 * D + d1] + gr[1][k] * J[1 * D + d1];
      if constexpr (D == 3) acc = acc + gr[2][k] * J[2 * D + d1];
      g[d1][k] = invDet * acc;
    }
}

// ---------------------------------------------------------------------------
// Viscous common normal flux at every flux point (face::calcViscousFlux,
// face.cpp:177-287), one thread per (element, flux point), from the state
// traces UfIn and the gradient traces dUf: boundaries take
// boundFace::applyViscousBCs (boundFace.cpp:279-320), advection-diffusion
// boundaries a zero outside gradient.  Right-side points evaluate it in the
// left element's frame, as the inviscid flux.  Fv[ele][field][fpt] (before
// the dA scaling), which the viscous stage kernel adds to its inviscid
// common flux: one thread per point keeps the stage kernel's registers for
// the rest of its work.
// ---------------------------------------------------------------------------
template <int D, int P, int NF, typename T, bool BND>
__global__ __launch_bounds__(256) void k_vflux(const T *__restrict__ Uf, const T *__restrict__ dUf,
                                               const T *__restrict__ fgeo, const int32_t *__restrict__ fnbr,
                                               const int8_t *__restrict__ finfo, int affine, Phys ph, int nEles,
                                               T *__restrict__ Fv) {
  using G = ElemGeo<D, P>;
  constexpr int NFP = G::NFP, NFPF = G::NFPF, NFACE = G::NFACE;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nEles * NFP) return;
  const size_t e = (size_t)(idx / NFP);
  const int q = (int)(idx - (long)e * NFP);
  const GeoAddr<D, P> ga(affine, 0);
  const int fi = finfo[e * NFACE + q / NFPF];
  const bool bnd = BND && fi >= FINFO_BOUNDARY;
  const bool right = !bnd && ((fi >> 5) & 1);
  T n[D], uo[NF], un[NF], go[D][NF], gn[D][NF], F[NF];
#pragma unroll
  for (int c = 0; c < D; c++) n[c] = fgeo[ga.fg(e, c, q)];
#pragma unroll
  for (int k = 0; k < NF; k++) {
    uo[k] = Uf[(e * NF + k) * NFP + q];
    F[k] = (T)0;
  }
  const T *Go = dUf + e * D * NF * NFP + q;
#pragma unroll
  for (int d = 0; d < D; d++)
#pragma unroll
    for (int k = 0; k < NF; k++) go[d][k] = Go[(d * NF + k) * NFP];
  int vb = 0;
  if (bnd) {
    const int bc = fi - FINFO_BOUNDARY;
    vb = (bc == 10) ? 2 : 1;
    boundaryState<D, NF, T>(ph, bc, uo, n, un);
    if constexpr (NF > 1) {
      boundaryGradient<D, NF, T>(ph, bc, uo, un, go, gn);
    } else {
#pragma unroll
      for (int d = 0; d < D; d++) gn[d][0] = (T)0;
    }
  } else {
    const size_t en = (size_t)fnbr[e * NFACE + q / NFPF];
    const int qn = neighbourFpt<D, P>(q, fi);
#pragma unroll
    for (int k = 0; k < NF; k++) un[k] = Uf[(en * NF + k) * NFP + qn];
    const T *Gn = dUf + en * D * NF * NFP + qn;
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int k = 0; k < NF; k++) gn[d][k] = Gn[(d * NF + k) * NFP];
  }
  if (right) ldgViscousFlux<D, NF, T>(ph, 0, un, uo, gn, go, n, F);
  else ldgViscousFlux<D, NF, T>(ph, vb, uo, un, go, gn, n, F);
#pragma unroll
  for (int k = 0; k < NF; k++) Fv[(e * NF + k) * NFP + q] = F[k];
}

// ---------------------------------------------------------------------------
// Gradient traces for the viscous path: LDG common solution at every flux
// point, corrected physical gradient at the solution points, and its
// extrapolation to the flux points (extrapolateGradU, solver.cpp:881-895),
// written to dUf[ele][dim][field][fpt] for the neighbours' viscous face flux.
// ---------------------------------------------------------------------------
template <typename T>
struct GradArgs {
  const T *__restrict__ Uin;
  const T *__restrict__ UfIn;
  const T *__restrict__ JG;
  const double *__restrict__ detJ;
  const T *__restrict__ fgeo;
  const int32_t *__restrict__ fnbr;
  const int8_t *__restrict__ finfo;
  const T *__restrict__ oned;
  T *__restrict__ dUf;
  int nEles;
  Phys ph;
  int affine;  // geometry layout (GeoAddr)
  T *dUs;      // debug output (null in production launches): the gradient at the spts [ele][dim][field][spt]
};

template <int D, int P, int NF, typename T, bool BND>
__global__ __launch_bounds__((ElemGeo<D, P>::NTH)) void k_grad(GradArgs<T> a) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, P1 = G::P1, TPE = G::TPE, EPB = G::EPB;
  constexpr int NFACE = G::NFACE, NQ = (NFP + TPE - 1) / TPE;
  __shared__ T sU[EPB][NF][NS];
  __shared__ T sC[EPB][NF][NFP];
  // the gradient at the solution points for its traces (the matrix-core
  // variant takes them from the lanes' registers instead)
  constexpr bool MG0 = FR_MFMA && D == 3 && P == 3 && sizeof(T) == 8;
  __shared__ T sG[MG0 ? 1 : EPB][D][NF][MG0 ? 1 : NS];
  __shared__ T s1[O::SIZE];
  const int b = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const int slot = threadIdx.x / TPE, t = threadIdx.x - slot * TPE;
  const int e = b * EPB + slot;
  const bool act = e < a.nEles;
  // hex P3 FP64: the state's face values, the gradient's derivative and the
  // gradient traces on the matrix cores (lane t handles flux point
  // mfmaFpt(t, r) in round r)
  constexpr bool MG = FR_MFMA && D == 3 && P == 3 && sizeof(T) == 8;
  const bool hasS = act && t < NS;
  for (int i = threadIdx.x; i < O::SIZE; i += blockDim.x) s1[i] = a.oned[i];
  if (act) {
    const T *Ue = a.Uin + (size_t)e * NF * NS;
    for (int i = t; i < NF * NS; i += TPE) (&sU[slot][0][0])[i] = Ue[i];
  }
  T pJ[D * D];
  T pdet = (T)1;
  const GeoAddr<D, P> ga(a.affine, t);
  if (hasS) {
#pragma unroll
    for (int c = 0; c < D * D; c++) pJ[c] = a.JG[ga.jg(e, c)];
    pdet = (T)a.detJ[ga.det(e)];
  }
  __syncthreads();
  // LDG jump at every flux point
#pragma unroll
  for (int r = 0; r < NQ; r++) {
    const int q = MG ? mfmaFpt(t, r) : t + r * TPE;
    T uo[NF];
    if constexpr (MG) {
      if (act) {
        T own[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) own[k] = sU[slot][k][t];
        mfmaFaceRound<NF, NS, T, false>(s1, &sU[slot][0][0], t, own, r, uo);
      }
    }
    if (act && q < NFP) {
      int base, stride, side, dn;
      fptLine<D, P>(q, base, stride, side, dn);
      const T *l1 = s1 + (side ? O::LR : O::LL);
      T un[NF], n[D], dUc[NF];
      if constexpr (!MG) {
#pragma unroll
        for (int k = 0; k < NF; k++) {
          T acc = (T)0;
#pragma unroll
          for (int m = 0; m < P1; m++) acc += l1[m] * sU[slot][k][base + m * stride];
          uo[k] = acc;
        }
      }
#pragma unroll
      for (int c = 0; c < D; c++) n[c] = a.fgeo[ga.fg(e, c, q)];
      const int fi = a.finfo[(size_t)e * NFACE + q / NFPF];
      const bool bnd = BND && fi >= FINFO_BOUNDARY;
      if (bnd) {
        boundaryState<D, NF, T>(a.ph, fi - FINFO_BOUNDARY, uo, n, un);
      } else {
        const T *Ur = a.UfIn + (size_t)a.fnbr[(size_t)e * NFACE + q / NFPF] * NF * NFP + neighbourFpt<D, P>(q, fi);
#pragma unroll
        for (int k = 0; k < NF; k++) un[k] = Ur[k * NFP];
      }
      ldgJump<D, NF, T>(a.ph, bnd, !bnd && ((fi >> 5) & 1), uo, un, n, dUc);
#pragma unroll
      for (int k = 0; k < NF; k++) sC[slot][k][q] = dUc[k];
    }
  }
  elementSync<TPE>();
  T g[D][NF];
  if (MG && act) {
    if constexpr (MG) {
      T own[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) own[k] = sU[slot][k][t];
      mfmaGradient<NF, NS, T, false>(s1, &sU[slot][0][0], own, &sC[slot][0][0], t, pJ, pdet, g);
      if (a.dUs)
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int k = 0; k < NF; k++) a.dUs[(((size_t)e * D + d) * NF + k) * NS + t] = g[d][k];
    }
  } else if (hasS) {
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
    correctedGradient<D, P, NF, T>(s1, &sU[slot][0][0], &sC[slot][0][0], t, si, sj, sk, pJ, pdet, g);
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int k = 0; k < NF; k++) sG[slot][d][k][t] = g[d][k];
    if (a.dUs)
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int k = 0; k < NF; k++) a.dUs[(((size_t)e * D + d) * NF + k) * NS + t] = g[d][k];
  }
  elementSync<TPE>();
  if (MG && act) {
    if constexpr (MG) {
      T *Ge = a.dUf + (size_t)e * D * NF * NFP;
#pragma unroll
      for (int r = 0; r < NQ; r++) {
        const int q = mfmaFpt(t, r);
#pragma unroll
        for (int d = 0; d < D; d++) {
          T v[NF];
          mfmaFaceRoundReg<NF, T>(s1, t, g[d], r, v);
          if (q < NFP)
#pragma unroll
            for (int k = 0; k < NF; k++) Ge[(d * NF + k) * NFP + q] = v[k];
        }
      }
    }
  } else if (act) {
    T *Ge = a.dUf + (size_t)e * D * NF * NFP;
#pragma unroll 1
    for (int q = t; q < NFP; q += TPE) {
      int base, stride, side, dn;
      fptLine<D, P>(q, base, stride, side, dn);
      const T *l1 = s1 + (side ? O::LR : O::LL);
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int k = 0; k < NF; k++) {
          T acc = (T)0;
#pragma unroll
          for (int m = 0; m < P1; m++) acc += l1[m] * sG[slot][d][k][base + m * stride];
          Ge[(d * NF + k) * NFP + q] = acc;
        }
    }
  }
}

// LDS of one stage-kernel workgroup (tile layout L).
template <int D, int P, int NF, typename T, int L = 0>
struct StageSmem {
  using G = ElemGeo<D, P, L>;
  T sU[G::EPB][NF][G::NSP];   // stage state (padded rows, padIdx)
  T sFd[G::EPB][NF][G::NSP];  // contravariant flux F~_d of one reference axis, later the updated state
  T sJ[G::EPB][NF][G::NFP];   // normal-flux jump Fn*dA - disFn at the flux points
  T s1[OneD<P>::SIZE];        // 1D factor table
};

// BND = false is the variant for meshes without boundary faces (no ghost
// states, which keeps its register allocation small).  VIS adds the viscous
// terms (LDG): the corrected gradient at the solution points is recomputed
// here from the state and the neighbours' traces exactly as k_grad forms it,
// and the viscous common flux reads the gradient traces k_grad wrote.
// FIN = true is the variant for the final stages (modes 2 and 3): the FP64
// base state, and in mode 2 the earlier stages' residuals, are prefetched.
// FIN = false serves modes 0 and 1 with the base in the residual precision.
// The stage computation of element block b (the elements b*EPB ..), shared
// by the per-mesh kernel and the batched persistent kernel.
// GM = 0 reads the metrics, detJ and face geometry from the per-point (or
// compact) arrays; GM = 1 (inviscid linear elements, StageArgs::nodal) forms
// them from the element's map coefficients geoN (LinearMap): the adjugate
// metrics and detJ at each solution point, and at each flux point the outward
// area-weighted normal, whose length is dA; a right-side point uses the
// negated own normal for the left element's (equal but for rounding).
// GM = 2 forms the solution-point metrics and reads the face geometry.
// L is the tile layout (ElemGeo): 1 only in the batched 2D kernels.
template <int D, int P, int NF, typename T, bool BND, bool VIS, bool FIN, int GM = 0, bool DBG = false, int L = 0>
__device__ __forceinline__ void stageBody(const StageArgs<T> &a, const int b, StageSmem<D, P, NF, T, L> &sm,
                                          const int tid) {
  static_assert(GM == 0 || !VIS, "the viscous stage reads the per-point geometry");
  static_assert(L == 0 || (D == 2 && !VIS), "layout 1 is the inviscid 2D batch layout");
  using G = ElemGeo<D, P, L>;
  using O = OneD<P>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, P1 = G::P1, TPE = G::TPE, EPB = G::EPB;
  constexpr int NFACE = G::NFACE;
  static_assert(TPE >= NS && TPE >= 2 * NFPF, "one solution point per thread");
  auto &sU = sm.sU;
  auto &sFd = sm.sFd;
  auto &sJ = sm.sJ;
  auto &s1 = sm.s1;
  // the modes a variant serves: the inviscid FIN variants 2 and 3, the other
  // inviscid variants 0 and 1 (the launches choose so), the viscous one all
  constexpr bool M01 = VIS || !FIN, M23 = VIS || FIN;
  // hex P3: the line contractions of (2) on the matrix cores
  constexpr bool MF = FR_MFMA && D == 3 && P == 3;
  // the face values of the state in (3) and (5) too: FP64 only (three FP32
  // 16x16x4 products per field cost more than the FP32 line sums they replace)
  constexpr bool MFF = MF && sizeof(T) == 8;
  // (2) with the correction of the extrapolated flux folded into the
  // derivative (OneD::DMF): no face values of F~, and the jump of (3) is the
  // common flux alone (the viscous variant's LDG jumps of (0), held in sJ, are
  // read in (1), before (2)'s barriers; the debug instance keeps the jump it
  // reports)
  constexpr bool FOLD = FR_FOLD && !DBG;
  const int slot = tid / TPE, t = tid - slot * TPE;
  const int tP = padIdx<P>(t);  // padded LDS index of this thread's solution point
  // flux point of this thread in round r of the face phases: the matrix-core
  // variant's lane order (mfmaFpt), otherwise t + r * TPE
  auto fq = [&](int r) { return MF ? mfmaFpt(t, r) : t + r * TPE; };
  const int e = b * EPB + slot;
  const bool act = e < a.nEles;
  const bool hasS = act && t < NS;
  // Faces between two elements of this block take the neighbour's trace from
  // its state in LDS (line sums), and their traces are not written to HBM (the
  // element order groups face neighbours into blocks).  The viscous variant
  // keeps the full trace exchange, which k_grad reads, and so does the FP32
  // matrix-core variant (hex P3 FP32 31.97 against 30.38 GDOF/s with it; the
  // FP64 one gains, 20.47 -> 21.02, since its step is bound by its bytes).
  const int blockBase = b * EPB;
  auto inBlock = [&](int en) {
    return FR_INBLOCK && (!MF || MFF) && !VIS && (unsigned)(en - blockBase) < (unsigned)EPB;
  };
  // Loads.  Vector-memory counters retire in issue order, so the face records
  // (which the neighbour-trace addresses depend on) go first, the state (or
  // the residual it is formed from) and the base state, detJ and dt (needed
  // in LDS before the barrier) next, then every other read of the element;
  // only those are waited for before the barrier, the rest stays in flight
  // through the LDS phases.  The loads are
  // branch-free (idle lanes read clamped, in-bounds indices; lanes with no
  // neighbour trace to fetch read the 1D table, which is cached), so that the
  // compiler's wait counts are exact.
  constexpr int NQ = (NFP + TPE - 1) / TPE;  // flux points per thread
  static_assert(O::SIZE <= G::NTH, "one thread per 1D table entry");
  // clamped element (wave-uniform when an element fills whole wavefronts, so
  // that its map coefficients come through scalar loads)
  size_t ec = (size_t)(act ? e : a.nEles - 1);
  if constexpr (TPE >= 64) ec = (size_t)__builtin_amdgcn_readfirstlane((int)ec);
  const int tc = t < NS ? t : NS - 1;  // clamped solution point
  int si, sj, sk;  // tensor index of this thread's solution point
  if (D == 3) {
    sk = tc / (P1 * P1);
    sj = (tc - P1 * P1 * sk) / P1;
    si = tc - P1 * sj - P1 * P1 * sk;
  } else {
    si = tc % P1;
    sj = tc / P1;
    sk = 0;
  }
  int pfi[NQ], pnb[NQ];  // face record, neighbour element (-1: boundary)
#pragma unroll
  for (int r = 0; r < NQ; r++) {
    const int qc = min(fq(r), NFP - 1);
    pfi[r] = a.finfo[ec * NFACE + qc / NFPF];
    pnb[r] = a.fnbr[ec * NFACE + qc / NFPF];
  }
  const T o1 = a.oned[min(tid, O::SIZE - 1)];
  const GeoAddr<D, P> ga(a.affine, tc);
  // Stage state at this thread's solution point: read from Uin, or (rprev
  // set) the previous stage's residual, from which the state is formed below.
  T ul[NF];
  {
    const T *src = (a.rprev ? a.rprev : a.Uin) + ec * NF * NS + tc;
#pragma unroll
    for (int k = 0; k < NF; k++) ul[k] = src[k * NS];
  }
  // stage base state, detJ and time step for the update (read in every mode:
  // in mode 0 the base is the input state, whose lines are already in cache)
  double pb[NF];
  T pJ[D * D];  // adjugate metrics of this solution point
  double pdet;
  if constexpr (GM == 0) {
    pdet = a.detJ[ga.det(ec)];
  } else {
    // metrics from the element's linear map at this solution point
    const T *cf = a.geoN + ec * LinearMap<D>::NC;
    const T r[3] = {a.oned[O::X1 + si], a.oned[O::X1 + sj], a.oned[O::X1 + sk]};
    T Jm[D][D], Gm[D][D];
    linearJacobian<D, T>(cf, r, Jm);
    pdet = (double)adjugate<D, T>(Jm, Gm);
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int d2 = 0; d2 < D; d2++) pJ[d * D + d2] = Gm[d][d2];
  }
  const double pdt = a.dtE[ec * a.dtStride];
#pragma unroll
  for (int k = 0; k < NF; k++)
    pb[k] = (FIN || (VIS && a.mode >= 2)) ? a.Ubase[(ec * NF + k) * NS + tc]
                                          : (double)a.UbaseT[(ec * NF + k) * NS + tc];
  // earlier stages' residuals for the final update, issued with the state:
  // they are folded into the base before the barrier, which frees their
  // registers for the LDS phases (the latest one is already in ul when rprev
  // is set)
  T pd[FIN ? 3 : 1][NF];
  if constexpr (FIN) {
    const T *dp[3] = {a.dprev0, a.dprev1, a.dprev2};
#pragma unroll
    for (int r = 0; r < 3; r++)
      if (a.mode == 2 && r < a.nprev && !(a.rprev && r == a.nprev - 1)) {
#pragma unroll
        for (int k = 0; k < NF; k++) pd[r][k] = __builtin_nontemporal_load(dp[r] + (ec * NF + k) * NS + tc);
      }
  }
  if constexpr (GM == 0) {
    // components c at a uniform stride from c = 0 (one 64-bit add per load)
    const T *jg0 = a.JG + ga.jg(ec, 0);
    const size_t js = (size_t)ga.gS;
#pragma unroll
    for (int c = 0; c < D * D; c++) pJ[c] = __builtin_nontemporal_load(jg0 + c * js);
  }
  T pn[NQ][D + 1], pu[NQ][NF];  // face normal + dA, neighbour trace
  if constexpr (GM != 1) {
#pragma unroll
    for (int r = 0; r < NQ; r++) {
      const int qc = min(fq(r), NFP - 1);
#pragma unroll
      for (int c = 0; c <= D; c++) pn[r][c] = __builtin_nontemporal_load(a.fgeo + ga.fg(ec, c, qc));
    }
  }
  // neighbour traces: prefetched here, or (DEF: the intermediate inviscid
  // variants when FR_DEFER_NBR is set) loaded at the start of the face phase,
  // which shortens their registers' live range to that phase
  auto loadNbr = [&]() {
#pragma unroll
    for (int r = 0; r < NQ; r++) {
      const int qc = min(fq(r), NFP - 1);
      const bool bndFace = BND && pfi[r] >= FINFO_BOUNDARY;
      const bool fetch = !bndFace && !inBlock(pnb[r]);
      const T *Ur = fetch ? a.UfIn + (size_t)pnb[r] * NF * NFP + neighbourFpt<D, P>(qc, pfi[r]) : a.oned;
      const int st = fetch ? NFP : 0;
#pragma unroll
      for (int k = 0; k < NF; k++) pu[r][k] = __builtin_nontemporal_load(Ur + k * st);
      if (bndFace) pnb[r] = -1;
    }
  };
  constexpr bool DEF = FR_DEFER_NBR && !VIS && !FIN;
  if constexpr (!DEF) loadNbr();
  if (tid < O::SIZE) s1[tid] = o1;
  // the FP64 base of the final stages is absolute; FP32 stage states hold the
  // energy less Phys::eOff
  const bool base64 = FIN || (VIS && a.mode >= 2);
  if (hasS) {
    // Us = base - rkPrev * rprev * dt / detJ: the previous stage's update
    // (the expression of step (4)), evaluated here instead of stored
    const double dtJ = pdt / pdet;
#pragma unroll
    for (int k = 0; k < NF; k++) {
      T v = ul[k];
      if (a.rprev) {
        double u = pb[k];
        u = u - a.rkPrev * (double)ul[k] * dtJ;
        if constexpr (hasEnergyOffset<T>())
          if (k == NF - 1 && base64) u = u - a.ph.eOff;
        v = (T)u;
      }
      sU[slot][k][tP] = v;
    }
  }
  if constexpr (FIN) {
    if (a.rprev) {
#pragma unroll
      for (int r = 0; r < 3; r++)
        if (r == a.nprev - 1) {
#pragma unroll
          for (int k = 0; k < NF; k++) pd[r][k] = ul[k];
        }
    }
    // Final update (mode 2): the earlier stages' terms are subtracted from
    // the base now (after the stage state above used it), in stage order as
    // at the update
    if (a.mode == 2) {
      const double dtJ = pdt / pdet;
#pragma unroll
      for (int k = 0; k < NF; k++) {
        double u = pb[k];
        if (a.nprev > 0) u -= a.c0 * (double)pd[0][k] * dtJ;
        if (a.nprev > 1) u -= a.c1 * (double)pd[1][k] * dtJ;
        if (a.nprev > 2) u -= a.c2 * (double)pd[2][k] * dtJ;
        pb[k] = u;
      }
    }
  }
  __syncthreads();

  // (0) viscous: LDG jumps at the flux points (held in sJ until the first
  // barrier of (2), which every read of them precedes)
  if constexpr (VIS) {
#pragma unroll
    for (int r = 0; r < NQ; r++) {
      const int q = fq(r);
      if (act && q < NFP) {
        int base, stride, side, dn;
        fptLine<D, P>(q, base, stride, side, dn);
        const int lb = padIdx<P>(base), ls = padStride<P>(stride);
        const T *l1 = s1 + (side ? O::LR : O::LL);
        T uo[NF], un[NF], dUc[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
          T acc = (T)0;
#pragma unroll
          for (int m = 0; m < P1; m++) acc += l1[m] * sU[slot][k][lb + m * ls];
          uo[k] = acc;
        }
        const int fi = pfi[r];
        const bool bnd = BND && fi >= FINFO_BOUNDARY;
        if (bnd) {
          boundaryState<D, NF, T>(a.ph, fi - FINFO_BOUNDARY, uo, pn[r], un);
        } else {
#pragma unroll
          for (int k = 0; k < NF; k++) un[k] = pu[r][k];
        }
        ldgJump<D, NF, T>(a.ph, bnd, !bnd && ((fi >> 5) & 1), uo, un, pn[r], dUc);
#pragma unroll
        for (int k = 0; k < NF; k++) sJ[slot][k][q] = dUc[k];
      }
    }
    elementSync<TPE>();
  }
  // (1) physical point flux (calcInviscidFlux_spts, solver.cpp:514-539; with
  // calcViscousFlux_spts, solver.cpp:562-596, added for VIS)
  T F[D][NF], div[NF];
  if (hasS) {
    T u[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) u[k] = sU[slot][k][tP];
    inviscidFlux<D, NF, T>(a.ph, u, F);
    if constexpr (VIS) {
      T g[D][NF], Fv[D][NF];
      if constexpr (MFF)
        mfmaGradient<NF, G::NSP, T, true>(s1, &sU[slot][0][0], u, &sJ[slot][0][0], t, pJ, (T)pdet, g);
      else
        correctedGradient<D, P, NF, T, true>(s1, &sU[slot][0][0], &sJ[slot][0][0], t, si, sj, sk, pJ, (T)pdet, g);
      viscousFlux<D, NF, T>(a.ph, u, g, Fv);
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int k = 0; k < NF; k++) F[d][k] = Fv[d][k] + F[d][k];
    }
  }
  // (2) per reference axis d: F~_d = sum_d2 JGinv(d,d2) F_d2, its derivative
  // along d (calcDivF_spts) and its extrapolation to the two faces normal to d
  // (extrapolateNormalFlux), stored negated as the start of the jump
  if constexpr (MF) {
    if (act) mfmaAxes<NF, T, FOLD>(sm, slot, t, pJ, F, div, DBG ? a.dbgF + (size_t)e * D * NF * NS : nullptr);
  } else {
#pragma unroll
    for (int d = 0; d < D; d++) {
      if (hasS) {
#pragma unroll
        for (int k = 0; k < NF; k++) {
          T acc = (T)0;
#pragma unroll
          for (int d2 = 0; d2 < D; d2++) acc += pJ[d * D + d2] * F[d2][k];
          sFd[slot][k][tP] = acc;
          if constexpr (DBG) a.dbgF[(((size_t)e * D + d) * NF + k) * NS + t] = acc;
        }
      }
      elementSync<TPE>();
      if (hasS) {
        const int id = (d == 0) ? si : (d == 1) ? sj : sk;
#pragma unroll
        for (int k = 0; k < NF; k++) {
          T acc = (T)0;
#pragma unroll
          for (int m = 0; m < P1; m++) {
            // LDS index of the line point
            const int sp = padIdx<P>((d == 0) ? m + P1 * (sj + P1 * sk)
                                     : (d == 1) ? si + P1 * (m + P1 * sk)
                                                : si + P1 * (sj + P1 * m));
            acc += s1[(FOLD ? O::DMF : O::DM) + id * P1 + m] * sFd[slot][k][sp];
          }
          div[k] = (d == 0) ? acc : acc + div[k];
        }
      }
      if (!FOLD && act && t < 2 * NFPF) {
        const int q = dirFace<D>(d, t / NFPF) * NFPF + t % NFPF;
        int base, stride, side, dn;
        fptLine<D, P>(q, base, stride, side, dn);
        const int lb = padIdx<P>(base), ls = padStride<P>(stride);
        const T *l1 = s1 + (side ? O::LR : O::LL);
        const T sg = side ? (T)1 : (T)-1;
        // the weights with the face sign and the negation of the stored value
        // applied once, ahead of the fields (exact: sign changes only)
        T lw[P1];
#pragma unroll
        for (int m = 0; m < P1; m++) lw[m] = -(l1[m] * sg);
#pragma unroll
        for (int k = 0; k < NF; k++) {
          T acc = (T)0;
#pragma unroll
          for (int m = 0; m < P1; m++) acc += lw[m] * sFd[slot][k][lb + m * ls];
          sJ[slot][k][q] = acc;
        }
      }
      elementSync<TPE>();
    }
  }

  // (3) common flux at every flux point (face::calcInviscidFlux,
  // face.cpp:132-175; intFace.cpp:85-120; boundFace.cpp:33-277)
  if constexpr (DEF) loadNbr();
#pragma unroll
  for (int r = 0; r < NQ; r++) {
    const int q = fq(r);
    // the state at the flux point
    T uo[NF];
    if constexpr (MFF) {
      if (act) {
        T own[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) own[k] = r == 0 ? sU[slot][k][tP] : (T)0;
        mfmaFaceRound<NF, G::NSP, T>(s1, &sU[slot][0][0], t, own, r, uo);
      }
    }
    if (act && q < NFP) {
      int base, stride, side, dn;
      fptLine<D, P>(q, base, stride, side, dn);
      if constexpr (!MFF) {
        const int lb = padIdx<P>(base), ls = padStride<P>(stride);
        const T *l1 = s1 + (side ? O::LR : O::LL);
#pragma unroll
        for (int k = 0; k < NF; k++) {
          T acc = (T)0;
#pragma unroll
          for (int m = 0; m < P1; m++) acc += l1[m] * sU[slot][k][lb + m * ls];
          uo[k] = acc;
        }
      }
      const int fi = pfi[r];
      const bool bnd = BND && fi >= FINFO_BOUNDARY;
      const bool right = !bnd && ((fi >> 5) & 1);
      T n[D], dA;
      if constexpr (GM != 1) {
#pragma unroll
        for (int c = 0; c < D; c++) n[c] = pn[r][c];
        dA = pn[r][D];
      } else {
        // reference coordinates of the point: -1 / +1 along the face's axis,
        // the line's abscissae across it
        const int bi = base % P1, bj = (base / P1) % P1, bk = base / (P1 * P1);
        const T one = side ? (T)1 : (T)-1;
        const T rq[3] = {dn == 0 ? one : s1[O::X1 + bi], dn == 1 ? one : s1[O::X1 + bj],
                         dn == 2 ? one : s1[O::X1 + bk]};
        T Jm[D][D], Gm[D][D];
        linearJacobian<D, T>(a.geoN + ec * LinearMap<D>::NC, rq, Jm);
        adjugate<D, T>(Jm, Gm);
        // outward normal times dA: row dn of the adjugate, signed by the side;
        // the left element's normal at a right-side point is its negation
        const T sg = (side != (int)right) ? (T)1 : (T)-1;
        T ss = (T)0;
#pragma unroll
        for (int c = 0; c < D; c++) {
          T g = Gm[0][c];
          if constexpr (D >= 2) g = (dn == 1) ? Gm[1][c] : g;
          if constexpr (D == 3) g = (dn == 2) ? Gm[2][c] : g;
          n[c] = sg * g;
          ss += g * g;
        }
        T rdA;
        rootAndInv(ss, dA, rdA);
#pragma unroll
        for (int c = 0; c < D; c++) n[c] = n[c] * rdA;
      }
      T un[NF], Fn[NF];
      if (bnd) {
        boundaryState<D, NF, T>(a.ph, fi - FINFO_BOUNDARY, uo, n, un);
      } else if (inBlock(pnb[r])) {
        // the neighbour's trace from its state in LDS, as it would extrapolate it
        int nbase, nstride, nside, ndn;
        fptLine<D, P>(neighbourFpt<D, P>(q, fi), nbase, nstride, nside, ndn);
        const int npb = padIdx<P>(nbase), nps = padStride<P>(nstride);
        const T *ln = s1 + (nside ? O::LR : O::LL);
        const T *Un = &sU[pnb[r] - blockBase][0][0];
#pragma unroll
        for (int k = 0; k < NF; k++) {
          T acc = (T)0;
#pragma unroll
          for (int m = 0; m < P1; m++) acc += ln[m] * Un[k * G::NSP + npb + m * nps];
          un[k] = acc;
        }
      } else {
#pragma unroll
        for (int k = 0; k < NF; k++) un[k] = pu[r][k];
      }
      // the left element's state first, per lane, so that a wavefront whose
      // points lie on both sides of their faces runs the common flux once,
      // not once per side (right implies !bnd)
      T ws;
      const bool hasWs = commonFluxSided<D, NF, T>(a.ph, bnd, right, uo, un, n, Fn, &ws);
      // the CFL wave speed at the face's left points (and boundary points);
      // the right element's points keep theirs (face.cpp:74-80)
      if (a.waveSp && hasWs && !right) a.waveSp[(size_t)e * NFP + q] = (double)ws;
      if constexpr (VIS) {
        // the viscous common flux of this point (k_vflux, before this launch)
#pragma unroll
        for (int k = 0; k < NF; k++) Fn[k] += a.Fv[((size_t)e * NF + k) * NFP + q];
      }
      const T scale = right ? -dA : dA;
#pragma unroll
      for (int k = 0; k < NF; k++) sJ[slot][k][q] = FOLD ? Fn[k] * scale : Fn[k] * scale + sJ[slot][k][q];
      if constexpr (DBG)
#pragma unroll
        for (int k = 0; k < NF; k++) {
          a.dbgFn[((size_t)e * NF + k) * NFP + q] = Fn[k] * scale;
          a.dbgJ[((size_t)e * NF + k) * NFP + q] = sJ[slot][k][q];
        }
    }
  }
  elementSync<TPE>();

  // (4) correction of the jump (correctDivFlux, solver.cpp:789-806) and the
  // stage update; the updated state goes to sFd (free after (2)) for its
  // traces, so that sU keeps the stage state other elements of the block read
  T utv[NF];  // the updated state at this thread's point, for the traces of (5)
  if (hasS) {
    int cq[NFACE], cside[NFACE], cax[NFACE];
    corrEntries<D, P>(t, cq, cside, cax);
    T cw[NFACE];  // correction weights, read once for all fields
#pragma unroll
    for (int f = 0; f < NFACE; f++) cw[f] = s1[(cside[f] ? O::GR : O::GL) + cax[f]];
    const size_t eo = (size_t)e * NF * NS;
#pragma unroll
    for (int k = 0; k < NF; k++) {
      T corr = (T)0;
#pragma unroll
      for (int f = 0; f < NFACE; f++) corr += cw[f] * sJ[slot][k][cq[f]];
      const T dv = corr + div[k];
      const size_t gi = eo + (size_t)k * NS + t;
      if (a.divF) __builtin_nontemporal_store(dv, a.divF + gi);
      if (M01 && a.mode == 0) continue;
      // stage terms RK * divF / detJ * dt, with dt / detJ formed once per point
      const double dtJ = pdt / pdet;
      double u = pb[k];
      // DEF: the base state read again here instead of kept in registers
      if constexpr (DEF) u = (double)a.UbaseT[gi];
      if (M23 && a.mode == 2) {
        // earlier residuals: already subtracted from pb (FIN, before the barrier) or
        // read here (viscous variant)
        if constexpr (!FIN) {
          const T r0 = (a.nprev > 0) ? a.dprev0[gi] : (T)0;
          const T r1 = (a.nprev > 1) ? a.dprev1[gi] : (T)0;
          const T r2 = (a.nprev > 2) ? a.dprev2[gi] : (T)0;
          if (a.nprev > 0) u -= a.c0 * (double)r0 * dtJ;
          if (a.nprev > 1) u -= a.c1 * (double)r1 * dtJ;
          if (a.nprev > 2) u -= a.c2 * (double)r2 * dtJ;
        }
        const double cl = (a.nprev == 0) ? a.c0 : (a.nprev == 1) ? a.c1 : (a.nprev == 2) ? a.c2 : a.c3;
        u -= cl * (double)dv * dtJ;
      } else {
        u = u - a.rk * (double)dv * dtJ;
      }
      // the updated state in the residual precision, for the traces: FP32
      // states hold the energy less Phys::eOff, and the final stages' FP64
      // base (hence u) is absolute
      double ut = u;
      if constexpr (hasEnergyOffset<T>())
        if (k == NF - 1 && base64) ut = u - a.ph.eOff;
      if (!M23 || (M01 && a.mode == 1)) {
        if (a.Unext) __builtin_nontemporal_store((T)u, a.Unext + gi);
      } else {
        __builtin_nontemporal_store(u, a.Uout + gi);
        if (a.Uout32) __builtin_nontemporal_store((float)ut, a.Uout32 + gi);
      }
      sFd[slot][k][tP] = (T)ut;
      utv[k] = (T)ut;
    }
  }
  if (M01 && a.mode == 0) return;
  elementSync<TPE>();

  // (5) face traces of the updated state for the next stage, for the faces
  // whose neighbour is in another block
  if (act) {
    T *Fe = a.UfOut + (size_t)e * NF * NFP;
#pragma unroll
    for (int r = 0; r < NQ; r++) {
      const int q = fq(r);
      T tv[NF];
      if constexpr (MFF) mfmaFaceRound<NF, G::NSP, T>(s1, &sFd[slot][0][0], t, utv, r, tv);
      if (q >= NFP) continue;
      if (!VIS && (pnb[r] < 0 || inBlock(pnb[r]))) continue;
      if constexpr (MFF) {
#pragma unroll
        for (int k = 0; k < NF; k++) __builtin_nontemporal_store(tv[k], Fe + k * NFP + q);
        continue;
      }
      int base, stride, side, dn;
      fptLine<D, P>(q, base, stride, side, dn);
      const int lb = padIdx<P>(base), ls = padStride<P>(stride);
      const T *l1 = s1 + (side ? O::LR : O::LL);
#pragma unroll
      for (int k = 0; k < NF; k++) {
        T acc = (T)0;
#pragma unroll
        for (int m = 0; m < P1; m++) acc += l1[m] * sFd[slot][k][lb + m * ls];
        __builtin_nontemporal_store(acc, Fe + k * NFP + q);
      }
    }
  }
}

// The inviscid final-stage variant asks for 3 waves per SIMD (at most 168
// VGPRs): it holds the most prefetched data, and at 2 waves per SIMD the
// final stage ran 12 % longer.  So do the hex P3 variants with the MFMA
// contractions, whose accumulators would otherwise take them to 2 waves.
// The viscous variant runs one wave per SIMD, except the FP64 hex P3 one,
// whose matrix-core gradient and line operators fit two without spilling.
template <int D, int P, int NF, typename T, bool BND, bool VIS, bool FIN, int GM = 0>
__global__ __launch_bounds__((ElemGeo<D, P>::NTH),
                             (VIS ? ((FR_MFMA && D == 3 && P == 3 && sizeof(T) == 8) ? 2 : 1)
                              : (FR_MFMA && D == 3 && P == 3) ? (FIN ? FR_MF_FIN_WAVES : FR_MF_WAVES)
                              : FIN                           ? 3
                                                              : FR_STAGE_MINW)) void k_stage(StageArgs<T> a) {
  __shared__ StageSmem<D, P, NF, T> sm;
  // Workgroups are dealt to the 8 XCDs in turn; remap so that each XCD walks a
  // contiguous range of elements and face neighbours share its L2.  The host
  // pads the grid to a multiple of 8.  rev walks the range backwards, so that
  // a launch after a forward one reads first what that one wrote last (still
  // in the memory-side cache).
  const int R = gridDim.x >> 3, j = blockIdx.x >> 3;
  const int b = (blockIdx.x & 7) * R + (a.rev ? R - 1 - j : j);
  stageBody<D, P, NF, T, BND, VIS, FIN, GM>(a, b, sm, (int)threadIdx.x);
}

// The residual evaluation (mode 0) with the debug outputs of StageArgs
// (dbgF, dbgFn, dbgJ) written: the slow path behind fr_get_array, FP64, the
// stored geometry arrays (GM = 0).
template <int D, int P, int NF, bool BND, bool VIS>
__global__ __launch_bounds__((ElemGeo<D, P>::NTH), 1) void k_stage_dbg(StageArgs<double> a) {
  __shared__ StageSmem<D, P, NF, double> sm;
  const int b = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  stageBody<D, P, NF, double, BND, VIS, false, 0, true>(a, b, sm, (int)threadIdx.x);
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
  if constexpr (D == 2) {
    double J2[2][2] = {{J[0][0], J[0][1]}, {J[1][0], J[1][1]}}, G2[2][2];
    det = adjugate<2, double>(J2, G2);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) G[i][j] = G2[i][j];
  } else {
    det = adjugate<3, double>(J, G);
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
                       const int8_t *__restrict__ finfo, int NFP, int NFPF, int nEles, double *__restrict__ fgeo) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nEles * NFP) return;
  const int e = (int)(idx / NFP), q = (int)(idx - (long)e * NFP);
  const int fi = finfo[(size_t)e * (NFP / NFPF) + q / NFPF];  // NFP / NFPF faces per element
  int se = e, sq = q;
  if (fi < FINFO_BOUNDARY && ((fi >> 5) & 1)) {
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
// FP32 copy of one FP64 state entry i of [ele][field][spt]: the energy field
// less the offset eOff (Phys::eOff), the other fields as they are.
__device__ __forceinline__ float stateToFloat(double u, long i, int NF, int NS, double eOff) {
  return (float)(((int)((i / NS) % NF) == NF - 1) ? u - eOff : u);
}

// U = U0 - rk * divF / detJ * dt; with copy != 0 first sets U0 = U
// (copyUspts_U0, solver.cpp:387-392).  U32 (optional) receives a float copy.
template <typename TD>
__global__ void k_rk_a(double *__restrict__ U, double *__restrict__ U0, const TD *__restrict__ divF,
                       const double *__restrict__ detJ, float *__restrict__ U32, double rk,
                       const double *__restrict__ dtE, int dtStride, int copy, int NF, int NS, long n,
                       double eOff) {
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
  const double v = u0 - rk * (double)divF[i] / detJ[e * NS + s] * dtE[e * dtStride];
  U[i] = v;
  if (U32) U32[i] = stateToFloat(v, i, NF, NS, eOff);
}

// Single timeStepB pass (solver.cpp:347-385) for the ABI entry point.
template <typename TD>
__global__ void k_rk_b1(double *__restrict__ U, const TD *__restrict__ dF, const double *__restrict__ detJ,
                        float *__restrict__ U32, double b, const double *__restrict__ dtE, int dtStride, int NF,
                        int NS, long n, double eOff) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long e = i / ((long)NF * NS);
  const int s = (int)(i % NS);
  const double u = U[i] - b * (double)dF[i] / detJ[e * NS + s] * dtE[e * dtStride];
  U[i] = u;
  if (U32) U32[i] = stateToFloat(u, i, NF, NS, eOff);
}

// FP32 copy of the FP64 state (stateToFloat).
static __global__ void k_state_to_float(const double *__restrict__ U, float *__restrict__ U32, int NF, int NS,
                                        double eOff, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) U32[i] = stateToFloat(U[i], i, NF, NS, eOff);
}

// CFL time step of each element from the current state, FP64: the traces
// (extrapolateU), the wave speed |v.n| + c (Euler) or |a.n| (advection) at
// every flux point (ele::calcWaveSpFpts, ele.cpp:1031-1084, static grids) and
// dt = CFL * limit(P) * 2 * vol / sum(waveSp * wSurf * dA) (ele::calcDt,
// ele.cpp:1086-1105).  One thread per element; runs only when asked for.
// waveSp[ele][fpt] keeps the wave speeds (the reference's waveSp_fpts) and
// dtNum[ele] the numerator CFL * limit(P) * 2 * vol for k_dt_ws.
template <int D, int P, int NF>
__global__ __launch_bounds__(256) void k_dt(const double *__restrict__ U, const double *__restrict__ ownN,
                                            const double *__restrict__ fgeo, const double *__restrict__ detJ,
                                            const double *__restrict__ wts, const double *__restrict__ wsurf,
                                            const double *__restrict__ oned, double factor, Phys ph, int nEles,
                                            double *__restrict__ dtE, double *__restrict__ waveSp,
                                            double *__restrict__ dtNum) {
  using G = ElemGeo<D, P>;
  using O = OneD<P>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, P1 = G::P1;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEles) return;
  double vol = 0;
  for (int s = 0; s < NS; s++) vol += detJ[(size_t)e * NS + s] * wts[s];
  const double *Ue = U + (size_t)e * NF * NS;
  double intWave = 0;
  for (int q = 0; q < NFP; q++) {
    double n[D];
    for (int c = 0; c < D; c++) n[c] = ownN[((size_t)e * D + c) * NFP + q];
    double ws;
    if constexpr (NF == 1) {
      double vN = ph.advect[0] * n[0] + ph.advect[1] * n[1];
      if (D == 3) vN += ph.advect[2] * n[2];
      ws = fabs(vN);
    } else {
      int base, stride, side, dn;
      fptLine<D, P>(q, base, stride, side, dn);
      const double *l1 = oned + (side ? O::LR : O::LL);
      double u[NF];
      for (int k = 0; k < NF; k++) {
        double acc = 0;
        for (int m = 0; m < P1; m++) acc += l1[m] * Ue[k * NS + base + m * stride];
        u[k] = acc;
      }
      const double rho = u[0], vx = u[1] / rho, vy = u[2] / rho, vz = (D == 3) ? u[3] / rho : 0.;
      const double rhoVSq = rho * (vx * vx + vy * vy + vz * vz);
      const double pr = (ph.gamma - 1) * (u[D + 1] - 0.5 * rhoVSq);
      double vN = vx * n[0] + vy * n[1];
      if (D == 3) vN += vz * n[2];
      const double csq = fmax(ph.gamma * pr / rho, 0.0);
      ws = fabs(vN) + sqrt(csq);
    }
    waveSp[(size_t)e * NFP + q] = ws;
    intWave += ws * wsurf[q % NFPF] * fgeo[((size_t)e * (D + 1) + D) * NFP + q];
  }
  dtE[e] = factor * vol / intWave;
  dtNum[e] = factor * vol;
}

// The CFL time step of every element from the stored wave speeds
// (ele::calcDt, ele.cpp:1086-1105, with the waveSp_fpts the face fluxes of
// the last residual left: solver::calcDt after the first iteration,
// solver.cpp:293-305), and the minimum over each block of elements into
// part[block].  One thread per element.
template <int NFP, int NFPF>
__global__ __launch_bounds__(256) void k_dt_ws(const double *__restrict__ waveSp, const double *__restrict__ fgeo,
                                               const double *__restrict__ wsurf, const double *__restrict__ dtNum,
                                               int D, int nEles, double *__restrict__ dtE, double *__restrict__ part) {
  __shared__ double red[256];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double dt = INFINITY;
  if (e < nEles) {
    const double *ws = waveSp + (size_t)e * NFP;
    const double *dA = fgeo + ((size_t)e * (D + 1) + D) * NFP;
    double intWave = 0;
    for (int q = 0; q < NFP; q++) intWave += ws[q] * wsurf[q % NFPF] * dA[q];
    dt = dtNum[e] / intWave;
    dtE[e] = dt;
  }
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// part[block] = the minimum of x over each block of 256 entries.
static __global__ __launch_bounds__(256) void k_block_min(const double *__restrict__ x, int n,
                                                          double *__restrict__ part) {
  __shared__ double red[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  red[threadIdx.x] = i < n ? x[i] : INFINITY;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dtG[0] = the minimum of part[0..n) (params->dt of solver::calcDt), one block.
static __global__ __launch_bounds__(256) void k_dt_min(const double *__restrict__ part, int n,
                                                       double *__restrict__ dtG) {
  __shared__ double red[256];
  double m = INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmin(m, part[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) dtG[0] = red[0];
}

// Affine-element test and compact geometry (GeoAddr): element e passes when
// its metrics and detJ agree with those of its first solution point, and each
// face's normal and dA with those of the face's first flux point, to `tol`
// relative to the element's largest metric entry, 1 (unit normals) and the
// first value (detJ, dA).  The compact arrays take the first point's values.
template <int D>
__global__ void k_affine(const double *__restrict__ JG, const double *__restrict__ detJ,
                         const double *__restrict__ fgeo, int NS, int NFP, int NFPF, int nFace, int nEles,
                         double tol, double *__restrict__ JGc, double *__restrict__ detJc,
                         double *__restrict__ fgeoc, int8_t *__restrict__ ok) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEles) return;
  bool good = true;
  double sc = 0.0;
  for (int c = 0; c < D * D; c++) sc = fmax(sc, fabs(JG[(e * D * D + c) * NS]));
  for (int c = 0; c < D * D; c++) {
    const double *g = JG + (e * D * D + c) * NS;
    JGc[e * D * D + c] = g[0];
    for (int p = 1; p < NS; p++) good = good && fabs(g[p] - g[0]) <= tol * sc;
  }
  const double *dj = detJ + e * NS;
  detJc[e] = dj[0];
  for (int p = 1; p < NS; p++) good = good && fabs(dj[p] - dj[0]) <= tol * fabs(dj[0]);
  for (int f = 0; f < nFace; f++)
    for (int c = 0; c <= D; c++) {
      const double *v = fgeo + (e * (D + 1) + c) * NFP + f * NFPF;
      fgeoc[(e * (D + 1) + c) * nFace + f] = v[0];
      const double s = (c < D) ? 1.0 : fabs(v[0]);
      for (int j = 1; j < NFPF; j++) good = good && fabs(v[j] - v[0]) <= tol * s;
    }
  ok[e] = good ? 1 : 0;
}

static __global__ void k_to_float(const double *__restrict__ a, float *__restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------
// Reductions (FP64).  One thread per element accumulates in the reference's
// loop order, then a block tree reduction; partial results per block go to
// `part` and are summed on the host in block order (deterministic).
// ---------------------------------------------------------------------------
// Maximum that keeps a NaN operand (fmax drops it): an infinity norm over a
// residual holding a NaN is NaN, so the NaN check of the reduction fires.
__host__ __device__ __forceinline__ double maxKeepNaN(double a, double b) { return (a != a || a > b) ? a : b; }

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
        else r[k] = maxKeepNaN(fabs(v) / J, r[k]);
      }
    }
  }
  for (int k = 0; k < 5; k++) red[threadIdx.x][k] = r[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 5; k++) {
        const double o = red[threadIdx.x + w][k];
        red[threadIdx.x][k] = (normType == 3) ? maxKeepNaN(red[threadIdx.x][k], o) : red[threadIdx.x][k] + o;
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
