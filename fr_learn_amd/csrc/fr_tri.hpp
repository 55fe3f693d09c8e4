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
// the new state.  In FP64 the operator products run on the matrix cores over
// all (element, field) columns of the workgroup (triStageMF), in FP32 as each
// thread's line sums (triStageBody).  The stage arguments and modes are those
// of k_stage; every face's traces go through HBM (no in-block exchange).
#pragma once

#include <type_traits>

#include "fr_batch.hpp"
#include "fr_kernels.hpp"

namespace frdev {

// LAY: tile layout, 0 = triTPE(P) threads per element (every per-mesh kernel),
// 1 = max(solution points, flux points) (a choice of the batched kernels'
// tuner).  The workgroup is 256 threads, of which NTH = EPB * TPE hold
// elements.
template <int P, int LAY = 0>
struct TriGeo {
  static constexpr int P1 = P + 1, NS = P1 * (P + 2) / 2, NFPF = P1, NFACE = 3, NFP = 3 * P1;
  static constexpr int TPE = triTPEL(P, LAY), EPB = 256 / TPE, NTH = EPB * TPE;
  static constexpr int DR = 0, DS = NS * NS, E = 2 * NS * NS, L = 2 * NS * NS + NFP * NS;
  static constexpr int TAB = 2 * NS * NS + 2 * NFP * NS;  // Dr | Ds | E | L
  // past TAB: Dr - L tNorm_r E and Ds - L tNorm_s E (the jump folded into the
  // divergence, fr_tri_ops.cpp), for the FP64 kernel
  static constexpr int DRF = TAB, DSF = TAB + NS * NS;
  static_assert(NS <= TPE && NFP <= TPE, "one point per thread");
};

// FP64 runs the operator products on the matrix cores (triStageMF), FP32 as
// line sums of each thread's row (triStageBody).
template <typename T>
constexpr bool triMfma() {
  return sizeof(T) == 8;
}

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

// The loads of one thread of a triangle's stage (element ec): at its
// solution point tc the metrics, detJ and base state, the element's time
// step; at its flux point qc the face record, normal, dA and the neighbour's
// trace (zero at boundaries).  The geometry comes from the per-point arrays,
// or (a.affine) from the compact ones: metrics and detJ per element, normal
// and dA per edge (a linear triangle's are constant).
template <int P, int NF, typename T, bool BND, bool FIN>
struct TriPoint {
  T pJ[4];
  double pdet, pdt, pb[NF];
  int fi;
  bool bnd;
  T n[2], dA, un[NF];
  // final stage: the earlier stages' residuals at tc, issued with the other
  // loads so that the update at the end of the stage does not wait on them
  T pd[FIN ? 3 : 1][NF];
  __device__ __forceinline__ TriPoint(const StageArgs<T> &a, const size_t ec, const int tc, const int qc) {
    using G = TriGeo<P>;
    constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, NFACE = G::NFACE;
    const int gS = a.affine ? 1 : NS, gP = a.affine ? 0 : tc;
#pragma unroll
    for (int c = 0; c < 4; c++) pJ[c] = a.JG[(ec * 4 + c) * gS + gP];
    pdet = a.detJ[ec * gS + gP];
    pdt = a.dtE[ec * a.dtStride];
#pragma unroll
    for (int k = 0; k < NF; k++)
      pb[k] = FIN ? a.Ubase[(ec * NF + k) * NS + tc] : (double)a.UbaseT[(ec * NF + k) * NS + tc];
    if constexpr (FIN) {
      const T *dp[3] = {a.dprev0, a.dprev1, a.dprev2};
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < NF; k++)
          pd[r][k] = (a.mode == 2 && r < a.nprev) ? dp[r][(ec * NF + k) * NS + tc] : (T)0;
    }
    fi = a.finfo[ec * NFACE + qc / NFPF];
    const int nbe = a.fnbr[ec * NFACE + qc / NFPF];
    bnd = BND && fi >= FINFO_BOUNDARY;
    const int fS = a.affine ? NFACE : NFP, fP = a.affine ? qc / NFPF : qc;
    n[0] = a.fgeo[(ec * 3 + 0) * fS + fP];
    n[1] = a.fgeo[(ec * 3 + 1) * fS + fP];
    dA = a.fgeo[(ec * 3 + 2) * fS + fP];
    if (!bnd && nbe >= 0) {
      const T *Ur = a.UfIn + (size_t)nbe * NF * NFP + neighbourFpt<2, P>(qc, fi);
#pragma unroll
      for (int k = 0; k < NF; k++) un[k] = Ur[k * NFP];
    } else {
#pragma unroll
      for (int k = 0; k < NF; k++) un[k] = (T)0;
    }
  }
  // contravariant flux F~_d at the solution point from the state u
  __device__ __forceinline__ void refFlux(const Phys &ph, const T (&u)[NF], T (&Fr)[2][NF]) const {
    T F[2][NF];
    inviscidFlux<2, NF, T>(ph, u, F);
#pragma unroll
    for (int d = 0; d < 2; d++)
#pragma unroll
      for (int k = 0; k < NF; k++) Fr[d][k] = pJ[d * 2 + 0] * F[0][k] + pJ[d * 2 + 1] * F[1][k];
  }
  // common normal flux times dA at the flux point from the own trace uo
  __device__ __forceinline__ void normalFlux(const Phys &ph, const T (&uo)[NF], T (&Fn)[NF]) const {
    T ub[NF], Fc[NF];
    const bool right = !bnd && ((fi >> 5) & 1);
    if (bnd) {
      boundaryState<2, NF, T>(ph, fi - FINFO_BOUNDARY, uo, n, ub);
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
    commonFlux<2, NF, T>(ph, bnd, sL, sR, n, Fc);
    const T scale = right ? -dA : dA;
#pragma unroll
    for (int k = 0; k < NF; k++) Fn[k] = Fc[k] * scale;
  }
  // the residual dv of field k at the solution point (global index gi): stored
  // in divF, and (modes 1-3) the stage update, returned and stored
  __device__ __forceinline__ T update(const StageArgs<T> &a, const int k, const size_t gi, const T dv) const {
    if (a.divF) a.divF[gi] = dv;
    if (a.mode == 0) return dv;
    const double dtJ = pdt / pdet;
    double u = pb[k];
    if (a.mode == 2) {
      // gi is this thread's point tc of field k (the prefetched residuals')
      const T r0 = FIN ? pd[0][k] : (a.nprev > 0) ? a.dprev0[gi] : (T)0;
      const T r1 = FIN ? pd[FIN ? 1 : 0][k] : (a.nprev > 1) ? a.dprev1[gi] : (T)0;
      const T r2 = FIN ? pd[FIN ? 2 : 0][k] : (a.nprev > 2) ? a.dprev2[gi] : (T)0;
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
    return (T)u;
  }
};

// One RK stage of element block b (EPB elements) with the arguments a, the
// operator products as each thread's line sums (FP32).  With FR_FOLD the
// divergence operators are the folded ones (TriGeo::DRF, DSF, copied into the
// table's Dr and Ds places) and the jump is the common flux alone.
template <int P, int NF, typename T, bool BND, bool FIN, int LAY = 0>
__device__ __forceinline__ void triStageBody(const StageArgs<T> &a, const int b, TriSmem<P, NF, T, LAY> &sm,
                                             const int tid) {
  using G = TriGeo<P, LAY>;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, TPE = G::TPE, EPB = G::EPB;
  const int slot = tid / TPE, t = tid - slot * TPE;
  const int e = b * EPB + slot;
  const bool act = slot < EPB && e < a.nEles;
  const bool hasS = act && t < NS, hasQ = act && t < NFP;
  const size_t ec = (size_t)(act ? e : a.nEles - 1);
  const int tc = t < NS ? t : NS - 1, qc = t < NFP ? t : NFP - 1;

  for (int i = tid; i < G::TAB; i += 256) sm.tab[i] = a.oned[(FR_FOLD && i < 2 * NS * NS) ? G::DRF + i : i];
  if (act)
    for (int i = t; i < NF * NS; i += TPE) (&sm.sU[slot][0][0])[i] = a.Uin[ec * NF * NS + i];
  const TriPoint<P, NF, T, BND, FIN> pt(a, ec, tc, qc);
  __syncthreads();

  const T *Dr = sm.tab + G::DR, *Ds = sm.tab + G::DS, *Em = sm.tab + G::E, *Lm = sm.tab + G::L;
  // (1) flux at the solution point, common flux at the flux point
  if (hasS) {
    T u[NF], F[2][NF];
#pragma unroll
    for (int k = 0; k < NF; k++) u[k] = sm.sU[slot][k][t];
    pt.refFlux(a.ph, u, F);
#pragma unroll
    for (int d = 0; d < 2; d++)
#pragma unroll
      for (int k = 0; k < NF; k++) sm.sF[slot][d][k][t] = F[d][k];
  }
  T Fn[NF];
  if (hasQ) {
    T uo[NF];
    // operator entries read once for all fields
#pragma unroll
    for (int k = 0; k < NF; k++) uo[k] = (T)0;
#pragma unroll 3
    for (int s = 0; s < NS; s++) {
      const T w = Em[t * NS + s];
#pragma unroll
      for (int k = 0; k < NF; k++) uo[k] += w * sm.sU[slot][k][s];
    }
    pt.normalFlux(a.ph, uo, Fn);
  }
  elementSync<TPE>();
  // (2) jump at the flux point: Fn dA - (E F~) . tNorm, or Fn dA alone with
  // the folded divergence operators
  if (FR_FOLD && hasQ) {
#pragma unroll
    for (int k = 0; k < NF; k++) sm.sJ[slot][k][t] = Fn[k];
  } else if (hasQ) {
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
    for (int k = 0; k < NF; k++) sm.sN[slot][k][t] = pt.update(a, k, eo + (size_t)k * NS + t, dvs[k] + cor[k]);
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

#ifndef FR_TRI_PADROWS
#define FR_TRI_PADROWS 1  // odd row-tile counts above 3 rounded up (balanced product items)
#endif

// Row tiles of 4 of a product with NR rows: odd counts above 3 are rounded up
// (FR_TRI_PADROWS) so that work items of 2 or 3 tiles divide them: the P5
// flux-point products get items of 3 independent accumulators, the same for
// every wavefront, instead of 1 (triangle P5 stage 137.7 -> 133.3 us).
constexpr int triRowTiles(int NR) {
  const int t = (NR + 3) / 4;
  return (FR_TRI_PADROWS && t > 3 && (t & 1)) ? t + 1 : t;
}

// Entries read from an NR x K operator table (row stride K) by the products:
// rows up to the last tile of 4, k up to the last step of 4.
constexpr int triOpSize(int NR, int K) { return (4 * triRowTiles(NR) - 1) * K + 4 * ((K + 3) / 4); }

// LDS of the FP64 triangle stage, whose operator products run on the matrix
// cores (triStageMF).  The data are stored zero-padded, the point dimension to
// a multiple of 4 and the columns (element slot, field) to a multiple of 16,
// so that every operand of the products is read at a constant offset from the
// lane's base, without predication.  The padding of the point dimension stays
// zero (written once per workgroup; the stage writes only the points
// themselves), so the operator tables need none along k: what a product reads
// past a row's end (the next row, or the zero tail) meets zero data.  Rows past
// the last give results that are not stored.  The data rows (columns of the
// products) have one more entry, an odd stride, so that the 16 columns one
// instruction reads fall in different LDS banks.
template <int P, int NF, int LAY>
struct TriSmemMF {
  using G = TriGeo<P, LAY>;
  static constexpr int NS = G::NS, NFP = G::NFP;
  static constexpr int NS4 = (NS + 3) & ~3, NQ4 = (NFP + 3) & ~3;
  static constexpr int SS = NS4 + 1, SQ = NQ4 + 1;     // data row strides
  static constexpr int NCP = (G::EPB * NF + 15) & ~15;  // columns
  double opE[triOpSize(NFP, NS)];  // extrapolation E (flux points x solution points)
  double opDr[triOpSize(NS, NS)];  // Dr - L tNorm_r E (the jump folded in, TriGeo::DRF)
  double opDs[triOpSize(NS, NS)];  // Ds - L tNorm_s E
  double opL[triOpSize(NS, NFP)];  // lifting L (solution points x flux points)
  double sU[NCP][SS];              // stage state of column (slot, field); later the divergence + correction
  double sF[2][NCP][SS];           // reference flux F~_r, F~_s; later the updated state in [0]
  double sJ[NCP][SQ];              // traces E U, then Fn dA; later the updated state's traces
};

// Row tiles per work item of a product with nrt tiles: a divisor, so that the
// items need no guards.
constexpr int triRpi(int nrt) { return nrt % 3 == 0 ? 3 : nrt % 2 == 0 ? 2 : 1; }

// One segment of a product on the matrix cores: the operator at op (row
// stride AST) against the data at dat (column stride BST) over KS steps of 4
// along k (a multiple of 4 at or past the segment's length).
template <int AST, int BST, int KS>
struct TriSeg {
  const double *op, *dat;
};

// acc_j += A(rows of tile rt0 + j) B(columns of group cg) over the segment,
// FP64 4x4x4 on the matrix cores (four 4x4 blocks per instruction).  Lane l:
// A row 4 rt + (l & 3), k 4 ks + (l >> 4) (the same in every block); B column
// 16 cg + (l & 15), k 4 ks + (l >> 4).  Every operand is at a constant offset
// from the lane's base.  rt0 is uniform per wavefront.
template <int NRT, int RPI, int AST, int BST, int KS>
__device__ __forceinline__ void triAcc(const TriSeg<AST, BST, KS> &sg, const int rt0, const int cg, const int lane,
                                       double (&acc)[RPI]) {
  const double *ap = sg.op + (4 * rt0 + (lane & 3)) * AST + (lane >> 4);
  const double *bp = sg.dat + (16 * cg + (lane & 15)) * BST + (lane >> 4);
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    const double bv = bp[4 * ks];
#pragma unroll
    for (int j = 0; j < RPI; j++) acc[j] = __builtin_amdgcn_mfma_f64_4x4x4f64(ap[4 * j * AST + 4 * ks], bv, acc[j], 0, 0, 0);
  }
}

// A product of the workgroup on the matrix cores: for every group of 16
// columns (element slot, field) and tile of 4 rows, the sum over the segments
// of A B, handed to fin(cg, rt, acc) (lane l holds column 16 cg + (l & 15),
// row 4 rt + (l >> 4)).  A work item is a column group and RPI row tiles (of
// NRT), taken by the NW wavefronts in turn; every lane of a wavefront takes
// part (the instructions read all 64 lanes' operands).
template <int NRT, int NCG, int NW, typename Fin, typename... Segs>
__device__ __forceinline__ void triProduct(const int wave, const int lane, Fin fin, const Segs &...segs) {
  constexpr int RPI = triRpi(NRT), NRG = NRT / RPI;
  for (int item = wave; item < NCG * NRG; item += NW) {
    const int cg = item / NRG, rt0 = (item - cg * NRG) * RPI;
    double acc[RPI];
#pragma unroll
    for (int j = 0; j < RPI; j++) acc[j] = 0.0;
    (triAcc<NRT, RPI>(segs, rt0, cg, lane, acc), ...);
#pragma unroll
    for (int j = 0; j < RPI; j++) fin(cg, rt0 + j, acc[j]);
  }
}

// One RK stage of element block b, FP64: the phases of triStageBody with the
// operator products on the matrix cores over the block's (slot, field)
// columns -- the divergence plus correction with the jump folded in,
// [Dr - L tNorm_r E | Ds - L tNorm_s E | L] [F~_r; F~_s; Fn dA] (no product
// for (E F~) . tNorm), and the new traces E U' (the element's own traces of
// this stage come from the trace buffer, as its neighbours') -- and the
// pointwise work (fluxes, common flux, update) one point per thread.
template <int P, int NF, bool BND, bool FIN, int LAY>
__device__ __forceinline__ void triStageMF(const StageArgs<double> &a, const int b, TriSmemMF<P, NF, LAY> &sm,
                                           const int tid) {
  using G = TriGeo<P, LAY>;
  using S = TriSmemMF<P, NF, LAY>;
  using T = double;
  constexpr int NS = G::NS, NFP = G::NFP, NFPF = G::NFPF, TPE = G::TPE, EPB = G::EPB;
  constexpr int NS4 = S::NS4, NQ4 = S::NQ4, SS = S::SS, SQ = S::SQ, NCG = S::NCP / 16;
  // the workgroup is 256 threads; those past EPB * TPE idle but for the products
  const int slot = tid / TPE, t = tid - slot * TPE;
  const int e = b * EPB + slot;
  const bool act = slot < EPB && e < a.nEles;
  const bool hasS = act && t < NS, hasQ = act && t < NFP;
  const size_t ec = (size_t)(act ? e : a.nEles - 1);
  const int tc = t < NS ? t : NS - 1, qc = t < NFP ? t : NFP - 1;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int lk = lane >> 4, lc = lane & 15;
  if (act)
    for (int i = t; i < NF * NS; i += TPE) {
      const int k = i / NS;
      sm.sU[slot * NF + k][i - k * NS] = a.Uin[ec * NF * NS + i];
    }
  const TriPoint<P, NF, T, BND, FIN> pt(a, ec, tc, qc);
  // this element's own traces from the trace buffer, as the neighbours' (every
  // face's traces go through HBM, written by the previous stage from the
  // product E U' of (5), or by the trace kernel) instead of a product E U here:
  // triangle P5 stage 114 -> 103 us, for 576 B more per element
  T uoG[NF];
#pragma unroll
  for (int k = 0; k < NF; k++) uoG[k] = a.UfIn[(ec * NF + k) * NFP + qc];
  __syncthreads();

  // (1) flux at the solution point
  if (hasS) {
    T u[NF], F[2][NF];
#pragma unroll
    for (int k = 0; k < NF; k++) u[k] = sm.sU[slot * NF + k][t];
    pt.refFlux(a.ph, u, F);
#pragma unroll
    for (int d = 0; d < 2; d++)
#pragma unroll
      for (int k = 0; k < NF; k++) sm.sF[d][slot * NF + k][t] = F[d][k];
  }
  // (2) common flux at the flux point, Fn dA into sJ
  if (hasQ) {
    T Fn[NF];
    pt.normalFlux(a.ph, uoG, Fn);
#pragma unroll
    for (int k = 0; k < NF; k++) sm.sJ[slot * NF + k][t] = Fn[k];
  }
  __syncthreads();
  // (3) divergence + correction with the jump folded in:
  // [Dr - L tNorm_r E | Ds - L tNorm_s E | L] [F~_r; F~_s; Fn dA] into sU
  triProduct<triRowTiles(NS), NCG, 4>(
      wave, lane,
      [&](int cg, int rt, double v) {
        if (4 * rt + lk < NS) sm.sU[16 * cg + lc][4 * rt + lk] = v;
      },
      TriSeg<NS, SS, NS4 / 4>{sm.opDr, &sm.sF[0][0][0]}, TriSeg<NS, SS, NS4 / 4>{sm.opDs, &sm.sF[1][0][0]},
      TriSeg<NFP, SQ, NQ4 / 4>{sm.opL, &sm.sJ[0][0]});
  __syncthreads();
  // (4) stage update, the new state into sF[0]
  if (hasS) {
    const size_t eo = ec * NF * NS;
#pragma unroll
    for (int k = 0; k < NF; k++)
      sm.sF[0][slot * NF + k][t] = pt.update(a, k, eo + (size_t)k * NS + t, sm.sU[slot * NF + k][t]);
  }
  if (a.mode == 0) return;
  __syncthreads();
  // (5) traces of the updated state E U'
  triProduct<triRowTiles(NFP), NCG, 4>(
      wave, lane,
      [&](int cg, int rt, double v) {
        if (4 * rt + lk < NFP) sm.sJ[16 * cg + lc][4 * rt + lk] = v;
      },
      TriSeg<NS, SS, NS4 / 4>{sm.opE, &sm.sF[0][0][0]});
  __syncthreads();
  if (hasQ) {
    T *Fe = a.UfOut + ec * NF * NFP;
#pragma unroll
    for (int k = 0; k < NF; k++) Fe[k * NFP + t] = sm.sJ[slot * NF + k][t];
  }
}

// The operator tables of TriSmemMF from the table Dr | Ds | E | L (zero
// tails), and zero data arrays (their point padding stays zero).
template <int P, int NF, int LAY>
__device__ __forceinline__ void triSetupMF(const double *tab, TriSmemMF<P, NF, LAY> &sm, const int tid) {
  using G = TriGeo<P, LAY>;
  using S = TriSmemMF<P, NF, LAY>;
  constexpr int NS = G::NS, NFP = G::NFP;
  for (int i = tid; i < triOpSize(NFP, NS); i += 256) sm.opE[i] = i < NFP * NS ? tab[G::E + i] : 0.0;
  for (int i = tid; i < triOpSize(NS, NS); i += 256) {
    sm.opDr[i] = i < NS * NS ? tab[G::DRF + i] : 0.0;
    sm.opDs[i] = i < NS * NS ? tab[G::DSF + i] : 0.0;
  }
  for (int i = tid; i < triOpSize(NS, NFP); i += 256) sm.opL[i] = i < NS * NFP ? tab[G::L + i] : 0.0;
  for (int i = tid; i < S::NCP * S::SS; i += 256) {
    (&sm.sU[0][0])[i] = 0.0;
    (&sm.sF[0][0][0])[i] = 0.0;
    (&sm.sF[1][0][0])[i] = 0.0;
  }
  for (int i = tid; i < S::NCP * S::SQ; i += 256) (&sm.sJ[0][0])[i] = 0.0;
  __syncthreads();
}

// The stage's LDS for T: FP64 the matrix-core layout.
template <int P, int NF, typename T, int LAY>
using TriSmemFor = typename std::conditional<triMfma<T>(), TriSmemMF<P, NF, LAY>, TriSmem<P, NF, T, LAY>>::type;

// Four waves per SIMD: 128 registers (the FP64 kernels need no more).
template <int P, int NF, typename T, bool BND, bool FIN>
__global__ __launch_bounds__(256, 4) void k_tri_stage(StageArgs<T> a) {
  __shared__ TriSmemFor<P, NF, T, 0> sm;
  const int b = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);  // XCD-contiguous blocks
  if constexpr (triMfma<T>()) {
    triSetupMF<P, NF, 0>(a.oned, sm, (int)threadIdx.x);
    triStageMF<P, NF, BND, FIN, 0>(a, b, sm, (int)threadIdx.x);
  } else {
    triStageBody<P, NF, T, BND, FIN>(a, b, sm, (int)threadIdx.x);
  }
}

// Batched: a persistent workgroup of 256 threads walks the triangle meshes'
// element blocks of order P (see k_batch_stage; rev walks them backwards),
// with each mesh's arguments.
template <int P, int NF, typename T, bool FIN, int LAY>
__global__ __launch_bounds__(256, 4) void k_batch_tri_stage(const StageArgs<T> *__restrict__ args,
                                                         const BatchItem *__restrict__ items, int nItems,
                                                         int rev) {
  __shared__ TriSmemFor<P, NF, T, LAY> sm;
  // the members share the order's operator table
  if constexpr (triMfma<T>())
    if (blockIdx.x < nItems) triSetupMF<P, NF, LAY>(args[items[blockIdx.x].inst].oned, sm, (int)threadIdx.x);
  for (int it = blockIdx.x; it < nItems; it += gridDim.x) {
    const BatchItem w = items[rev ? nItems - 1 - it : it];
    if constexpr (triMfma<T>()) triStageMF<P, NF, true, FIN, LAY>(args[w.inst], w.block, sm, itemThreadIdx());
    else triStageBody<P, NF, T, true, FIN, LAY>(args[w.inst], w.block, sm, itemThreadIdx());
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
