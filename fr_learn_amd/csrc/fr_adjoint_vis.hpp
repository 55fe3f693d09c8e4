// fr_adjoint_vis.hpp -- launches of the viscous (LDG) adjoint passes: the
// forward gradient traces, k_vjp_state_vis and k_vjp_ldg.  Instantiated per
// (nDims, nFields) in fr_adjoint_v2.hip, fr_adjoint_v3.hip and
// fr_adjoint_v3ad.hip, which compile in parallel (the dual-number kernels
// are the slowest part of the build).
#pragma once

#include "fr_adjoint.hpp"
#include "fr_ctx.hpp"

namespace frimpl {

template <int D, int NF>
hipError_t vjpViscousLaunch(fr_ctx *c, const double *Us, const double *lam, double *out, bool bnd);

template <int D, int P, int NF>
hipError_t vjpViscousLaunchP(fr_ctx *c, const double *Us, const double *lam, double *out, bool bnd) {
  using G = frdev::ElemGeo<D, P>;
  const unsigned nbg = (unsigned)(((c->nE + G::EPB - 1) / G::EPB + 7) / 8 * 8);
  frdev::GradArgs<double> ga{Us, c->adjTr, c->JG64, c->detJ, c->fgeo64, c->fnbr, c->finfo, c->oned64, c->adjQf,
                             c->nE, c->ph, /*affine=*/0};
  const unsigned nbq = (unsigned)(((long)c->nE * G::NFP + 255) / 256);
  if (bnd) {
    hipLaunchKernelGGL((frdev::k_grad<D, P, NF, double, true>), dim3(nbg), dim3(G::NTH), 0, c->stream, ga);
    hipLaunchKernelGGL((frdev::k_vjp_face_vis<D, P, NF, true>), dim3(nbq), dim3(256), 0, c->stream, lam, c->adjTr,
                       c->adjQf, c->adjA, c->fgeo64, c->fnbr, c->finfo, c->oned64, c->ph, c->nE, c->adjUo,
                       c->adjQo, c->adjCj);
    hipLaunchKernelGGL((frdev::k_vjp_state_vis<D, P, NF, true>), dim3(c->nE), dim3(G::TPE), 0, c->stream, Us, lam,
                       c->adjUo, c->adjQo, c->adjCj, c->JG64, c->detJ, c->finfo, c->oned64, c->ph, c->nE, out,
                       c->adjB);
    hipLaunchKernelGGL((frdev::k_vjp_ldg<D, P, NF, true>), dim3(c->nE), dim3(G::TPE), 0, c->stream, Us, c->adjB,
                       c->fgeo64, c->fnbr, c->finfo, c->oned64, c->ph, c->nE, out);
  } else {
    hipLaunchKernelGGL((frdev::k_grad<D, P, NF, double, false>), dim3(nbg), dim3(G::NTH), 0, c->stream, ga);
    hipLaunchKernelGGL((frdev::k_vjp_face_vis<D, P, NF, false>), dim3(nbq), dim3(256), 0, c->stream, lam, c->adjTr,
                       c->adjQf, c->adjA, c->fgeo64, c->fnbr, c->finfo, c->oned64, c->ph, c->nE, c->adjUo,
                       c->adjQo, c->adjCj);
    hipLaunchKernelGGL((frdev::k_vjp_state_vis<D, P, NF, false>), dim3(c->nE), dim3(G::TPE), 0, c->stream, Us, lam,
                       c->adjUo, c->adjQo, c->adjCj, c->JG64, c->detJ, c->finfo, c->oned64, c->ph, c->nE, out,
                       c->adjB);
    hipLaunchKernelGGL((frdev::k_vjp_ldg<D, P, NF, false>), dim3(c->nE), dim3(G::TPE), 0, c->stream, Us, c->adjB,
                       c->fgeo64, c->fnbr, c->finfo, c->oned64, c->ph, c->nE, out);
  }
  return hipGetLastError();
}

#define FR_VJP_VIS_INSTANCE(DD, NFF)                                                                    \
  template <>                                                                                           \
  hipError_t vjpViscousLaunch<DD, NFF>(fr_ctx * c, const double *Us, const double *lam, double *out,    \
                                       bool bnd) {                                                      \
    switch (c->P) {                                                                                     \
      case 1: return vjpViscousLaunchP<DD, 1, NFF>(c, Us, lam, out, bnd);                               \
      case 2: return vjpViscousLaunchP<DD, 2, NFF>(c, Us, lam, out, bnd);                               \
      case 3: return vjpViscousLaunchP<DD, 3, NFF>(c, Us, lam, out, bnd);                               \
      case 4: return vjpViscousLaunchP<DD, 4, NFF>(c, Us, lam, out, bnd);                               \
      case 5: return vjpViscousLaunchP<DD, 5, NFF>(c, Us, lam, out, bnd);                               \
      default: return hipErrorInvalidValue;                                                             \
    }                                                                                                   \
  }

template <>
hipError_t vjpViscousLaunch<2, 4>(fr_ctx *c, const double *Us, const double *lam, double *out, bool bnd);
template <>
hipError_t vjpViscousLaunch<2, 1>(fr_ctx *c, const double *Us, const double *lam, double *out, bool bnd);
template <>
hipError_t vjpViscousLaunch<3, 5>(fr_ctx *c, const double *Us, const double *lam, double *out, bool bnd);
template <>
hipError_t vjpViscousLaunch<3, 1>(fr_ctx *c, const double *Us, const double *lam, double *out, bool bnd);

}  // namespace frimpl
