// fr_adjoint.hip -- launches of the adjoint kernels (fr_adjoint.hpp).
#include "fr_adjoint.hpp"
#include "fr_adjoint_vis.hpp"
#include "fr_ctx.hpp"

namespace frimpl {

// out = (d divF / dU)^T lam at the state Us (device arrays, device layout)
int vjpResidualDev(fr_ctx *c, const double *Us, const double *lam, double *out) {
  const bool bnd = c->conn.nIntFaces < (int)c->conn.faces.size();
  hipError_t err = hipSuccess;
  bool ok = dispatch(c->D, c->P, c->NF, [&](auto Dc, auto Pc, auto NFc) {
    constexpr int D = decltype(Dc)::value, P = decltype(Pc)::value, NF = decltype(NFc)::value;
    using G = frdev::ElemGeo<D, P>;
    const long nq = (long)c->nE * G::NFP;
    hipLaunchKernelGGL((frdev::k_trace<D, P, NF, double>), dim3(nblk(nq)), dim3(256), 0, c->stream, Us, c->adjTr,
                       c->oned64, c->nE);
    hipLaunchKernelGGL((frdev::k_vjp_faces<D, P, NF>), dim3(nblk(nq)), dim3(256), 0, c->stream, lam, c->fgeo64,
                       c->finfo, c->oned64, c->nE, c->adjA);
    if (c->cfg.viscous) {
      err = hipGetLastError();
      return;
    }
    const unsigned nb = (unsigned)((c->nE + G::EPB - 1) / G::EPB);
    if (bnd)
      hipLaunchKernelGGL((frdev::k_vjp_state<D, P, NF, true>), dim3(nb), dim3(G::NTH), 0, c->stream, Us, lam,
                         c->adjTr, c->adjA, c->JG64, c->fgeo64, c->fnbr, c->finfo, c->oned64, c->ph, c->nE, out);
    else
      hipLaunchKernelGGL((frdev::k_vjp_state<D, P, NF, false>), dim3(nb), dim3(G::NTH), 0, c->stream, Us, lam,
                         c->adjTr, c->adjA, c->JG64, c->fgeo64, c->fnbr, c->finfo, c->oned64, c->ph, c->nE, out);
    err = hipGetLastError();
  });
  if (!ok) return fail(FR_EINVAL, "no kernel instance for this (nDims, order, nFields)");
  if (err != hipSuccess) return fail(FR_EHIP, std::string("adjoint launch: ") + hipGetErrorString(err));
  if (c->cfg.viscous) {
    if (c->D == 2)
      err = (c->NF == 1) ? vjpViscousLaunch<2, 1>(c, Us, lam, out, bnd) : vjpViscousLaunch<2, 4>(c, Us, lam, out, bnd);
    else
      err = (c->NF == 1) ? vjpViscousLaunch<3, 1>(c, Us, lam, out, bnd) : vjpViscousLaunch<3, 5>(c, Us, lam, out, bnd);
    if (err != hipSuccess) return fail(FR_EHIP, std::string("viscous adjoint launch: ") + hipGetErrorString(err));
  }
  return 0;
}

}  // namespace frimpl
