// fr_stage_dbg.hip -- the debug residual evaluation behind fr_get_array: the
// stage kernel with its debug outputs (k_stage_dbg), after the gradient and
// viscous-flux kernels on the viscous path, FP64 quads and hexes.  Its own
// translation unit, so that these instances compile beside the production
// ones.
#include "fr_stage_impl.hpp"

namespace frimpl {

int launchStageDbg(fr_ctx *c, const frdev::StageArgs<double> &a, double *dUs) {
  if (c->tri || c->prec != FR_PREC_FP64) return fail(FR_EINVAL, "debug residual: FP64 quads and hexes only");
  const bool bnd = c->conn.nIntFaces < (int)c->conn.faces.size();
  const bool vis = c->cfg.viscous != 0;
  hipError_t err = hipSuccess;
  bool ok = dispatch(c->D, c->P, c->NF, [&](auto Dc, auto Pc, auto NFc) {
    constexpr int D = decltype(Dc)::value, P = decltype(Pc)::value, NF = decltype(NFc)::value;
    using G = frdev::ElemGeo<D, P>;
    const unsigned nb = (unsigned)(((c->nE + G::EPB - 1) / G::EPB + 7) / 8 * 8);
    if (vis) {
      frdev::GradArgs<double> g{a.Uin,  a.UfIn,  a.JG,   a.detJ,  a.fgeo, a.fnbr,
                                a.finfo, a.oned, (double *)c->dUf, a.nEles, a.ph, a.affine, dUs};
      if (bnd) hipLaunchKernelGGL((frdev::k_grad<D, P, NF, double, true>), dim3(nb), dim3(G::NTH), 0, c->stream, g);
      else hipLaunchKernelGGL((frdev::k_grad<D, P, NF, double, false>), dim3(nb), dim3(G::NTH), 0, c->stream, g);
      const unsigned nq = (unsigned)(((long)c->nE * G::NFP + 255) / 256);
      if (bnd)
        hipLaunchKernelGGL((frdev::k_vflux<D, P, NF, double, true>), dim3(nq), dim3(256), 0, c->stream, a.UfIn,
                           (const double *)c->dUf, a.fgeo, a.fnbr, a.finfo, a.affine, a.ph, a.nEles,
                           (double *)a.Fv);
      else
        hipLaunchKernelGGL((frdev::k_vflux<D, P, NF, double, false>), dim3(nq), dim3(256), 0, c->stream, a.UfIn,
                           (const double *)c->dUf, a.fgeo, a.fnbr, a.finfo, a.affine, a.ph, a.nEles,
                           (double *)a.Fv);
      if ((err = hipGetLastError()) != hipSuccess) return;
      if (bnd) hipLaunchKernelGGL((frdev::k_stage_dbg<D, P, NF, true, true>), dim3(nb), dim3(G::NTH), 0, c->stream, a);
      else hipLaunchKernelGGL((frdev::k_stage_dbg<D, P, NF, false, true>), dim3(nb), dim3(G::NTH), 0, c->stream, a);
    } else {
      if (bnd) hipLaunchKernelGGL((frdev::k_stage_dbg<D, P, NF, true, false>), dim3(nb), dim3(G::NTH), 0, c->stream, a);
      else hipLaunchKernelGGL((frdev::k_stage_dbg<D, P, NF, false, false>), dim3(nb), dim3(G::NTH), 0, c->stream, a);
    }
    err = hipGetLastError();
  });
  if (!ok) return fail(FR_EINVAL, "no kernel instance for this (nDims, order, nFields)");
  if (err != hipSuccess) return fail(FR_EHIP, std::string("debug stage launch: ") + hipGetErrorString(err));
  return 0;
}

}  // namespace frimpl
