// fr_stage_impl.hpp -- launches of the stage kernel (and, on the viscous path,
// the gradient kernel before it); instantiated per precision in
// fr_stage64.hip and fr_stage32.hip so that the two compile in parallel.
#pragma once

#include "fr_ctx.hpp"

namespace frimpl {

// Launch of the stage kernel variant for `mode` with arguments `a` (and the
// gradient kernel before it on the viscous path).
template <typename T>
int launchStageArgs(fr_ctx *c, const frdev::StageArgs<T> &a, int mode) {
  if (c->tri) return launchTriStage<T>(c, a, mode);
  const bool bnd = c->conn.nIntFaces < (int)c->conn.faces.size();
  const bool vis = c->cfg.viscous != 0;
  hipError_t err = hipSuccess;
  bool ok = dispatch(c->D, c->P, c->NF, [&](auto Dc, auto Pc, auto NFc) {
    constexpr int D = decltype(Dc)::value, P = decltype(Pc)::value, NF = decltype(NFc)::value;
    using G = frdev::ElemGeo<D, P>;
    // grid padded to a multiple of the 8 XCDs (the kernels remap block ids)
    const unsigned nb = (unsigned)(((c->nE + G::EPB - 1) / G::EPB + 7) / 8 * 8);
    if (vis) {
      frdev::GradArgs<T> g{a.Uin,  a.UfIn,  a.JG,   a.detJ,  a.fgeo, a.fnbr,
                            a.finfo, a.oned, (T *)c->dUf, a.nEles, a.ph, a.affine};
      {
        TimeScope ts(c, 3);
        if (bnd) hipLaunchKernelGGL((frdev::k_grad<D, P, NF, T, true>), dim3(nb), dim3(G::NTH), 0, c->stream, g);
        else hipLaunchKernelGGL((frdev::k_grad<D, P, NF, T, false>), dim3(nb), dim3(G::NTH), 0, c->stream, g);
        const unsigned nq = (unsigned)(((long)c->nE * G::NFP + 255) / 256);
        if (bnd)
          hipLaunchKernelGGL((frdev::k_vflux<D, P, NF, T, true>), dim3(nq), dim3(256), 0, c->stream, a.UfIn,
                             (const T *)c->dUf, a.fgeo, a.fnbr, a.finfo, a.affine, a.ph, a.nEles, (T *)a.Fv);
        else
          hipLaunchKernelGGL((frdev::k_vflux<D, P, NF, T, false>), dim3(nq), dim3(256), 0, c->stream, a.UfIn,
                             (const T *)c->dUf, a.fgeo, a.fnbr, a.finfo, a.affine, a.ph, a.nEles, (T *)a.Fv);
      }
      if ((err = hipGetLastError()) != hipSuccess) return;
      TimeScope ts(c, 0);
      if (bnd) hipLaunchKernelGGL((frdev::k_stage<D, P, NF, T, true, true, false>), dim3(nb), dim3(G::NTH), 0,
                                  c->stream, a);
      else hipLaunchKernelGGL((frdev::k_stage<D, P, NF, T, false, true, false>), dim3(nb), dim3(G::NTH), 0,
                              c->stream, a);
    } else {
      TimeScope ts(c, 0);
      if (mode >= 2) {
        if (bnd) hipLaunchKernelGGL((frdev::k_stage<D, P, NF, T, true, false, true>), dim3(nb), dim3(G::NTH), 0,
                                    c->stream, a);
        else hipLaunchKernelGGL((frdev::k_stage<D, P, NF, T, false, false, true>), dim3(nb), dim3(G::NTH), 0,
                                c->stream, a);
      } else {
        if (bnd) hipLaunchKernelGGL((frdev::k_stage<D, P, NF, T, true, false, false>), dim3(nb), dim3(G::NTH), 0,
                                    c->stream, a);
        else hipLaunchKernelGGL((frdev::k_stage<D, P, NF, T, false, false, false>), dim3(nb), dim3(G::NTH), 0,
                                c->stream, a);
      }
    }
    err = hipGetLastError();
  });
  if (!ok) return fail(FR_EINVAL, "no kernel instance for this (nDims, order, nFields)");
  if (err != hipSuccess) return fail(FR_EHIP, std::string("stage launch: ") + hipGetErrorString(err));
  return 0;
}

}  // namespace frimpl
