// fr_stage_impl.hpp -- launches of the stage kernel (and, on the viscous path,
// the gradient kernel before it); instantiated per precision in
// fr_stage64.hip and fr_stage32.hip so that the two compile in parallel.
#pragma once

#include "fr_ctx.hpp"

namespace frimpl {

// Launch of the stage kernel variant for `mode` with arguments `a` (and the
// gradient kernel before it on the viscous path), quads (DD = 2) or hexes
// (DD = 3); the four (precision, DD) instances compile in their own
// translation units (fr_stage{64,32}_{2,3}d.hip).
template <typename T, int DD>
int launchStageArgsD(fr_ctx *c, const frdev::StageArgs<T> &a, int mode) {
  const bool bnd = c->conn.nIntFaces < (int)c->conn.faces.size();
  const bool vis = c->cfg.viscous != 0;
  hipError_t err = hipSuccess;
  bool ok = dispatch(c->D, c->P, c->NF, [&](auto Dc, auto Pc, auto NFc) {
    constexpr int D = decltype(Dc)::value, P = decltype(Pc)::value, NF = decltype(NFc)::value;
    if constexpr (D != DD) {
      err = hipErrorInvalidValue;
      return;
    } else {
    using G = frdev::ElemGeo<D, P>;
    // grid padded to a multiple of the 8 XCDs (the kernels remap block ids)
    const unsigned nb = (unsigned)(((c->nE + G::EPB - 1) / G::EPB + 7) / 8 * 8);
    if (vis) {
      frdev::GradArgs<T> g{a.Uin,  a.UfIn,  a.JG,   a.detJ,  a.fgeo, a.fnbr,
                            a.finfo, a.oned, (T *)c->dUf, a.nEles, a.ph, a.affine, nullptr};
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
      // the inviscid variants: boundary faces or not, final stage or not,
      // geometry from the stored arrays or from the nodal map (GM)
      auto launch = [&](auto Bc, auto Fc, auto Gc) {
        hipLaunchKernelGGL((frdev::k_stage<D, P, NF, T, decltype(Bc)::value, false, decltype(Fc)::value,
                                           decltype(Gc)::value>),
                           dim3(nb), dim3(G::NTH), 0, c->stream, a);
      };
      auto byGeo = [&](auto Bc, auto Fc) {
        if (a.nodal == 1) launch(Bc, Fc, std::integral_constant<int, 1>{});
        else if (a.nodal == 2) launch(Bc, Fc, std::integral_constant<int, 2>{});
        else launch(Bc, Fc, std::integral_constant<int, 0>{});
      };
      auto byFin = [&](auto Bc) {
        if (mode >= 2) byGeo(Bc, std::true_type{});
        else byGeo(Bc, std::false_type{});
      };
      if (bnd) byFin(std::true_type{});
      else byFin(std::false_type{});
    }
    err = hipGetLastError();
    }
  });
  if (!ok) return fail(FR_EINVAL, "no kernel instance for this (nDims, order, nFields)");
  if (err != hipSuccess) return fail(FR_EHIP, std::string("stage launch: ") + hipGetErrorString(err));
  return 0;
}

template <typename T>
int launchStageArgs(fr_ctx *c, const frdev::StageArgs<T> &a, int mode) {
  if (c->tri) return launchTriStage<T>(c, a, mode);
  return c->D == 3 ? launchStageArgsD<T, 3>(c, a, mode) : launchStageArgsD<T, 2>(c, a, mode);
}

}  // namespace frimpl
