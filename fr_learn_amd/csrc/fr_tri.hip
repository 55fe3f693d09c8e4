// fr_tri.hip -- launches of the triangle kernels (fr_tri.hpp).
#include "fr_ctx.hpp"
#include "fr_tri.hpp"

namespace frimpl {

namespace {

// f(integral_constant<int, P>{}, integral_constant<int, NF>{}) for the
// context's order (1..5) and field count (1 advection, 4 Euler).
template <class F>
bool triDispatch(int P, int NF, F &&f) {
  auto withNF = [&](auto Pc) {
    if (NF == 1) f(Pc, std::integral_constant<int, 1>{});
    else if (NF == 4) f(Pc, std::integral_constant<int, 4>{});
    else return false;
    return true;
  };
  switch (P) {
    case 1: return withNF(std::integral_constant<int, 1>{});
    case 2: return withNF(std::integral_constant<int, 2>{});
    case 3: return withNF(std::integral_constant<int, 3>{});
    case 4: return withNF(std::integral_constant<int, 4>{});
    case 5: return withNF(std::integral_constant<int, 5>{});
  }
  return false;
}

}  // namespace

template <typename T>
int launchTriStage(fr_ctx *c, const frdev::StageArgs<T> &a, int mode) {
  const bool bnd = c->conn.nIntFaces < (int)c->conn.faces.size();
  hipError_t err = hipSuccess;
  bool ok = triDispatch(c->P, c->NF, [&](auto Pc, auto NFc) {
    constexpr int P = decltype(Pc)::value, NF = decltype(NFc)::value;
    using G = frdev::TriGeo<P>;
    const unsigned nb = (unsigned)(((c->nE + G::EPB - 1) / G::EPB + 7) / 8 * 8);
    TimeScope ts(c, 0);
    if (mode >= 2) {
      if (bnd) hipLaunchKernelGGL((frdev::k_tri_stage<P, NF, T, true, true>), dim3(nb), dim3(256), 0, c->stream, a);
      else hipLaunchKernelGGL((frdev::k_tri_stage<P, NF, T, false, true>), dim3(nb), dim3(256), 0, c->stream, a);
    } else {
      if (bnd) hipLaunchKernelGGL((frdev::k_tri_stage<P, NF, T, true, false>), dim3(nb), dim3(256), 0, c->stream, a);
      else hipLaunchKernelGGL((frdev::k_tri_stage<P, NF, T, false, false>), dim3(nb), dim3(256), 0, c->stream, a);
    }
    err = hipGetLastError();
  });
  if (!ok) return fail(FR_EINVAL, "no triangle kernel for this (order, nFields)");
  if (err != hipSuccess) return fail(FR_EHIP, std::string("triangle stage launch: ") + hipGetErrorString(err));
  return 0;
}

template <typename T>
int launchTriTrace(fr_ctx *c, const T *U, T *Uf, const T *tab) {
  hipError_t err = hipSuccess;
  bool ok = triDispatch(c->P, c->NF, [&](auto Pc, auto NFc) {
    constexpr int P = decltype(Pc)::value, NF = decltype(NFc)::value;
    TimeScope ts(c, 1);
    hipLaunchKernelGGL((frdev::k_tri_trace<P, NF, T>), dim3(nblk((long)c->nE * frdev::TriGeo<P>::NFP)), dim3(256),
                       0, c->stream, U, Uf, tab, c->nE);
    err = hipGetLastError();
  });
  if (!ok) return fail(FR_EINVAL, "no triangle kernel for this (order, nFields)");
  if (err != hipSuccess) return fail(FR_EHIP, std::string("triangle trace launch: ") + hipGetErrorString(err));
  return 0;
}

int launchTriDt(fr_ctx *c, const double *wsurf, double factor, hipStream_t st) {
  hipError_t err = hipSuccess;
  bool ok = triDispatch(c->P, c->NF, [&](auto Pc, auto NFc) {
    constexpr int P = decltype(Pc)::value, NF = decltype(NFc)::value;
    hipLaunchKernelGGL((frdev::k_tri_dt<P, NF>), dim3(nblk(c->nE)), dim3(256), 0, st, c->U, c->ownN,
                       c->fgeo64, c->detJ, c->wts, wsurf, c->oned64, factor, c->ph, c->nE, c->dtE);
    err = hipGetLastError();
  });
  if (!ok) return fail(FR_EINVAL, "no triangle kernel for this (order, nFields)");
  if (err != hipSuccess) return fail(FR_EHIP, std::string("triangle dt launch: ") + hipGetErrorString(err));
  return 0;
}

template int launchTriStage<double>(fr_ctx *, const frdev::StageArgs<double> &, int);
template int launchTriStage<float>(fr_ctx *, const frdev::StageArgs<float> &, int);
template int launchTriTrace<double>(fr_ctx *, const double *, double *, const double *);
template int launchTriTrace<float>(fr_ctx *, const float *, float *, const float *);

}  // namespace frimpl
