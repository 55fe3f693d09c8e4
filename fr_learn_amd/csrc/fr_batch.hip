// fr_batch.hip -- the batched persistent stage launches (kernels in fr_batch.hpp).
#include "fr_ctx.hpp"
#include "fr_tri.hpp"

namespace frimpl {

// Stage s of the meshes of group gOnly (0: every group, one after the
// other); parity < 0: the residual-only tuning launch.
template <typename T>
int batchLaunch(fr_batch *b, int parity, int s, int grid, hipStream_t st, int gOnly) {
  const size_t n = b->ctx.size();
  const frdev::StageArgs<T> *args = (parity < 0) ? (const frdev::StageArgs<T> *)b->argsTune
                                                 : (const frdev::StageArgs<T> *)b->args[parity] + (size_t)s * n;
  int mode = 0;
  double rk;
  if (parity >= 0) stageMode(b->ctx[0], s, mode, rk);
  const bool fin = mode >= 2;
  for (int grp = 10; grp >= 1; grp--) {
    const int cnt = b->itemCount[grp];
    if (!cnt || (gOnly && grp != gOnly)) continue;
    const frdev::BatchItem *it = b->items + b->itemStart[grp];
    const unsigned g = (unsigned)std::min(grid, cnt);
    if (grp > 5) {
      const int P = grp - 5;
      auto launch = [&](auto Pc, auto NFc) {
        constexpr int PP = decltype(Pc)::value, NF = decltype(NFc)::value;
        if (fin)
          hipLaunchKernelGGL((frdev::k_batch_tri_stage<PP, NF, T, true>), dim3(g), dim3(256), 0, st, args, it, cnt);
        else
          hipLaunchKernelGGL((frdev::k_batch_tri_stage<PP, NF, T, false>), dim3(g), dim3(256), 0, st, args, it, cnt);
      };
      auto withNF = [&](auto Pc) {
        if (b->NF == 1) launch(Pc, std::integral_constant<int, 1>{});
        else launch(Pc, std::integral_constant<int, 4>{});
      };
      switch (P) {
        case 1: withNF(std::integral_constant<int, 1>{}); break;
        case 2: withNF(std::integral_constant<int, 2>{}); break;
        case 3: withNF(std::integral_constant<int, 3>{}); break;
        case 4: withNF(std::integral_constant<int, 4>{}); break;
        default: withNF(std::integral_constant<int, 5>{}); break;
      }
      HIP_TRY(hipGetLastError());
      continue;
    }
    const int P = grp;
    dispatch(b->D, P, b->NF, [&](auto Dc, auto Pc, auto NFc) {
      constexpr int D = decltype(Dc)::value, PP = decltype(Pc)::value, NF = decltype(NFc)::value;
      if (fin)
        hipLaunchKernelGGL((frdev::k_batch_stage<D, PP, NF, T, true>), dim3(g), dim3(256), 0, st, args, it, cnt);
      else
        hipLaunchKernelGGL((frdev::k_batch_stage<D, PP, NF, T, false>), dim3(g), dim3(256), 0, st, args, it, cnt);
    });
    HIP_TRY(hipGetLastError());
  }
  return 0;
}

int batchLaunchStage(fr_batch *b, int parity, int s, int grid, hipStream_t st, int gOnly) {
  return b->prec == FR_PREC_FP32 ? batchLaunch<float>(b, parity, s, grid, st, gOnly)
                                 : batchLaunch<double>(b, parity, s, grid, st, gOnly);
}

}  // namespace frimpl
