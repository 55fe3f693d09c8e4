// fr_batch.hip -- the batched persistent stage launches (kernels in fr_batch.hpp).
#include "fr_ctx.hpp"

namespace frimpl {

// Stage s of the meshes of order P (P = 0: every order, one after the
// other); parity < 0: the residual-only tuning launch.
template <typename T>
int batchLaunch(fr_batch *b, int parity, int s, int grid, hipStream_t st, int Ponly) {
  const size_t n = b->ctx.size();
  const frdev::StageArgs<T> *args = (parity < 0) ? (const frdev::StageArgs<T> *)b->argsTune
                                                 : (const frdev::StageArgs<T> *)b->args[parity] + (size_t)s * n;
  int mode = 0;
  double rk;
  if (parity >= 0) stageMode(b->ctx[0], s, mode, rk);
  const bool fin = mode >= 2;
  for (int P = 5; P >= 1; P--) {
    const int cnt = b->itemCount[P];
    if (!cnt || (Ponly && P != Ponly)) continue;
    const frdev::BatchItem *it = b->items + b->itemStart[P];
    const unsigned g = (unsigned)std::min(grid, cnt);
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

int batchLaunchStage(fr_batch *b, int parity, int s, int grid, hipStream_t st, int Ponly) {
  return b->prec == FR_PREC_FP32 ? batchLaunch<float>(b, parity, s, grid, st, Ponly)
                                 : batchLaunch<double>(b, parity, s, grid, st, Ponly);
}

}  // namespace frimpl
