// fr_batch.hip -- the batched persistent stage launches (kernels in fr_batch.hpp).
#include "fr_ctx.hpp"
#include "fr_tri.hpp"

namespace frimpl {

// Kernel group of a member: its order P, plus 5 for triangles, plus 10 for
// hexes.
int batchGroup(const fr_ctx *c) { return c->P + (c->tri ? 5 : c->D == 3 ? 10 : 0); }

void batchGroupKind(int g, int &D, int &P, bool &tri) {
  tri = g > 5 && g <= 10;
  D = g > 10 ? 3 : 2;
  P = g - (g > 10 ? 10 : g > 5 ? 5 : 0);
}

// Elements per work item of member c in tile layout lay.
int batchEpb(const fr_batch *, const fr_ctx *c, int lay) {
  const int tpe = c->tri ? frdev::triTPEL(c->P, lay) : frdev::tensorTPE(c->D, c->P, lay);
  return tpe >= 256 ? 1 : 256 / tpe;
}

namespace {

template <int P, int NF, typename T, int L>
void triLaunch(bool fin, unsigned g, hipStream_t st, const frdev::StageArgs<T> *args, const frdev::BatchItem *it,
               int cnt, int rev) {
  // 256 threads in either layout (the matrix-core products use whole wavefronts)
  if (fin)
    hipLaunchKernelGGL((frdev::k_batch_tri_stage<P, NF, T, true, L>), dim3(g), dim3(256), 0, st, args, it, cnt, rev);
  else
    hipLaunchKernelGGL((frdev::k_batch_tri_stage<P, NF, T, false, L>), dim3(g), dim3(256), 0, st, args, it, cnt, rev);
}

template <int D, int P, int NF, typename T, int L>
void tensorLaunch(bool fin, unsigned g, hipStream_t st, const frdev::StageArgs<T> *args,
                  const frdev::BatchItem *it, int cnt, int rev) {
  constexpr int NTH = frdev::ElemGeo<D, P, L>::NTH;
  if (fin)
    hipLaunchKernelGGL((frdev::k_batch_stage<D, P, NF, T, true, L>), dim3(g), dim3(NTH), 0, st, args, it, cnt, rev);
  else
    hipLaunchKernelGGL((frdev::k_batch_stage<D, P, NF, T, false, L>), dim3(g), dim3(NTH), 0, st, args, it, cnt, rev);
}

// Stage s of the meshes of group grp in tile layout lay with a persistent
// grid of `grid` workgroups; parity < 0: the residual-only tuning launch.
template <typename T>
int launchGroup(fr_batch *b, int parity, int s, hipStream_t st, int grp, int lay, int grid) {
  const size_t n = b->ctx.size();
  const frdev::StageArgs<T> *args = (parity < 0) ? (const frdev::StageArgs<T> *)b->argsTune
                                                 : (const frdev::StageArgs<T> *)b->args[parity] + (size_t)s * n;
  int mode = 0;
  double rk;
  if (parity >= 0) stageMode(b->ctx[0], s, mode, rk);
  const bool fin = mode >= 2;
  const int rev = parity >= 0 ? (s & 1) : 0;  // the stages alternate their item order
  const int cnt = b->itemCount[lay][grp];
  if (!cnt) return 0;
  const frdev::BatchItem *it = b->items[lay] + b->itemStart[lay][grp];
  const unsigned g = (unsigned)std::max(1, std::min(grid, cnt));
  int gD, gP;
  bool gTri;
  batchGroupKind(grp, gD, gP, gTri);
  const int gNF = b->groupNF[grp];
  if (gTri) {
    auto withL = [&](auto Pc, auto NFc) {
      constexpr int PP = decltype(Pc)::value, NF = decltype(NFc)::value;
      if (lay == 1) triLaunch<PP, NF, T, 1>(fin, g, st, args, it, cnt, rev);
      else triLaunch<PP, NF, T, 0>(fin, g, st, args, it, cnt, rev);
    };
    auto withNF = [&](auto Pc) {
      if (gNF == 1) withL(Pc, std::integral_constant<int, 1>{});
      else withL(Pc, std::integral_constant<int, 4>{});
    };
    switch (gP) {
      case 1: withNF(std::integral_constant<int, 1>{}); break;
      case 2: withNF(std::integral_constant<int, 2>{}); break;
      case 3: withNF(std::integral_constant<int, 3>{}); break;
      case 4: withNF(std::integral_constant<int, 4>{}); break;
      default: withNF(std::integral_constant<int, 5>{}); break;
    }
  } else {
    dispatch(gD, gP, gNF, [&](auto Dc, auto Pc, auto NFc) {
      constexpr int D = decltype(Dc)::value, PP = decltype(Pc)::value, NF = decltype(NFc)::value;
      if constexpr (D == 2) {
        if (lay == 1) {
          tensorLaunch<D, PP, NF, T, 1>(fin, g, st, args, it, cnt, rev);
          return;
        }
      }
      tensorLaunch<D, PP, NF, T, 0>(fin, g, st, args, it, cnt, rev);
    });
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace

int batchLaunchGroup(fr_batch *b, int parity, int s, hipStream_t st, int grp, int lay, int grid) {
  return b->prec == FR_PREC_FP32 ? launchGroup<float>(b, parity, s, st, grp, lay, grid)
                                 : launchGroup<double>(b, parity, s, st, grp, lay, grid);
}

}  // namespace frimpl
