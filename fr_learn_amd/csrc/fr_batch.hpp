// fr_batch.hpp -- persistent stage launches for many meshes at once.
//
// Each work item is one element block of one mesh (instance); the meshes may
// differ in size, order P (1..5) and boundaries, and share the dimension, the
// equation and the RK scheme.  Per RK stage there is one persistent launch
// per order present (and, in fr_tri.hpp, per triangle order): a workgroup
// walks that group's item list with a stride
// of the grid size and runs the same stage computation as the per-mesh
// kernel (stageBody) for each item, with that mesh's arguments.
#pragma once

#include "fr_kernels.hpp"

namespace frdev {

struct BatchItem {
  int inst;   // mesh
  int block;  // element block of that mesh
};

// The thread index, made opaque to the compiler for one item of a persistent
// loop: index arithmetic derived from it is then recomputed per item instead
// of being hoisted out of the loop and kept live across it (which doubles the
// register allocation of the stage computation and halves occupancy).
__device__ __forceinline__ int itemThreadIdx() {
  int t = (int)threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}


// One kernel per order keeps each order's register allocation (a single
// kernel switching over the orders is allocated for their union and spills).
// L is the tile layout (ElemGeo); a workgroup is ElemGeo<D, P, L>::NTH
// threads (at most 256), and an item is a block of EPB elements of that
// layout.  rev walks the items backwards (the stages alternate, so that a
// stage reads first what the one before wrote last).
template <int D, int P, int NF, typename T, bool FIN, int L>
__global__ __launch_bounds__(256) void k_batch_stage(const StageArgs<T> *__restrict__ args,
                                                     const BatchItem *__restrict__ items, int nItems, int rev) {
  __shared__ StageSmem<D, P, NF, T, L> sm;
  for (int it = blockIdx.x; it < nItems; it += gridDim.x) {
    const BatchItem w = items[rev ? nItems - 1 - it : it];
    stageBody<D, P, NF, T, true, false, FIN, 0, false, L>(args[w.inst], w.block, sm, itemThreadIdx());
    __syncthreads();  // the next item reuses the LDS
  }
}

}  // namespace frdev
