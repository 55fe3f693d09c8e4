// fr_batch.hpp -- one persistent launch per RK stage for many meshes.
//
// Each work item is one element block of one mesh (instance); the meshes may
// differ in size, order P (1..5) and boundaries, and share the dimension, the
// equation and the RK scheme.  A workgroup walks the item list with a
// stride of the grid size and runs the same stage computation as the
// per-mesh kernel (stageBody) for each item, with that mesh's arguments and
// its order's LDS layout carved from one shared buffer.
#pragma once

#include "fr_kernels.hpp"

namespace frdev {

template <int A, int B>
struct CMax {
  static constexpr int value = A > B ? A : B;
};

// LDS bytes of the largest order's stage layout
template <int D, int NF, typename T>
struct BatchSmem {
  static constexpr int bytes =
      CMax<CMax<CMax<(int)sizeof(StageSmem<D, 1, NF, T>), (int)sizeof(StageSmem<D, 2, NF, T>)>::value,
                CMax<(int)sizeof(StageSmem<D, 3, NF, T>), (int)sizeof(StageSmem<D, 4, NF, T>)>::value>::value,
           (int)sizeof(StageSmem<D, 5, NF, T>)>::value;
};

struct BatchItem {
  int inst;   // mesh
  int block;  // element block of that mesh
};

// Every order's workgroup is 256 threads (ElemGeo NTH), so one launch shape
// serves all of them.
template <int D, int NF, typename T, bool FIN>
__global__ __launch_bounds__(256) void k_batch_stage(const StageArgs<T> *__restrict__ args,
                                                     const int *__restrict__ order,
                                                     const BatchItem *__restrict__ items, int nItems) {
  static_assert(ElemGeo<D, 1>::NTH == 256 && ElemGeo<D, 2>::NTH == 256 && ElemGeo<D, 3>::NTH == 256 &&
                    ElemGeo<D, 4>::NTH == 256 && ElemGeo<D, 5>::NTH == 256,
                "every order runs 256-thread workgroups");
  __shared__ __align__(16) unsigned char smem[BatchSmem<D, NF, T>::bytes];
  for (int it = blockIdx.x; it < nItems; it += gridDim.x) {
    const BatchItem w = items[it];
    const StageArgs<T> &a = args[w.inst];
    switch (order[w.inst]) {
      case 1: stageBody<D, 1, NF, T, true, false, FIN>(a, w.block, *reinterpret_cast<StageSmem<D, 1, NF, T> *>(smem)); break;
      case 2: stageBody<D, 2, NF, T, true, false, FIN>(a, w.block, *reinterpret_cast<StageSmem<D, 2, NF, T> *>(smem)); break;
      case 3: stageBody<D, 3, NF, T, true, false, FIN>(a, w.block, *reinterpret_cast<StageSmem<D, 3, NF, T> *>(smem)); break;
      case 4: stageBody<D, 4, NF, T, true, false, FIN>(a, w.block, *reinterpret_cast<StageSmem<D, 4, NF, T> *>(smem)); break;
      default: stageBody<D, 5, NF, T, true, false, FIN>(a, w.block, *reinterpret_cast<StageSmem<D, 5, NF, T> *>(smem)); break;
    }
    __syncthreads();  // the next item reuses the LDS
  }
}

}  // namespace frdev
