// fr_stage32.hip -- the stage-kernel launches in FP32: the quad kernels here,
// the hex kernels in fr_stage32_3d.hip (compiled in parallel).
#include "fr_stage_impl.hpp"

namespace frimpl {
extern template int launchStageArgsD<float, 3>(fr_ctx *c, const frdev::StageArgs<float> &a, int mode);
template int launchStageArgsD<float, 2>(fr_ctx *c, const frdev::StageArgs<float> &a, int mode);
template int launchStageArgs<float>(fr_ctx *c, const frdev::StageArgs<float> &a, int mode);
}  // namespace frimpl
