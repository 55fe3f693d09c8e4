// fr_stage64.hip -- the stage-kernel launches in FP64: the quad kernels here,
// the hex kernels in fr_stage64_3d.hip (compiled in parallel).
#include "fr_stage_impl.hpp"

namespace frimpl {
extern template int launchStageArgsD<double, 3>(fr_ctx *c, const frdev::StageArgs<double> &a, int mode);
template int launchStageArgsD<double, 2>(fr_ctx *c, const frdev::StageArgs<double> &a, int mode);
template int launchStageArgs<double>(fr_ctx *c, const frdev::StageArgs<double> &a, int mode);
}  // namespace frimpl
