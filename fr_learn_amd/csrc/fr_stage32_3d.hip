// fr_stage32_3d.hip -- the FP32 hex stage-kernel launches.
#include "fr_stage_impl.hpp"

namespace frimpl {
template int launchStageArgsD<float, 3>(fr_ctx *c, const frdev::StageArgs<float> &a, int mode);
}  // namespace frimpl
