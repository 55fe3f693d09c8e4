// fr_stage32.hip -- the stage-kernel launches in FP32.
#include "fr_stage_impl.hpp"

namespace frimpl {
template int launchStageArgs<float>(fr_ctx *c, const frdev::StageArgs<float> &a, int mode);
}  // namespace frimpl
