// fr_stage64.hip -- the stage-kernel launches in FP64.
#include "fr_stage_impl.hpp"

namespace frimpl {
template int launchStageArgs<double>(fr_ctx *c, const frdev::StageArgs<double> &a, int mode);
}  // namespace frimpl
