// fr_stage64_3d.hip -- the FP64 hex stage-kernel launches.
#include "fr_stage_impl.hpp"

namespace frimpl {
template int launchStageArgsD<double, 3>(fr_ctx *c, const frdev::StageArgs<double> &a, int mode);
}  // namespace frimpl
