// fr_adjoint_v3ad.hip -- viscous adjoint launches, 3D advection-diffusion.
#include "fr_adjoint_vis.hpp"

namespace frimpl {
FR_VJP_VIS_INSTANCE(3, 1)
}  // namespace frimpl
