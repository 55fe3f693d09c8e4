// fr_adjoint_v2.hip -- viscous adjoint launches, 2D Navier-Stokes and advection-diffusion.
#include "fr_adjoint_vis.hpp"

namespace frimpl {
FR_VJP_VIS_INSTANCE(2, 4)
FR_VJP_VIS_INSTANCE(2, 1)
}  // namespace frimpl
