// fr_adjoint_v3.hip -- viscous adjoint launches, 3D Navier-Stokes.
#include "fr_adjoint_vis.hpp"

namespace frimpl {
FR_VJP_VIS_INSTANCE(3, 5)
}  // namespace frimpl
