// ipg_snap_zy_j8.hip -- instantiations of the FP32 Z kernel launcher (ipg_snap_zy.h):
// twojmax 8, adjoint (Z, energy, Y).
#include "ipg_snap_zy.h"

namespace ipg {
IPG_ZY_LAUNCHER(template, 8, false)
}  // namespace ipg
