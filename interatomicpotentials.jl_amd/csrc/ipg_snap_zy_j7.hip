// ipg_snap_zy_j7.hip -- instantiations of the FP32 Z kernel launcher (ipg_snap_zy.h):
// twojmax 7, adjoint and descriptors.
#include "ipg_snap_zy.h"

namespace ipg {
IPG_ZY_LAUNCHER(template, 7, false)
IPG_ZY_LAUNCHER(template, 7, true)
}  // namespace ipg
