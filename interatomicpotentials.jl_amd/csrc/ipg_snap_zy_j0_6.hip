// ipg_snap_zy_j0_6.hip -- instantiations of the FP32 Z kernel launcher (ipg_snap_zy.h):
// twojmax 0 to 6, adjoint and descriptors.
#include "ipg_snap_zy.h"

namespace ipg {
IPG_ZY_LAUNCHER(template, 0, false)
IPG_ZY_LAUNCHER(template, 0, true)
IPG_ZY_LAUNCHER(template, 1, false)
IPG_ZY_LAUNCHER(template, 1, true)
IPG_ZY_LAUNCHER(template, 2, false)
IPG_ZY_LAUNCHER(template, 2, true)
IPG_ZY_LAUNCHER(template, 3, false)
IPG_ZY_LAUNCHER(template, 3, true)
IPG_ZY_LAUNCHER(template, 4, false)
IPG_ZY_LAUNCHER(template, 4, true)
IPG_ZY_LAUNCHER(template, 5, false)
IPG_ZY_LAUNCHER(template, 5, true)
IPG_ZY_LAUNCHER(template, 6, false)
IPG_ZY_LAUNCHER(template, 6, true)
}  // namespace ipg
