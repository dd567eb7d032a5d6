// ipg_snap_zy_j8_desc.hip -- instantiations of the FP32 Z kernel launcher (ipg_snap_zy.h):
// twojmax 8, descriptors (Z, B).
#include "ipg_snap_zy.h"

namespace ipg {
IPG_ZY_LAUNCHER(template, 8, true)
}  // namespace ipg
