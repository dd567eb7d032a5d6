// ipg_snap_k_f32.hip -- SNAP kernel instantiations for S = float.
#include "ipg_snap_kernels.h"

namespace ipg {
IPG_SNAP_INSTANTIATE(float)
}  // namespace ipg
