// ipg_snap_k_f32.hip -- SNAP kernel instantiations for S = float (U and
// force kernels on the atom-blocked layout; Z/B/Y is ipg_snap_zy.hip).
#include "ipg_snap_kernels.h"

namespace ipg {
IPG_SNAP_INSTANTIATE_UD(float)
}  // namespace ipg
