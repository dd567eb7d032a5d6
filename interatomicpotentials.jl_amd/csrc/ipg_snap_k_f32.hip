// ipg_snap_k_f32.hip -- SNAP kernel instantiation for S = float: the
// neighbor-parallel U kernel of the batched descriptors (the single-system
// FP32 U and force kernels are in ipg_snap_f32.hip, Z/B/Y in ipg_snap_zy.hip).
#include "ipg_snap_kernels.h"

namespace ipg {
IPG_SNAP_INSTANTIATE_U(float)
}  // namespace ipg
