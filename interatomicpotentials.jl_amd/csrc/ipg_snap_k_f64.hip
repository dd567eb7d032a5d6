// ipg_snap_k_f64.hip -- SNAP kernel instantiations for S = double.
#include "ipg_snap_kernels.h"

namespace ipg {
IPG_SNAP_INSTANTIATE(double)
}  // namespace ipg
