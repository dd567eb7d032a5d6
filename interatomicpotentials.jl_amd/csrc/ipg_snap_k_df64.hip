// ipg_snap_k_df64.hip -- SNAP kernel instantiations for S = Dual<double>.
#include "ipg_snap_kernels.h"

namespace ipg {
IPG_SNAP_INSTANTIATE(Dual<double>)
}  // namespace ipg
