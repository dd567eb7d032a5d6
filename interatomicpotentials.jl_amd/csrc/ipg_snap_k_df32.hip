// ipg_snap_k_df32.hip -- SNAP kernel instantiations for S = Dual<float>.
#include "ipg_snap_kernels.h"

namespace ipg {
IPG_SNAP_INSTANTIATE(Dual<float>)
}  // namespace ipg
