// ipg_snap_plain.hip -- launcher of the per-block Y coefficient table
// (beta entries whose dB/dr half-sums use each Z block, with their
// multiplicities) used by the FP64 and Dual adjoint paths.
#include "ipg_snap_kernels.h"

namespace ipg {

template <typename T>
void snap_launch_ycoef(ipg_ctx* ctx, const SnapDev& P, const double* beta, T* ycoef) {
  const int nt = P.nspecies * P.nblk;
  launch(ctx, "snap_ycoef", [&] { k_snap_ycoef<T><<<grid_for(nt, 256), 256, 0, ctx->stream>>>(P, beta, ycoef); });
}
template void snap_launch_ycoef<float>(ipg_ctx*, const SnapDev&, const double*, float*);
template void snap_launch_ycoef<double>(ipg_ctx*, const SnapDev&, const double*, double*);

}  // namespace ipg
