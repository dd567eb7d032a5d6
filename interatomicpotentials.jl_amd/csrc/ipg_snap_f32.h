// ipg_snap_f32.h -- launchers of the FP32 U and force kernels of the
// single-system adjoint step (ipg_snap_f32.hip), 2J <= 8, on the
// atom-blocked U/Y layout [ceil(n/32)][idxu_max][32][2].
#pragma once
#include "ipg_snap.h"

namespace ipg {

// U (ulisttot) by Wigner rows.
void snap_launch_ui_rows32(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, float* U);
// dE_i/dr_ij per pair from the Y of snap_launch_zy32.
void snap_launch_dedr_rev32(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const float* Y, double* dedr);
// The same reverse-mode force kernel for S = double, Dual<float> or
// Dual<double>, over the Y of snap_launch_zby<S> ([n][idxu_max][2] of S):
// dedr gets the values and, for dual numbers, dedr_t (if not NULL) their
// tangents along R.tang.
template <typename S>
void snap_launch_dedr_rev(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const S* Y, double* dedr,
                          double* dedr_t);

}  // namespace ipg
