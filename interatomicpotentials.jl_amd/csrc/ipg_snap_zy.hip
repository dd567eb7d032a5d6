// ipg_snap_zy.hip -- dispatch of the FP32 Z kernels (ipg_snap_zy.h) over
// twojmax: the adjoint Z/energy/Y kernel and the descriptor Z/B kernel.
#include "ipg_snap_zy.h"

namespace ipg {

#define IPG_ZY_EXTERN(J) IPG_ZY_LAUNCHER(extern template, J, false) IPG_ZY_LAUNCHER(extern template, J, true)
IPG_ZY_EXTERN(0)
IPG_ZY_EXTERN(1)
IPG_ZY_EXTERN(2)
IPG_ZY_EXTERN(3)
IPG_ZY_EXTERN(4)
IPG_ZY_EXTERN(5)
IPG_ZY_EXTERN(6)
IPG_ZY_EXTERN(7)
IPG_ZY_EXTERN(8)

template <bool DESC>
static void launch_zy(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const float* U, const double* beta, float* Y,
                      double* eatom, float* Z, double* B) {
  switch (P.J) {
    case 0: launch_zy_j<0, DESC>(ctx, P, R, U, beta, Y, eatom, Z, B); break;
    case 1: launch_zy_j<1, DESC>(ctx, P, R, U, beta, Y, eatom, Z, B); break;
    case 2: launch_zy_j<2, DESC>(ctx, P, R, U, beta, Y, eatom, Z, B); break;
    case 3: launch_zy_j<3, DESC>(ctx, P, R, U, beta, Y, eatom, Z, B); break;
    case 4: launch_zy_j<4, DESC>(ctx, P, R, U, beta, Y, eatom, Z, B); break;
    case 5: launch_zy_j<5, DESC>(ctx, P, R, U, beta, Y, eatom, Z, B); break;
    case 6: launch_zy_j<6, DESC>(ctx, P, R, U, beta, Y, eatom, Z, B); break;
    case 7: launch_zy_j<7, DESC>(ctx, P, R, U, beta, Y, eatom, Z, B); break;
    default: launch_zy_j<8, DESC>(ctx, P, R, U, beta, Y, eatom, Z, B); break;
  }
}

void snap_launch_zy32(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const float* U, const double* beta, float* Y,
                      double* eatom) {
  if (R.n == 0) return;
  IPG_REQUIRE(!P.bnorm, "internal: bnorm is not supported by the FP32 Z kernels");
  launch(ctx, "snap_zy_fp32", [&] { launch_zy<false>(ctx, P, R, U, beta, Y, eatom, nullptr, nullptr); });
}

void snap_launch_zb32(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const float* U, float* Z, double* B) {
  if (R.n == 0) return;
  IPG_REQUIRE(!P.bnorm, "internal: bnorm is not supported by the FP32 Z kernels");
  launch(ctx, "snap_zb_fp32", [&] { launch_zy<true>(ctx, P, R, U, nullptr, nullptr, nullptr, Z, B); });
}

}  // namespace ipg
