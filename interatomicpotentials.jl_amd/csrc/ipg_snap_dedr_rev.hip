// ipg_snap_dedr_rev.hip -- the reverse-mode force kernel (the sweep of
// ipg_wigner_adj.h) in FP64 and on dual numbers, over the Y of the generic
// Z/Y kernel ([n][idxu_max][2] of S).
//
// dE_i/dr_ij = dsfac r_hat Phi + sfac grad_r Phi with Phi = sum_half
// 2 w(e) Re(conj(u_e) Y_i,e) (the per-pair force term of compute_ui.jl:
// 186-326 contracted with Y), grad_r Phi from the adjoints of the
// Cayley-Klein parameters through their FP64 Jacobian.
//
// On dual numbers (the pullback of the forces along a direction gF) every
// input of that expression carries its tangent along r_ij(e) = r_ij + e
// (gF_j - gF_i) and Y_i(e) (the tangent of Y from the dual Z/Y kernel), so
// the tangent of dE_i/dr_ij is exact: the second derivatives of the Wigner
// recursion with respect to (a, b) come from the sweep's own dual
// arithmetic (forward mode over the reverse sweep), with no derivative
// rows of du/dr formed.
#include "ipg_snap_f32.h"
#include "ipg_snap_kernels.h"
#include "ipg_wigner_adj.h"

namespace ipg {

namespace wadj {
template <typename T>
struct Base<Dual<T>> {
  using type = T;
};
}  // namespace wadj

// Pins a dual accumulator where it is written (both parts).
__device__ __forceinline__ void pin(Dual<double>& x) { asm volatile("" : "+v"(x.v), "+v"(x.t)); }

// Makes v depend on x (an empty asm), for the sweep's ordering handles.
__device__ __forceinline__ void tie_to(double& v, double x) { asm volatile("; tie %1" : "+v"(v) : "v"(x)); }
template <typename T>
__device__ __forceinline__ void tie_to(Dual<T>& v, Dual<T> x) {
  asm volatile("; tie %2" : "+v"(v.v), "+v"(v.t) : "v"(x.v));
}
__device__ __forceinline__ double value_part(double x) { return x; }
template <typename T>
__device__ __forceinline__ T value_part(Dual<T> x) {
  return x.v;
}

// Seeds of the sweep: Y_i in the layout of the generic Z/Y kernel,
// [jju][re, im] of S, unweighted (the sweep applies 2 w(e) N(e)).
template <typename S>
struct PlainYSeed {
  static constexpr bool weighted = false;
  const S* y;
  struct C {
    S x, y;
  };
  __device__ __forceinline__ int order(S x) const {
    int z = 0;
    asm volatile("; order %1" : "+v"(z) : "v"(value_part(x)));
    return z;
  }
  __device__ __forceinline__ S tie(S v, S x) const {
    tie_to(v, x);
    return v;
  }
  template <int jl, int mb, int ma>
  __device__ __forceinline__ C operator()(IC<jl>, IC<mb>, IC<ma>, int tok) const {
    const S* p = y + 2 * (ub_of(jl) + (jl + 1) * mb + ma) + tok;
    return C{p[0], p[1]};
  }
};

// Waves per SIMD: two in FP64 and on FP32 duals; FP64 duals need the
// whole register file of one.
template <typename S>
constexpr int rev_waves() {
  return std::is_same<S, Dual<double>>::value ? 1 : 2;
}

template <int J, typename S>
__global__ __launch_bounds__(256, rev_waves<S>()) void k_snap_dedr_rev(SnapDev P, SnapRun R, const S* __restrict__ Y,
                                                       double* __restrict__ dedr, double* __restrict__ dedr_t) {
  using W = typename WideOf<S>::type;  // double, or Dual<double>
  constexpr int IDXU = ub_of(J + 1);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (R.d_np ? *R.d_np : R.npairs)) return;
  const int32_t i = R.center[q], j = R.nbr[q];
  const int si = snap_species(R.species, i, P.nspecies), sj = snap_species(R.species, j, P.nspecies);
  W x, y, z;
  load_pair<W>(R, q, i, j, x, y, z);
  const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
  CK<W> d;
  cayley_klein_impl<W, true>(x, y, z, rcut, P.weight[sj], P, d);
  const wadj::AB<S> c{narrow<S>(d.a_r), narrow<S>(d.a_i), narrow<S>(d.b_r), narrow<S>(d.b_i)};
  // the FP64 Jacobian rows wait in LDS while the sweep runs:
  // [k][thread] for k = (d a_r, d a_i, d b_r, d b_i) x 3, r_hat x 3, sfac, dsfac
  __shared__ W jac[17][256];
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    jac[k][t] = d.da_r[k];
    jac[3 + k][t] = d.da_i[k];
    jac[6 + k][t] = d.db_r[k];
    jac[9 + k][t] = d.db_i[k];
    jac[12 + k][t] = d.u[k];
  }
  jac[15][t] = d.sfac;
  jac[16][t] = d.dsfac;
  PlainYSeed<S> seed{Y + (int64_t)i * IDXU * 2};
  wadj::AccW<W> acc;
  wadj::vjp<S, J>(c, seed, acc);
  const W g0 = jac[16][t] * acc.phi, sf = jac[15][t];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const W v = g0 * jac[12 + k][t] +
                sf * (acc.ar * jac[k][t] + acc.ai * jac[3 + k][t] + acc.br * jac[6 + k][t] + acc.bi * jac[9 + k][t]);
    dedr[3 * q + k] = value_of(v);
    if (dedr_t) dedr_t[3 * q + k] = tangent_of(v);
  }
}

template <typename S>
void snap_launch_dedr_rev(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const S* Y, double* dedr,
                          double* dedr_t) {
  if (R.npairs == 0) return;
  IPG_REQUIRE(P.J <= 8, "internal: reverse-mode force kernel compiled for 2J <= 8");
  const unsigned g = grid_for(R.npairs, 256);
  const char* nm = IsDual<S>::value ? "snap_dedr_dual" : "snap_dedr_fp64";
  launch(ctx, nm, [&] {
    switch (P.J) {
      case 0: k_snap_dedr_rev<0, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 1: k_snap_dedr_rev<1, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 2: k_snap_dedr_rev<2, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 3: k_snap_dedr_rev<3, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 4: k_snap_dedr_rev<4, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 5: k_snap_dedr_rev<5, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 6: k_snap_dedr_rev<6, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 7: k_snap_dedr_rev<7, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      default: k_snap_dedr_rev<8, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
    }
  });
}

template void snap_launch_dedr_rev<double>(ipg_ctx*, const SnapDev&, const SnapRun&, const double*, double*,
                                           double*);
template void snap_launch_dedr_rev<Dual<float>>(ipg_ctx*, const SnapDev&, const SnapRun&, const Dual<float>*,
                                                double*, double*);
template void snap_launch_dedr_rev<Dual<double>>(ipg_ctx*, const SnapDev&, const SnapRun&, const Dual<double>*,
                                                 double*, double*);

}  // namespace ipg
