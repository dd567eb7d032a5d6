// ipg_snap_f32.hip -- the FP32 U and force kernels of the single-system
// adjoint step, both built on the rescaled Wigner row chains of
// ipg_wigner_adj.h (2J <= 8, atom-blocked U/Y layout [n/32][idxu_max][32]).
#include "ipg_snap_f32.h"
#include "ipg_snap_kernels.h"
#include "ipg_wigner_adj.h"

namespace ipg {

// ---------------------------------------------------------------------------
// K1 in FP32: ulisttot (compute_ui.jl:5-44, add_uarraytot :155-175) by
// Wigner rows.  A block holds UR_ATOMS atoms and one wave per row
// mb = 0..J/2; lane 4k+s of a wave takes every 4th pair of atom k, starting
// at its s-th pair.  Every row chain forms the same number of entries (the
// staircase of the rows below, then the row), so the waves of a block stay
// balanced, and a lane keeps its row's sums w_j sfac v_e (v = u / N) in
// registers over its pairs: no cross-lane reduction per neighbor.  The
// Cayley-Klein parameters of the block's pairs are formed once, in FP64
// and rounded to FP32, into a per-pair scratch array read by the block's
// waves (L1/L2 resident).  At the end the four lanes of an atom add
// their sums by quad permutations (the same order in every lane, so the
// result is deterministic), each entry is scaled by N, wself goes on the
// diagonal and the right half is written by the inversion symmetry.
// ---------------------------------------------------------------------------
constexpr int UR_ATOMS = 16;  // atoms per block

// sin and cos of an FP64 argument of moderate size (|x| < 1e3): Cody-Waite
// reduction by pi/2 (the quotient's product with each part of pi/2 formed
// exactly by an FMA), then the Taylor series of sin and cos on
// [-pi/4, pi/4] through x^17 and x^16 (truncation below 1e-17).  Relative
// error about 1e-16, far below the FP32 rounding these kernels apply to the
// Cayley-Klein parameters, with a fraction of the general routines'
// registers and instructions.
__device__ __forceinline__ void sincos_small(double x, double& sn, double& cs) {
  constexpr double PIO2_HI = 1.5707963267948966;        // pi/2 rounded to double
  constexpr double PIO2_LO = 6.123233995736766e-17;     // pi/2 - PIO2_HI
  constexpr double TWO_OVER_PI = 0.6366197723675814;
  const double k = rint(x * TWO_OVER_PI);
  double r = fma(-k, PIO2_HI, x);
  r = fma(-k, PIO2_LO, r);
  const double r2 = r * r;
  // Horner in r^2 with coefficients (-1)^n / (2n+1)! and (-1)^n / (2n)!
  double ps = -1.0 / 355687428096000.0;  // -1/17!
  ps = fma(ps, r2, 1.0 / 1307674368000.0);
  ps = fma(ps, r2, -1.0 / 6227020800.0);
  ps = fma(ps, r2, 1.0 / 39916800.0);
  ps = fma(ps, r2, -1.0 / 362880.0);
  ps = fma(ps, r2, 1.0 / 5040.0);
  ps = fma(ps, r2, -1.0 / 120.0);
  ps = fma(ps, r2, 1.0 / 6.0);
  const double s0 = fma(-r * r2, ps, r);
  double pc = 1.0 / 20922789888000.0;  // 1/16!
  pc = fma(pc, r2, -1.0 / 87178291200.0);
  pc = fma(pc, r2, 1.0 / 479001600.0);
  pc = fma(pc, r2, -1.0 / 3628800.0);
  pc = fma(pc, r2, 1.0 / 40320.0);
  pc = fma(pc, r2, -1.0 / 720.0);
  pc = fma(pc, r2, 1.0 / 24.0);
  pc = fma(pc, r2, -0.5);
  const double c0 = fma(r2, pc, 1.0);
  const int q = (int)k & 3;
  sn = q == 0 ? s0 : q == 1 ? c0 : q == 2 ? -s0 : -c0;
  cs = q == 0 ? c0 : q == 1 ? -s0 : q == 2 ? -c0 : s0;
}

// Cayley-Klein parameters (and, with DERIV, their gradients) in FP64 as
// cayley_klein_impl<double> forms them (compute_ui.jl:191-216, 328-347,
// switching function :350-380), with sincos_small for the two angles.
template <bool DERIV>
__device__ __forceinline__ void ck64(double x, double y, double z, double rcut, double wj, const SnapDev& P,
                                     CK<double>& c) {
  const double rsq = x * x + y * y + z * z;
  const double r = sqrt(rsq);
  const double rscale0 = P.rfac0 * M_PI / (rcut - P.rmin0);
  double sn, cs;
  sincos_small((r - P.rmin0) * rscale0, sn, cs);
  const double z0 = r * cs / sn;
  const double r0inv = 1.0 / sqrt(r * r + z0 * z0);
  c.a_r = z0 * r0inv;
  c.a_i = -z * r0inv;
  c.b_r = y * r0inv;
  c.b_i = -x * r0inv;
  double sfac = 1.0, dsfac = 0.0;
  if (P.switch_flag && r > P.rmin0) {
    const double rcutfac = M_PI / (rcut - P.rmin0);
    double sa, ca;
    sincos_small((r - P.rmin0) * rcutfac, sa, ca);
    sfac = r > rcut ? 0.0 : 0.5 * (ca + 1.0);
    dsfac = -0.5 * (sa * rcutfac);
  }
  c.sfac = sfac * wj;
  c.dsfac = dsfac * wj;
  if (DERIV) {
    const double dz0dr = z0 / r - (r * rscale0) * (rsq + z0 * z0) / rsq;
    const double rinv = 1.0 / r;
    c.u[0] = x * rinv;
    c.u[1] = y * rinv;
    c.u[2] = z * rinv;
    const double dr0invdr = -r0inv * r0inv * r0inv * (r + z0 * dz0dr);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double dr0inv = dr0invdr * c.u[k], dz0 = dz0dr * c.u[k];
      c.da_r[k] = dz0 * r0inv + z0 * dr0inv;
      c.da_i[k] = -z * dr0inv;
      c.db_r[k] = y * dr0inv;
      c.db_i[k] = -x * dr0inv;
    }
    c.da_i[2] = c.da_i[2] - r0inv;
    c.db_i[0] = c.db_i[0] - r0inv;
    c.db_r[1] = c.db_r[1] + r0inv;
  }
}

// Emitter of one row chain: adds w v_e to the row's sums.
template <int MB>
struct RowSum {
  float* accr;
  float* acci;
  float w;
  template <int j, int ma>
  __device__ __forceinline__ void operator()(IC<j>, IC<ma>, float vr, float vi) {
    constexpr int k = wadj::row_index(MB, j, ma);
    accr[k] = accr[k] + w * vr;
    acci[k] = acci[k] + w * vi;
  }
  __device__ __forceinline__ wadj::AB<float> next(const wadj::AB<float>& c) const { return c; }
};

template <int J>
__global__ __launch_bounds__(64 * (J / 2 + 1)) void k_snap_ui_rows32(SnapDev P, SnapRun R, float2* __restrict__ Ub,
                                                                     float4* __restrict__ ckab,
                                                                     float* __restrict__ ckw) {
  constexpr int NR = J / 2 + 1, NT = 64 * NR;
  constexpr int IDXU = ub_of(J + 1);
  constexpr int EMAX = wadj::row_count(J, 0);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int s = lane & 3;
  const int64_t a0 = (int64_t)blockIdx.x * UR_ATOMS, a = a0 + (lane >> 2);
  const bool valid = a < R.n;
  const int64_t aend = a0 + UR_ATOMS < R.n ? a0 + UR_ATOMS : R.n;
  if (wv == 0 && s == 0 && valid && R.species && R.flag && (unsigned)R.species[a] >= (unsigned)P.nspecies)
    atomicOr(R.flag, 16);
  // the Cayley-Klein parameters of the block's pairs, once per pair, before
  // any row sum is live (the block's own pairs: no other block reads them)
  for (int64_t p = R.off[a0] + tid; p < R.off[aend]; p += NT) {
    const int32_t i = R.center[p], j = R.nbr[p];
    const int si = snap_species(R.species, i, P.nspecies), sj = snap_species(R.species, j, P.nspecies);
    const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
    CK<double> w;
    ck64<false>(R.rvec[3 * p], R.rvec[3 * p + 1], R.rvec[3 * p + 2], rcut, P.weight[sj], P, w);
    ckab[p] = make_float4((float)w.a_r, (float)w.a_i, (float)w.b_r, (float)w.b_i);
    ckw[p] = (float)w.sfac;
  }
  __syncthreads();
  float accr[EMAX], acci[EMAX];
#pragma unroll
  for (int k = 0; k < EMAX; k++) accr[k] = acci[k] = 0.f;
  const int64_t q0 = valid ? R.off[a] : 0, q1 = valid ? R.off[a + 1] : 0;
  static_for<0, NR>([&](auto mbc) {
    constexpr int MB = decltype(mbc)::value;
    if (wv != MB) return;
#pragma unroll 1
    for (int64_t q = q0 + s; q < q1; q += 4) {
      const float4 ab = ckab[q];
      RowSum<MB> emit{accr, acci, ckw[q]};
      wadj::row_forward<float, J, MB>(wadj::AB<float>{ab.x, ab.y, ab.z, ab.w}, emit);
    }
  });
  // the four lanes of an atom add their sums (quad permutations, the same
  // order in every lane); lane s writes the entries k = s mod 4, each entry
  // reduced and stored before the next so that few values are live
  static_for<0, NR>([&](auto mbc) {
    constexpr int MB = decltype(mbc)::value;
    if (wv != MB) return;
    float2* out = Ub + (valid ? (a >> 5) * (int64_t)IDXU * 32 + (a & 31) : 0);
    const float wself = (float)P.wself;
    auto put = [&](auto jc, auto mc) {
      constexpr int j = decltype(jc)::value, ma = decltype(mc)::value;
      constexpr int k = wadj::row_index(MB, j, ma);
      float vr = accr[k], vi = acci[k];
      vr = vr + dpp_v<0xB1>(vr);
      vi = vi + dpp_v<0xB1>(vi);
      vr = vr + dpp_v<0x4E>(vr);
      vi = vi + dpp_v<0x4E>(vi);
      asm volatile("" : "+v"(vr), "+v"(vi));
      if (!valid || s != k % 4) return;
      constexpr float nm = (float)wadj::norm(j, MB, ma);
      const float ur = nm * vr, ui = nm * vi;
      constexpr bool diag = ma == MB;
      out[32 * (ub_of(j) + (j + 1) * MB + ma)] = make_float2(diag ? ur + wself : ur, ui);
      // (j-ma, j-mb) = (-1)^(ma+mb) conj((ma, mb)), unless the entry is its own mirror
      if constexpr (!(2 * MB == j && 2 * ma == j)) {
        constexpr bool neg = (ma + MB) & 1;
        const float mr = neg ? -ur : ur, mi = neg ? ui : -ui;
        out[32 * (ub_of(j) + (j + 1) * (j - MB) + (j - ma))] = make_float2(diag ? mr + wself : mr, mi);
      }
    };
    static_for<0, MB + 1>([&](auto mc) { put(IC<2 * MB>{}, mc); });
    static_for<2 * MB + 1, J + 1>([&](auto jc) {
      static_for<0, decltype(jc)::value + 1>([&](auto mc) { put(jc, mc); });
    });
  });
}

template <int J>
static void launch_ui_rows(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, float* U, float4* ab, float* w) {
  const unsigned g = (unsigned)((R.n + UR_ATOMS - 1) / UR_ATOMS);
  k_snap_ui_rows32<J><<<g, 64 * (J / 2 + 1), 0, ctx->stream>>>(P, R, reinterpret_cast<float2*>(U), ab, w);
}

void snap_launch_ui_rows32(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, float* U) {
  if (R.n == 0) return;
  IPG_REQUIRE(P.J <= 8, "internal: FP32 U kernel compiled for 2J <= 8");
  const int64_t np = R.npairs > 0 ? R.npairs : 1;
  float4* ab = scratch_t<float4>(ctx, "snap_ck32_ab", np);
  float* w = scratch_t<float>(ctx, "snap_ck32_w", np);
  launch(ctx, "snap_ui_fp32", [&] {
    switch (P.J) {
      case 0: launch_ui_rows<0>(ctx, P, R, U, ab, w); break;
      case 1: launch_ui_rows<1>(ctx, P, R, U, ab, w); break;
      case 2: launch_ui_rows<2>(ctx, P, R, U, ab, w); break;
      case 3: launch_ui_rows<3>(ctx, P, R, U, ab, w); break;
      case 4: launch_ui_rows<4>(ctx, P, R, U, ab, w); break;
      case 5: launch_ui_rows<5>(ctx, P, R, U, ab, w); break;
      case 6: launch_ui_rows<6>(ctx, P, R, U, ab, w); break;
      case 7: launch_ui_rows<7>(ctx, P, R, U, ab, w); break;
      default: launch_ui_rows<8>(ctx, P, R, U, ab, w); break;
    }
  });
}

// ---------------------------------------------------------------------------
// K3 in FP32: the same dE_i/dr_ij by the reverse-mode sweep of
// ipg_wigner_adj.h.  Phi = 2 sum_half w(e) Re(conj(u_e) Y_i,e) and its
// adjoints with respect to the Cayley-Klein parameters (a, b) come from one
// backward pass through the recursion (no du/dr rows); then
//   dE_i/dr_ij = dsfac r_hat Phi + sfac grad_r Phi,
// with grad_r Phi = Re(a_bar) da_r + Im(a_bar) da_i + Re(b_bar) db_r +
// Im(b_bar) db_i from the FP64 Jacobian of (a, b).  One thread per pair.
// ---------------------------------------------------------------------------
// Seeds of the sweep: Y_i,e in the atom-blocked layout, already weighted by
// the Z/Y kernel.  order() launders a zero offset through an empty asm that
// depends on x, so a level's reads cannot be hoisted above the point where
// the sweep reaches it; tie() does the same for a value, so a row's
// recursion starts after the previous row.
struct YSeed {
  static constexpr bool weighted = true;
  const float2* y;
  __device__ __forceinline__ int order(float x) const {
    int z = 0;
    asm volatile("; order %1" : "+v"(z) : "v"(x));
    return z;
  }
  __device__ __forceinline__ float tie(float v, float x) const {
    asm volatile("; tie %1" : "+v"(v) : "v"(x));
    return v;
  }
  template <int jl, int mb, int ma>
  __device__ __forceinline__ float2 operator()(IC<jl>, IC<mb>, IC<ma>, int tok) const {
    return y[32 * (ub_of(jl) + (jl + 1) * mb + ma) + tok];
  }
};

template <int J>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_snap_dedr_rev32(SnapDev P, SnapRun R, const float* __restrict__ Y,
                                                         double* __restrict__ dedr) {
  constexpr int IDXU = ub_of(J + 1);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (R.d_np ? *R.d_np : R.npairs)) return;
  const int32_t i = R.center[q], j = R.nbr[q];
  const int si = snap_species(R.species, i, P.nspecies), sj = snap_species(R.species, j, P.nspecies);
  const double x = R.rvec[3 * q], y = R.rvec[3 * q + 1], z = R.rvec[3 * q + 2];
  const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
  CK<double> d;
  ck64<true>(x, y, z, rcut, P.weight[sj], P, d);
  const wadj::AB<float> c{(float)d.a_r, (float)d.a_i, (float)d.b_r, (float)d.b_i};
  // the FP64 Jacobian rows wait in LDS while the sweep runs:
  // [k][thread] for k = (d a_r, d a_i, d b_r, d b_i) x 3, r_hat x 3, sfac, dsfac
  __shared__ double jac[17][256];
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
  // the atom-blocked Y of the Z/Y kernel: [i/32][jju][i%32]
  YSeed seed{reinterpret_cast<const float2*>(Y) + ((int64_t)(i >> 5) * IDXU * 32 + (i & 31))};
  wadj::Acc acc;
  wadj::vjp<float, J>(c, seed, acc);
  const double g0 = jac[16][t] * acc.phi, sf = jac[15][t];
#pragma unroll
  for (int k = 0; k < 3; k++)
    dedr[3 * q + k] = g0 * jac[12 + k][t] + sf * (acc.ar * jac[k][t] + acc.ai * jac[3 + k][t] +
                                                  acc.br * jac[6 + k][t] + acc.bi * jac[9 + k][t]);
}

void snap_launch_dedr_rev32(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const float* Y, double* dedr) {
  if (R.npairs == 0) return;
  IPG_REQUIRE(P.J <= 8, "internal: FP32 force kernel compiled for 2J <= 8");
  const unsigned g = grid_for(R.npairs, 256);
  launch(ctx, "snap_dedr_fp32", [&] {
    switch (P.J) {
      case 0: k_snap_dedr_rev32<0><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr); break;
      case 1: k_snap_dedr_rev32<1><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr); break;
      case 2: k_snap_dedr_rev32<2><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr); break;
      case 3: k_snap_dedr_rev32<3><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr); break;
      case 4: k_snap_dedr_rev32<4><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr); break;
      case 5: k_snap_dedr_rev32<5><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr); break;
      case 6: k_snap_dedr_rev32<6><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr); break;
      case 7: k_snap_dedr_rev32<7><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr); break;
      default: k_snap_dedr_rev32<8><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr); break;
    }
  });
}

}  // namespace ipg
