// ipg_snap_dbdr.hip -- per-pair dB_i/dr_ij for every coefficient
// (compute_dbidrj, compute_b.jl:104-342, after compute_duidrj,
// compute_ui.jl:186-347), the plain route of compute_all_descriptors
// (snap.jl:250-265).
//
// One block per center atom stages the atom's Z in LDS.  Its lanes take
// (pair, component) items: each lane runs the Wigner u / du row chain of its
// pair in registers with the one derivative component it owns, scales du as
// compute_duarray's tail does (:300-325), and, for every emitted left-half
// entry of level l, adds the entry's share to every coefficient whose
// half-sums use level l.  Which coefficients those are, the Z entry each one
// reads and its factor are compile-time tables, so the inner work is a fixed
// sequence of LDS reads at constant offsets and FMAs into per-lane
// accumulators.
//
// dB_(j1,j2,j) = 2 [ S(j; Z(j1,j2,j)) + (j+1)/(j1+1) S(j1; Z(j,j2,j1))
//                   + (j+1)/(j2+1) S(j2; Z(j,j1,j2)) ]
// with S(l; Z) the half-column sum of Re(conj(dU_l) Z) (compute_b.jl:36-62
// rule: rows mb < l/2, then the middle row up to its centre at weight 1/2).
#include "ipg_snap_kernels.h"

namespace ipg {
namespace dbdr {

constexpr int cmin(int a, int b) { return a < b ? a : b; }

// Z block start (jjz) of block (a, c, l), a >= c, in the table order of
// types/prebuilt_arrays.jl (j1 >= j2, j = j1-j2 .. min(J, j1+j2) step 2).
constexpr int zblock(int J, int a, int c, int l) {
  int off = 0;
  for (int x = 0; x <= J; x++)
    for (int y = 0; y <= x; y++)
      for (int z = x - y; z <= cmin(J, x + y); z += 2) {
        if (x == a && y == c && z == l) return off;
        off += (z / 2 + 1) * (z + 1);
      }
  return -1;
}
constexpr int ncoeff(int J) {
  int n = 0;
  for (int x = 0; x <= J; x++)
    for (int y = 0; y <= x; y++)
      for (int z = x - y; z <= cmin(J, x + y); z += 2)
        if (z >= x) n++;
  return n;
}
constexpr int idxz_max(int J) {
  int n = 0;
  for (int x = 0; x <= J; x++)
    for (int y = 0; y <= x; y++)
      for (int z = x - y; z <= cmin(J, x + y); z += 2) n += (z / 2 + 1) * (z + 1);
  return n;
}

// The (coefficient, Z block, factor) uses of each level l.
template <int J>
struct Uses {
  static constexpr int MAXU = 3 * 64;
  int n[J + 1] = {};
  int jjb[J + 1][MAXU] = {};
  int zb[J + 1][MAXU] = {};
  double fac[J + 1][MAXU] = {};
  constexpr Uses() {
    int b = 0;
    for (int j1 = 0; j1 <= J; j1++)
      for (int j2 = 0; j2 <= j1; j2++)
        for (int j = j1 - j2; j <= cmin(J, j1 + j2); j += 2) {
          if (j < j1) continue;
          add(j, b, zblock(J, j1, j2, j), 2.0);
          add(j1, b, zblock(J, j, j2, j1), 2.0 * (j + 1) / (j1 + 1));
          add(j2, b, zblock(J, j, j1, j2), 2.0 * (j + 1) / (j2 + 1));
          b++;
        }
  }
  constexpr void add(int l, int b, int z, double f) {
    jjb[l][n[l]] = b;
    zb[l][n[l]] = z;
    fac[l][n[l]] = f;
    n[l]++;
  }
};

template <typename S>
__device__ __forceinline__ S pick3(const S (&v)[3], int k) {
  return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]);
}

}  // namespace dbdr

constexpr int DBDR_THREADS = 128;

// Occupancy: 3 waves per SIMD in FP32 and 2 in FP64.  A few accumulators
// then spill to scratch, which measured faster than the spill-free 2 and 1
// waves (and than 4 and 3, where the spills grow).
template <int J, typename S>
__global__ __launch_bounds__(DBDR_THREADS, sizeof(S) == 4 ? 3 : 2) void k_snap_dbdr(SnapDev P, SnapRun R, const S* __restrict__ Zg,
                                                            S* __restrict__ dblist) {
  constexpr int NZ = dbdr::idxz_max(J);
  constexpr int NB = dbdr::ncoeff(J);
  constexpr dbdr::Uses<J> U{};
  __shared__ S Zs[2 * NZ];
  const int64_t a = blockIdx.x;
  const S* zg = Zg + a * (int64_t)NZ * 2;
  for (int e = threadIdx.x; e < 2 * NZ; e += blockDim.x) Zs[e] = zg[e];
  __syncthreads();
  const int64_t q0 = R.off[a], q1 = R.off[a + 1];
  const int si = snap_species(R.species, a, P.nspecies);
  for (int64_t it = threadIdx.x; it < 3 * (q1 - q0); it += blockDim.x) {
    const int64_t q = q0 + it / 3;
    const int k = (int)(it % 3);
    const int32_t jn = R.nbr[q];
    const int sj = snap_species(R.species, jn, P.nspecies);
    typename WideOf<S>::type x, y, z;
    load_pair<typename WideOf<S>::type>(R, q, (int32_t)a, jn, x, y, z);
    const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
    CK<S> c;
    cayley_klein<S, true>(x, y, z, rcut, P.weight[sj], P, c);
    // the derivative along component k only, moved to slot 0 of the chain
    const S uk = dbdr::pick3(c.u, k);
    c.da_r[0] = dbdr::pick3(c.da_r, k);
    c.da_i[0] = dbdr::pick3(c.da_i, k);
    c.db_r[0] = dbdr::pick3(c.db_r, k);
    c.db_i[0] = dbdr::pick3(c.db_i, k);
    S acc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) acc[b] = S(0);
    wigner_chain<J, S, 1>(c, [&](auto jc, auto mbc, auto mac, auto hc, auto, S u_r, S u_i, auto& dur, auto& dui) {
      constexpr int l = decltype(jc)::value, mb = decltype(mbc)::value, ma = decltype(mac)::value;
      constexpr double hw = decltype(hc)::value ? 0.5 : 1.0;
      // du <- dsfac u uhat + sfac du (sfac, dsfac include w_j)
      const S dr = c.dsfac * u_r * uk + c.sfac * dur[0][ma];
      const S di = c.dsfac * u_i * uk + c.sfac * dui[0][ma];
      static_for<0, U.n[l]>([&](auto uc) {
        constexpr int t = decltype(uc)::value;
        constexpr int e = U.zb[l][t] + mb * (l + 1) + ma;
        constexpr S f = (S)(U.fac[l][t] * hw);
        acc[U.jjb[l][t]] += f * (dr * Zs[2 * e] + di * Zs[2 * e + 1]);
      });
      // keep each entry's Z loads next to their use: the chain is one basic
      // block, and unbounded load hoisting overflows the register file
      __builtin_amdgcn_sched_barrier(0);
    });
    S* out = dblist + q * (int64_t)NB * 3 + k;
#pragma unroll
    for (int b = 0; b < NB; b++) out[3 * b] = acc[b];
  }
}

template <int J, typename S>
static void launch_dbdr_j(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const S* Z, S* dblist) {
  const int threads = R.dbdr_threads == 64 ? 64 : DBDR_THREADS;  // lanes loop over the (pair, component) items
  k_snap_dbdr<J, S><<<(unsigned)R.n, threads, 0, ctx->stream>>>(P, R, Z, dblist);
}

template <typename S>
void snap_launch_dbdr(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const S* Z, S* dblist) {
  if (R.n == 0 || R.npairs == 0) return;
  IPG_REQUIRE(!P.bnorm, "internal: bnorm is not supported by the dB kernel");
  launch(ctx, sizeof(S) == 4 ? "snap_dbdr_fp32" : "snap_dbdr_fp64", [&] {
    switch (P.J) {
      case 0: launch_dbdr_j<0, S>(ctx, P, R, Z, dblist); break;
      case 1: launch_dbdr_j<1, S>(ctx, P, R, Z, dblist); break;
      case 2: launch_dbdr_j<2, S>(ctx, P, R, Z, dblist); break;
      case 3: launch_dbdr_j<3, S>(ctx, P, R, Z, dblist); break;
      case 4: launch_dbdr_j<4, S>(ctx, P, R, Z, dblist); break;
      case 5: launch_dbdr_j<5, S>(ctx, P, R, Z, dblist); break;
      case 6: launch_dbdr_j<6, S>(ctx, P, R, Z, dblist); break;
      case 7: launch_dbdr_j<7, S>(ctx, P, R, Z, dblist); break;
      default: launch_dbdr_j<8, S>(ctx, P, R, Z, dblist); break;
    }
  });
}
template void snap_launch_dbdr<double>(ipg_ctx*, const SnapDev&, const SnapRun&, const double*, double*);
template void snap_launch_dbdr<float>(ipg_ctx*, const SnapDev&, const SnapRun&, const float*, float*);

}  // namespace ipg
