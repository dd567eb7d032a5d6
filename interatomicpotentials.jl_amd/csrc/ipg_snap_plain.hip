// ipg_snap_plain.hip -- the plain FP64 descriptor-derivative path: per-pair
// dB_i/dr_ij (compute_dbidrj, compute_b.jl:166-404) for every coefficient,
// as the reference's compute_all_descriptors forms it (snap.jl:250-265).
//
// One wave per center atom.  Its Z (from the Z/B kernel) is staged in LDS.
// For each neighbor the wave runs the U/dU recursion level by level, with
// lanes over the left-half entries of a level (compute_duarray,
// compute_ui.jl:186-326), keeps u and du of all left-half rows in LDS,
// applies the switching-function scaling (:300-325), then lanes over the
// (jjb, k) outputs form the three half-sums of Re(conj(du) Z).
#include "ipg_snap_kernels.h"

namespace ipg {

__device__ __forceinline__ void mirror8(const double* __restrict__ src, bool neg, double* v) {
  // (-1)^p conj(src) for u and the three du components (re, im pairs)
#pragma unroll
  for (int t = 0; t < 4; t++) {
    v[2 * t] = neg ? -src[2 * t] : src[2 * t];
    v[2 * t + 1] = neg ? src[2 * t + 1] : -src[2 * t + 1];
  }
}

__global__ __launch_bounds__(64) void k_snap_dbdr_plain(SnapDev P, SnapRun R, const double* __restrict__ Zg,
                                                        double* __restrict__ dblist) {
  extern __shared__ __align__(16) double sm[];
  double* Z = sm;                     // 2 * idxz_max
  double* st = Z + 2 * P.idxz_max;    // 8 doubles per left-half entry: u(re,im), du_k(re,im) k=0..2
  const int lane = threadIdx.x;
  const int64_t a = blockIdx.x;
  const double* zg = Zg + a * (int64_t)P.idxz_max * 2;
  for (int e = lane; e < 2 * P.idxz_max; e += 64) Z[e] = zg[e];
  const int si = R.species ? R.species[a] : 0;
  const int nst = P.hoff[P.J + 1];
  __syncthreads();
  for (int64_t q = R.off[a]; q < R.off[a + 1]; q++) {
    const int32_t jn = R.nbr[q];
    const int sj = R.species ? R.species[jn] : 0;
    const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
    const double* rv = R.rvec + 3 * q;
    CK<double> c;
    cayley_klein<double, true>(rv[0], rv[1], rv[2], rcut, P.weight[sj], P, c);
    if (lane == 0) {
      st[0] = 1.0;
      for (int t = 1; t < 8; t++) st[t] = 0.0;
    }
    __syncthreads();
    for (int jl = 1; jl <= P.J; jl++) {
      const int w = jl + 1, nj = (jl / 2 + 1) * w;
      const int ho = P.hoff[jl], hp = P.hoff[jl - 1];
      for (int e = lane; e < nj; e += 64) {
        const int mb = e / w, ma = e % w;
        double pm1[8], p0[8];  // level jl-1, row mb, columns ma-1 and ma
        const bool mir = 2 * mb > jl - 1;
        if (ma >= 1) {
          const int x = ma - 1;
          if (!mir) {
            for (int t = 0; t < 8; t++) pm1[t] = st[(hp + mb * jl + x) * 8 + t];
          } else {
            mirror8(st + (hp + (mb - 1) * jl + (jl - 1 - x)) * 8, (mb + x) & 1, pm1);
          }
        }
        if (ma <= jl - 1) {
          const int x = ma;
          if (!mir) {
            for (int t = 0; t < 8; t++) p0[t] = st[(hp + mb * jl + x) * 8 + t];
          } else {
            mirror8(st + (hp + (mb - 1) * jl + (jl - 1 - x)) * 8, (mb + x) & 1, p0);
          }
        }
        double nv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (ma >= 1) {
          const double rb = sqrt((double)ma / (double)(jl - mb));
          const double pr = pm1[0], pi = pm1[1];
          nv[0] = -rb * (c.b_r * pr + c.b_i * pi);
          nv[1] = -rb * (c.b_r * pi - c.b_i * pr);
          for (int k = 0; k < 3; k++) {
            const double dr = pm1[2 + 2 * k], di = pm1[3 + 2 * k];
            nv[2 + 2 * k] = -rb * (c.db_r[k] * pr + c.db_i[k] * pi + c.b_r * dr + c.b_i * di);
            nv[3 + 2 * k] = -rb * (c.db_r[k] * pi - c.db_i[k] * pr + c.b_r * di - c.b_i * dr);
          }
        }
        if (ma <= jl - 1) {
          const double ra = sqrt((double)(jl - ma) / (double)(jl - mb));
          const double pr = p0[0], pi = p0[1];
          nv[0] += ra * (c.a_r * pr + c.a_i * pi);
          nv[1] += ra * (c.a_r * pi - c.a_i * pr);
          for (int k = 0; k < 3; k++) {
            const double dr = p0[2 + 2 * k], di = p0[3 + 2 * k];
            nv[2 + 2 * k] += ra * (c.da_r[k] * pr + c.da_i[k] * pi + c.a_r * dr + c.a_i * di);
            nv[3 + 2 * k] += ra * (c.da_r[k] * pi - c.da_i[k] * pr + c.a_r * di - c.a_i * dr);
          }
        }
        for (int t = 0; t < 8; t++) st[(ho + e) * 8 + t] = nv[t];
      }
      __syncthreads();
      if ((jl & 1) == 0) {
        // middle row of an even level (compute_ui.jl:270-297)
        const int mbm = jl / 2;
        const int e = lane + mbm + 1;  // columns mbm+1 .. jl
        if (e <= jl) {
          double v[8];
          mirror8(st + (ho + mbm * w + (2 * mbm - e)) * 8, (mbm + e) & 1, v);
          for (int t = 0; t < 8; t++) st[(ho + mbm * w + e) * 8 + t] = v[t];
        }
        if (lane == 63) {
          double* cen = st + (ho + mbm * w + mbm) * 8;
          for (int t = 1; t < 8; t += 2) cen[t] = -cen[t];
        }
        __syncthreads();
      }
    }
    // du <- dsfac u uhat + sfac du (sfac, dsfac include w_j)
    for (int e = lane; e < nst; e += 64) {
      double* s = st + e * 8;
      const double ur = s[0], ui = s[1];
      for (int k = 0; k < 3; k++) {
        s[2 + 2 * k] = c.dsfac * ur * c.u[k] + c.sfac * s[2 + 2 * k];
        s[3 + 2 * k] = c.dsfac * ui * c.u[k] + c.sfac * s[3 + 2 * k];
      }
    }
    __syncthreads();
    for (int o = lane; o < 3 * P.idxb_max; o += 64) {
      const int jjb = o / 3, k = o % 3;
      const BTerm bt = P.bterm[jjb];
      double tot = 0.0;
      for (int t = 0; t < 3; t++) {
        const int l = bt.l[t];
        int e = P.hoff[l], jjz = bt.jjz0[t];
        double sum = 0.0;
        for (int mb = 0; 2 * mb < l; mb++)
          for (int ma = 0; ma <= l; ma++, e++, jjz++)
            sum += st[e * 8 + 2 + 2 * k] * Z[2 * jjz] + st[e * 8 + 3 + 2 * k] * Z[2 * jjz + 1];
        if ((l & 1) == 0) {
          const int mb = l / 2;
          for (int ma = 0; ma < mb; ma++, e++, jjz++)
            sum += st[e * 8 + 2 + 2 * k] * Z[2 * jjz] + st[e * 8 + 3 + 2 * k] * Z[2 * jjz + 1];
          sum += (st[e * 8 + 2 + 2 * k] * Z[2 * jjz] + st[e * 8 + 3 + 2 * k] * Z[2 * jjz + 1]) * 0.5;
        }
        tot += bt.fac[t] * sum;
      }
      dblist[(q * P.ncoeff + jjb) * 3 + k] = 2.0 * tot;
    }
    __syncthreads();
  }
}

void snap_launch_dbdr_plain(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const double* Z, double* dblist) {
  if (R.n == 0 || R.npairs == 0) return;
  size_t sm = sizeof(double) * (2 * (size_t)P.idxz_max + 8 * (size_t)P.hoff[P.J + 1]);
  launch(ctx, "snap_dbdr_fp64", [&] {
    if (sm > 64 * 1024)
      IPG_HIP(hipFuncSetAttribute((const void*)k_snap_dbdr_plain, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)sm));
    k_snap_dbdr_plain<<<(unsigned)R.n, 64, sm, ctx->stream>>>(P, R, Z, dblist);
  });
}

template <typename T>
void snap_launch_ycoef(ipg_ctx* ctx, const SnapDev& P, const double* beta, T* ycoef) {
  const int nt = P.nspecies * P.nblk;
  launch(ctx, "snap_ycoef", [&] { k_snap_ycoef<T><<<grid_for(nt, 256), 256, 0, ctx->stream>>>(P, beta, ycoef); });
}
template void snap_launch_ycoef<float>(ipg_ctx*, const SnapDev&, const double*, float*);
template void snap_launch_ycoef<double>(ipg_ctx*, const SnapDev&, const double*, double*);

}  // namespace ipg
