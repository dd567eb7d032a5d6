// ipg_snap_gen.hip -- SNAP over runtime index tables, in FP64: the
// multi-element form (chem_flag = true, which forces bnorm_flag) and
// 2J = 9, 10, the configurations the compile-time kernels leave out.
//
// One wave per atom.  The Wigner recursion of one neighbor (and, for dB/dr,
// its derivative) is spread over the wave's lanes level by level, the
// previous level held in LDS; the lanes own disjoint entries, so every sum
// runs in the reference's order and the results do not depend on timing.
//
// The multi-element arithmetic is the reference's as written:
//  - U is accumulated per neighbor element, wself on the center atom's
//    element block (compute_ui.jl:5-44, :131-175);
//  - compute_zi (compute_z.jl:5-62) zeroes the Z block of each element pair
//    but accumulates every pair into the first block, dividing it by j+1
//    after each pair (bnorm), so only that block is non-zero: Z0 =
//    ((Z_00/(j+1) + Z_01)/(j+1) + ...)/(j+1);
//  - compute_bi (compute_b.jl:5-95) then gives B = 2 sum U_e3 . Z0 for the
//    triples (0, 0, e3), zero for the others, and the bzero shift of its
//    last element loop only (each pass recomputes blist);
//  - compute_dbidrj (compute_b.jl:104-342) with Z0 fills the triples
//    (0, 0, e3), (e3, 0, 0) and (0, e3, 0) for a neighbor of element e3.
#include "ipg_snap_driver.h"
#include "ipg_snap_kernels.h"

namespace ipg {

constexpr int GEN_LANES = 64;

// Cayley-Klein parameters of compute_ui.jl:25-29 with z0 = r / tan(theta0)
// (compute_ui.jl:23-28, compute_uarray :50-60).
struct GenCK {
  double a_r, a_i, b_r, b_i, r;
};
__device__ __forceinline__ GenCK gen_ck_u(const SnapDev& P, double x, double y, double z, double rcut) {
  GenCK c;
  const double rsq = x * x + y * y + z * z;
  const double r = sqrt(rsq);
  const double theta0 = (r - P.rmin0) * P.rfac0 * M_PI / (rcut - P.rmin0);
  const double z0 = r / tan(theta0);
  const double r0inv = 1.0 / sqrt(r * r + z0 * z0);
  c.a_r = r0inv * z0;
  c.a_i = -r0inv * z;
  c.b_r = r0inv * y;
  c.b_i = -r0inv * x;
  c.r = r;
  return c;
}

// compute_sfac (compute_ui.jl:350-364).
__device__ __forceinline__ double gen_sfac(const SnapDev& P, double r, double rcut) {
  if (!P.switch_flag) return 1.0;
  if (r <= P.rmin0) return 1.0;
  if (r > rcut) return 0.0;
  const double rcutfac = M_PI / (rcut - P.rmin0);
  return 0.5 * (cos((r - P.rmin0) * rcutfac) + 1.0);
}

__device__ __forceinline__ double gen_rootpq(const SnapDev& P, int p, int q) { return P.rootpq[p * (P.J + 1) + q]; }

// Right half of level j by the inversion symmetry of compute_ui.jl:103-122:
// u(j-mb, j-ma) = (-1)^(mb+ma) conj(u(mb, ma)) for the rows 2 mb < j; on
// the middle row of an even level the reference's in-place copy leaves the
// left entries, mirrors them onto the right and conjugates the centre.
// NK = 0: U only; NK = 3: the derivative rows dr/di [jju][3] as well
// (compute_ui.jl:292-310, same pattern).
template <int NK>
__device__ __forceinline__ void gen_mirror(int j, int base, int lane, double* ur, double* ui, double* dr,
                                           double* di) {
  const int w = j + 1;
  const int nl = (j / 2 + 1) * w;
  for (int t = lane; t < nl; t += GEN_LANES) {
    const int mb = t / w, ma = t % w;
    const int src = base + t;
    const int dst = base + (j - mb) * w + (j - ma);
    const bool pos = ((mb + ma) & 1) == 0;
    if (2 * mb < j || 2 * ma < j) {
      ur[dst] = pos ? ur[src] : -ur[src];
      ui[dst] = pos ? -ui[src] : ui[src];
#pragma unroll
      for (int k = 0; k < NK; k++) {
        dr[3 * dst + k] = pos ? dr[3 * src + k] : -dr[3 * src + k];
        di[3 * dst + k] = pos ? -di[3 * src + k] : di[3 * src + k];
      }
    } else if (2 * ma == j) {  // centre of the middle row: conj in place
      ui[src] = -ui[src];
#pragma unroll
      for (int k = 0; k < NK; k++) di[3 * src + k] = -di[3 * src + k];
    }
    // middle row, 2 ma > j: the reference writes these left entries back
    // from their mirror images unchanged
  }
}

// compute_uarray (compute_ui.jl:50-129) into ur/ui (LDS, idxu_max each).
// Entry (j, mb, ma) of the left half is -sqrt(ma/(j-mb)) b u(j-1, mb, ma-1)
// (assigned at step ma-1) plus sqrt((j-ma)/(j-mb)) conj(a) u(j-1, mb, ma)
// (added at step ma), in that order.
__device__ void gen_uarray(const SnapDev& P, const GenCK& c, double* ur, double* ui, int lane) {
  if (lane == 0) {
    ur[0] = 1.0;
    ui[0] = 0.0;
  }
  __syncthreads();
  for (int j = 1; j <= P.J; j++) {
    const int jju = P.idxu_block[j], jjup = P.idxu_block[j - 1];
    const int w = j + 1, nl = (j / 2 + 1) * w;
    for (int t = lane; t < nl; t += GEN_LANES) {
      const int mb = t / w, ma = t % w;
      double vr = 0.0, vi = 0.0;
      if (ma > 0) {
        const double rp = gen_rootpq(P, ma, j - mb);
        const double pr = ur[jjup + j * mb + ma - 1], pi = ui[jjup + j * mb + ma - 1];
        vr = -rp * (c.b_r * pr + c.b_i * pi);
        vi = -rp * (c.b_r * pi - c.b_i * pr);
      }
      if (ma < j) {
        const double rp = gen_rootpq(P, j - ma, j - mb);
        const double pr = ur[jjup + j * mb + ma], pi = ui[jjup + j * mb + ma];
        vr = vr + rp * (c.a_r * pr + c.a_i * pi);
        vi = vi + rp * (c.a_r * pi - c.a_i * pr);
      }
      ur[jju + t] = vr;
      ui[jju + t] = vi;
    }
    __syncthreads();
    gen_mirror<0>(j, jju, lane, ur, ui, nullptr, nullptr);
    __syncthreads();
  }
}

// U per element: utot[i][e][jju][2] = wself delta(ma, mb) on the center's
// element block (every block with wselfall_flag) plus sum over neighbors of
// element e of sfac w_j u(r_ij) (compute_ui.jl:5-44, :131-175).
__global__ __launch_bounds__(GEN_LANES) void k_gen_u(SnapDev P, SnapRun R, double* __restrict__ utot) {
  extern __shared__ double sh[];
  const int nu = P.idxu_max;
  double* ur = sh;
  double* ui = sh + nu;
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const int si = snap_species(R.species, i, P.nspecies);
  const int ielem = P.chem ? si : 0;
  double* ut = utot + (size_t)i * P.nel * nu * 2;
  for (int e = 0; e < P.nel; e++)
    for (int k = lane; k < nu; k += GEN_LANES) {
      const int m = P.jmm[k];
      const bool diag = ((m >> 8) & 255) == (m >> 16);
      ut[(e * nu + k) * 2] = (diag && (e == ielem || P.wselfall)) ? P.wself : 0.0;
      ut[(e * nu + k) * 2 + 1] = 0.0;
    }
  for (int64_t q = R.off[i]; q < R.off[i + 1]; q++) {
    const int32_t j = R.nbr[q];
    const int sj = snap_species(R.species, j, P.nspecies);
    const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
    const GenCK c = gen_ck_u(P, R.rvec[3 * q], R.rvec[3 * q + 1], R.rvec[3 * q + 2], rcut);
    gen_uarray(P, c, ur, ui, lane);
    double sfac = gen_sfac(P, c.r, rcut);
    sfac *= P.weight[sj];
    const int e = P.chem ? sj : 0;
    for (int k = lane; k < nu; k += GEN_LANES) {
      ut[(e * nu + k) * 2] += sfac * ur[k];
      ut[(e * nu + k) * 2 + 1] += sfac * ui[k];
    }
    __syncthreads();  // the next neighbor overwrites ur/ui
  }
}

// Z of the first element pair as compute_zi leaves it (see the file head),
// then B (compute_bi) and optionally E_i = beta_s(i) . B_i.  One wave per
// atom; lane per Z entry, then lane per (e3, jjb).
__global__ __launch_bounds__(GEN_LANES) void k_gen_zb(SnapDev P, SnapRun R, const double* __restrict__ utot,
                                                     double* __restrict__ Z, double* __restrict__ B,
                                                     const double* __restrict__ beta, double* __restrict__ eatom) {
  __shared__ double epart[GEN_LANES];
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const int nu = P.idxu_max, nel = P.nel;
  const double* ut = utot + (size_t)i * nel * nu * 2;
  double* zi = Z + (size_t)i * P.idxz_max * 2;
  for (int jjz = lane; jjz < P.idxz_max; jjz += GEN_LANES) {
    const ZEnt zt = P.zent[jjz];
    const double* cg = P.cg_d + zt.cgoff;
    double z_r = 0.0, z_i = 0.0;
    for (int e1 = 0; e1 < nel; e1++)
      for (int e2 = 0; e2 < nel; e2++) {
        const double* u1 = ut + (size_t)e1 * nu * 2;
        const double* u2 = ut + (size_t)e2 * nu * 2;
        int jju1 = P.idxu_block[zt.j1] + (zt.j1 + 1) * zt.mb1min;
        int jju2 = P.idxu_block[zt.j2] + (zt.j2 + 1) * zt.mb2max;
        int icgb = zt.mb1min * (zt.j2 + 1) + zt.mb2max;
        for (int ib = 0; ib < zt.nb; ib++) {
          double s_r = 0.0, s_i = 0.0;
          int ma1 = zt.ma1min, ma2 = zt.ma2max;
          int icga = zt.ma1min * (zt.j2 + 1) + zt.ma2max;
          for (int ia = 0; ia < zt.na; ia++) {
            const double a_r = u1[2 * (jju1 + ma1)], a_i = u1[2 * (jju1 + ma1) + 1];
            const double b_r = u2[2 * (jju2 + ma2)], b_i = u2[2 * (jju2 + ma2) + 1];
            s_r += cg[icga] * (a_r * b_r - a_i * b_i);
            s_i += cg[icga] * (a_r * b_i + a_i * b_r);
            ma1++;
            ma2--;
            icga += zt.j2;
          }
          z_r += cg[icgb] * s_r;
          z_i += cg[icgb] * s_i;
          jju1 += zt.j1 + 1;
          jju2 -= zt.j2 + 1;
          icgb += zt.j2;
        }
        if (P.bnorm) {
          z_r /= (zt.j + 1);
          z_i /= (zt.j + 1);
        }
      }
    zi[2 * jjz] = z_r;
    zi[2 * jjz + 1] = z_i;
  }
  __syncthreads();  // Z of this atom, written by the wave's lanes
  const int nb = P.idxb_max, ntrip = nel * nel * nel;
  const int si = snap_species(R.species, i, P.nspecies);
  const double* bt = beta ? beta + (size_t)si * P.ncoeff : nullptr;
  const int last = (nel - 1) * (nel * nel + nel + 1);  // bzero triple of the last element pass
  double esum = 0.0;
  for (int t = lane; t < ntrip * nb; t += GEN_LANES) {
    const int it = t / nb, jjb = t % nb;
    const int j1 = P.idxb3[3 * jjb], j2 = P.idxb3[3 * jjb + 1], j = P.idxb3[3 * jjb + 2];
    double b = 0.0;
    if (it < nel) {  // (0, 0, e3 = it): Z0 against U_e3
      const double* u = ut + (size_t)it * nu * 2;
      int jjz = P.idxz_block[(j1 * (P.J + 1) + j2) * (P.J + 1) + j];
      int jju = P.idxu_block[j];
      double sumzu = 0.0;
      for (int mb = 0; 2 * mb < j; mb++)
        for (int ma = 0; ma <= j; ma++) {
          sumzu += u[2 * jju] * zi[2 * jjz] + u[2 * jju + 1] * zi[2 * jjz + 1];
          jjz++;
          jju++;
        }
      if (j % 2 == 0) {
        const int mb = j / 2;
        for (int ma = 0; ma < mb; ma++) {
          sumzu += u[2 * jju] * zi[2 * jjz] + u[2 * jju + 1] * zi[2 * jjz + 1];
          jjz++;
          jju++;
        }
        sumzu += 0.5 * (u[2 * jju] * zi[2 * jjz] + u[2 * jju + 1] * zi[2 * jjz + 1]);
      }
      b = 2.0 * sumzu;
    }
    if (P.bzero && (P.wselfall || it == last)) b -= P.bzero_j[j];
    if (B) B[(size_t)i * P.ncoeff + t] = b;
    if (bt) esum += bt[t] * b;
  }
  if (!eatom) return;
  // E_i: each lane's terms in coefficient order, then the lanes in order
  epart[lane] = esum;
  __syncthreads();
  if (lane == 0) {
    double e = 0.0;
    for (int k = 0; k < GEN_LANES; k++) e += epart[k];
    eatom[i] = e;
  }
}

// dB_i/dr_ij of every pair of atom i: U and dU of the pair (compute_ui.jl:
// 186-347; the derivative recursion takes its Cayley-Klein parameters from
// z0 = r cos/sin as compute_duidrj does, the U rows from compute_uarray's
// z0 = r / tan), then the three half-sums of compute_b.jl:145-337 with the
// center's Z0, written to the full row [ncoeff][3] of the pair.
__global__ __launch_bounds__(GEN_LANES) void k_gen_dbdr(SnapDev P, SnapRun R, const double* __restrict__ Z,
                                                       double* __restrict__ dblist) {
  extern __shared__ double sh[];
  const int nu = P.idxu_max, nb = P.idxb_max, nel = P.nel;
  double* ur = sh;
  double* ui = ur + nu;
  double* dr = ui + nu;      // [jju][3]
  double* di = dr + 3 * nu;  // [jju][3]
  double* sterm = di + 3 * nu;  // [jjb][term][3]
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const int si = snap_species(R.species, i, P.nspecies);
  const double* zi = Z + (size_t)i * P.idxz_max * 2;
  const int jd = P.J + 1;
  for (int64_t q = R.off[i]; q < R.off[i + 1]; q++) {
    const int32_t j = R.nbr[q];
    const int sj = snap_species(R.species, j, P.nspecies);
    const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
    const double x = R.rvec[3 * q], y = R.rvec[3 * q + 1], z = R.rvec[3 * q + 2];
    const GenCK cu = gen_ck_u(P, x, y, z, rcut);
    gen_uarray(P, cu, ur, ui, lane);
    CK<double> c;
    cayley_klein_impl<double, true>(x, y, z, rcut, P.weight[sj], P, c);
    // compute_duarray (compute_ui.jl:186-326)
    if (lane < 3) {
      dr[lane] = 0.0;
      di[lane] = 0.0;
    }
    __syncthreads();
    for (int jl = 1; jl <= P.J; jl++) {
      const int jju = P.idxu_block[jl], jjup = P.idxu_block[jl - 1];
      const int w = jl + 1, nl = (jl / 2 + 1) * w;
      for (int t = lane; t < nl; t += GEN_LANES) {
        const int mb = t / w, ma = t % w;
        double vr[3] = {0.0, 0.0, 0.0}, vi[3] = {0.0, 0.0, 0.0};
        if (ma > 0) {
          const double rp = gen_rootpq(P, ma, jl - mb);
          const int p = jjup + jl * mb + ma - 1;
          for (int k = 0; k < 3; k++) {
            vr[k] = -rp * (c.db_r[k] * ur[p] + c.db_i[k] * ui[p] + c.b_r * dr[3 * p + k] + c.b_i * di[3 * p + k]);
            vi[k] = -rp * (c.db_r[k] * ui[p] - c.db_i[k] * ur[p] + c.b_r * di[3 * p + k] - c.b_i * dr[3 * p + k]);
          }
        }
        if (ma < jl) {
          const double rp = gen_rootpq(P, jl - ma, jl - mb);
          const int p = jjup + jl * mb + ma;
          for (int k = 0; k < 3; k++) {
            vr[k] = vr[k] + rp * (c.da_r[k] * ur[p] + c.da_i[k] * ui[p] + c.a_r * dr[3 * p + k] + c.a_i * di[3 * p + k]);
            vi[k] = vi[k] + rp * (c.da_r[k] * ui[p] - c.da_i[k] * ur[p] + c.a_r * di[3 * p + k] - c.a_i * dr[3 * p + k]);
          }
        }
        for (int k = 0; k < 3; k++) {
          dr[3 * (jju + t) + k] = vr[k];
          di[3 * (jju + t) + k] = vi[k];
        }
      }
      __syncthreads();
      // the mirror step of compute_ui.jl:292-310 on dU only (U is complete)
      const int nl2 = (jl / 2 + 1) * w;
      for (int t = lane; t < nl2; t += GEN_LANES) {
        const int mb = t / w, ma = t % w;
        const int src = jju + t, dst = jju + (jl - mb) * w + (jl - ma);
        const bool pos = ((mb + ma) & 1) == 0;
        if (2 * mb < jl || 2 * ma < jl) {
          for (int k = 0; k < 3; k++) {
            dr[3 * dst + k] = pos ? dr[3 * src + k] : -dr[3 * src + k];
            di[3 * dst + k] = pos ? -di[3 * src + k] : di[3 * src + k];
          }
        } else if (2 * ma == jl) {
          for (int k = 0; k < 3; k++) di[3 * src + k] = -di[3 * src + k];
        }
      }
      __syncthreads();
    }
    // switching-function scaling of the left-half rows (compute_ui.jl:312-325)
    for (int t = lane; t < nu; t += GEN_LANES) {
      const int m = P.jmm[t];
      const int jl = m & 255, mb = (m >> 8) & 255;
      if (2 * mb > jl) continue;
      for (int k = 0; k < 3; k++) {
        dr[3 * t + k] = c.dsfac * ur[t] * c.u[k] + c.sfac * dr[3 * t + k];
        di[3 * t + k] = c.dsfac * ui[t] * c.u[k] + c.sfac * di[3 * t + k];
      }
    }
    __syncthreads();
    // half-sums of conj(dU) Z0 for the three roles of each triple
    for (int t = lane; t < 3 * nb; t += GEN_LANES) {
      const int jjb = t / 3, role = t % 3;
      const int j1 = P.idxb3[3 * jjb], j2 = P.idxb3[3 * jjb + 1], jj = P.idxb3[3 * jjb + 2];
      int jl, blk;
      if (role == 0) jl = jj, blk = (j1 * jd + j2) * jd + jj;
      else if (role == 1) jl = j1, blk = (jj * jd + j2) * jd + j1;
      else jl = j2, blk = (jj * jd + j1) * jd + j2;
      int jjz = P.idxz_block[blk];
      int jju = P.idxu_block[jl];
      double s[3] = {0.0, 0.0, 0.0};
      for (int mb = 0; 2 * mb < jl; mb++)
        for (int ma = 0; ma <= jl; ma++) {
          for (int k = 0; k < 3; k++) s[k] += dr[3 * jju + k] * zi[2 * jjz] + di[3 * jju + k] * zi[2 * jjz + 1];
          jjz++;
          jju++;
        }
      if (jl % 2 == 0) {
        const int mb = jl / 2;
        for (int ma = 0; ma < mb; ma++) {
          for (int k = 0; k < 3; k++) s[k] += dr[3 * jju + k] * zi[2 * jjz] + di[3 * jju + k] * zi[2 * jjz + 1];
          jjz++;
          jju++;
        }
        for (int k = 0; k < 3; k++) s[k] += (dr[3 * jju + k] * zi[2 * jjz] + di[3 * jju + k] * zi[2 * jjz + 1]) * 0.5;
      }
      for (int k = 0; k < 3; k++) sterm[(jjb * 3 + role) * 3 + k] = s[k];
    }
    __syncthreads();
    // the pair's row: compute_b.jl:207 assigns the first term, :268-270 and
    // :333-335 add the other two (times (j+1)/(j1+1), (j+1)/(j2+1) without
    // bnorm_flag); with chem_flag the three land in (0,0,e3), (e3,0,0) and
    // (0,e3,0)
    const int e3 = P.chem ? sj : 0;
    const int t1 = e3, t2 = e3 * nel * nel, t3 = e3 * nel;
    double* row = dblist + (size_t)q * P.ncoeff * 3;
    for (int t = lane; t < P.ncoeff * 3; t += GEN_LANES) {
      const int k = t % 3, cc = t / 3;
      const int it = cc / nb, jjb = cc % nb;
      double v = 0.0;
      if (it == t1 || it == t2 || it == t3) {
        const int j1 = P.idxb3[3 * jjb], j2 = P.idxb3[3 * jjb + 1], jj = P.idxb3[3 * jjb + 2];
        const double f1 = P.bnorm ? 1.0 : (jj + 1) / (j1 + 1.0);
        const double f2 = P.bnorm ? 1.0 : (jj + 1) / (j2 + 1.0);
        const double* st = sterm + jjb * 9;
        if (it == t1) v = 2.0 * st[k];
        if (it == t2) v += P.bnorm ? 2.0 * st[3 + k] : 2.0 * st[3 + k] * f1;
        if (it == t3) v += P.bnorm ? 2.0 * st[6 + k] : 2.0 * st[6 + k] * f2;
      }
      row[t] = v;
    }
    __syncthreads();  // the next pair overwrites the LDS rows
  }
}

void snap_gen_descriptors(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, double* U, double* Z, double* B,
                          const double* beta, double* eatom) {
  if (R.n == 0) return;
  const unsigned g = (unsigned)R.n;
  const size_t shu = sizeof(double) * 2 * (size_t)P.idxu_max;
  launch(ctx, "snap_gen_u", [&] { k_gen_u<<<g, GEN_LANES, shu, ctx->stream>>>(P, R, U); });
  launch(ctx, "snap_gen_zb", [&] { k_gen_zb<<<g, GEN_LANES, 0, ctx->stream>>>(P, R, U, Z, B, beta, eatom); });
}

void snap_gen_dbdr(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const double* Z, double* dblist) {
  if (R.n == 0) return;
  const size_t sh = sizeof(double) * (8 * (size_t)P.idxu_max + 9 * (size_t)P.idxb_max);
  launch(ctx, "snap_gen_dbdr",
         [&] { k_gen_dbdr<<<(unsigned)R.n, GEN_LANES, sh, ctx->stream>>>(P, R, Z, dblist); });
}

}  // namespace ipg
