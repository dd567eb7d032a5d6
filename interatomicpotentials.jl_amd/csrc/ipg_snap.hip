// ipg_snap.hip -- SNAP tables, entry points and the deterministic gathers.
//
// Table semantics follow types/prebuilt_arrays.jl:48-321 (index blocks in
// the same j1 >= j2, j = |j1-j2|..min(J, j1+j2) step 2 order; the
// Clebsch-Gordan coefficients from the Racah closed form); the derived
// tables (Y coefficients, dB half-sum terms) are this engine's own.
#include "ipg_snap_driver.h"
#include "ipg_snap_f32.h"
#include "ipg_snap_kernels.h"

#include <algorithm>
#include <functional>

namespace ipg {

// ---------------------------------------------------------------------------
// host tables
// ---------------------------------------------------------------------------
static long double lfact(int n) {
  long double r = 1.0L;
  for (int k = 2; k <= n; k++) r *= (long double)k;
  return r;
}

// <j1/2 a/2, j2/2 b/2 | j/2 (a+b)/2> by the Racah formula (arguments doubled).
static double clebsch_gordan2(int j1, int j2, int j, int a, int b) {
  const int c = a + b;
  if (c < -j || c > j) return 0.0;
  const long double tri = lfact((j1 + j2 - j) / 2) * lfact((j1 - j2 + j) / 2) * lfact((-j1 + j2 + j) / 2) /
                          lfact((j1 + j2 + j) / 2 + 1);
  const long double nrm = (long double)(j + 1) * lfact((j1 + a) / 2) * lfact((j1 - a) / 2) * lfact((j2 + b) / 2) *
                          lfact((j2 - b) / 2) * lfact((j + c) / 2) * lfact((j - c) / 2);
  const int zmin = std::max(0, std::max(-(j - j2 + a) / 2, -(j - j1 - b) / 2));
  const int zmax = std::min((j1 + j2 - j) / 2, std::min((j1 - a) / 2, (j2 + b) / 2));
  long double s = 0.0L;
  for (int z = zmin; z <= zmax; z++) {
    const long double den = lfact(z) * lfact((j1 + j2 - j) / 2 - z) * lfact((j1 - a) / 2 - z) * lfact((j2 + b) / 2 - z) *
                            lfact((j - j2 + a) / 2 + z) * lfact((j - j1 - b) / 2 + z);
    s += ((z & 1) ? -1.0L : 1.0L) / den;
  }
  return (double)(std::sqrt(tri) * std::sqrt(nrm) * s);
}

void snap_build_host(SnapHost& h, const ipg_snap_params* p) {
  IPG_REQUIRE(p != nullptr, "params is NULL");
  IPG_REQUIRE(p->twojmax >= 0, "twojmax must be >= 0");
  // init_clebsch_gordan (prebuilt_arrays.jl:257-312) multiplies Int64
  // factorials; from 2J = 11 on the products overflow and the reference
  // stops with a DomainError (sqrt of a negative product) or an
  // OverflowError (factorial(21))
  if (p->twojmax > SNAP_JMAX_ALL)
    fail(IPG_ERR_UNSUPPORTED, "twojmax > 10: the reference's Int64 Clebsch-Gordan factorials overflow");
  IPG_REQUIRE(p->nspecies >= 1 && p->nspecies <= 8, "nspecies must be in [1, 8]");
  IPG_REQUIRE(p->radii && p->weight, "radii/weight is NULL");
  // snap.jl:24-27
  IPG_REQUIRE((p->chem_flag != 0) == (p->bnorm_flag != 0),
              "chem_flag and bnorm_flag should be the same, setting bnorm_flag == chem_flag");
  IPG_REQUIRE(std::isfinite(p->rcutfac) && p->rcutfac > 0, "rcutfac must be finite and > 0");
  IPG_REQUIRE(std::isfinite(p->rfac0) && std::isfinite(p->rmin0), "rfac0/rmin0 must be finite");
  for (int s = 0; s < p->nspecies; s++) {
    IPG_REQUIRE(std::isfinite(p->radii[s]) && p->radii[s] > 0, "radii must be finite and > 0");
    IPG_REQUIRE(std::isfinite(p->weight[s]), "weight must be finite");
    IPG_REQUIRE((p->radii[s] + p->radii[s]) * p->rcutfac > p->rmin0, "cutoff must exceed rmin0");
  }
  const int J = p->twojmax, jd = J + 1;
  h.J = J;
  h.nspecies = p->nspecies;
  h.bnorm = p->bnorm_flag != 0;
  h.bzero = p->bzero_flag != 0;
  h.switch_flag = p->switch_flag != 0;
  h.wselfall = p->wselfall_flag != 0;
  h.rcutfac = p->rcutfac;
  h.rmin0 = p->rmin0;
  h.rfac0 = p->rfac0;
  h.radii.assign(p->radii, p->radii + p->nspecies);
  h.weight.assign(p->weight, p->weight + p->nspecies);
  auto b3 = [&](int x, int y, int z) { return (x * jd + y) * jd + z; };
  h.idxu_block.assign(jd + 1, 0);
  for (int j = 0; j < jd; j++) h.idxu_block[j + 1] = h.idxu_block[j] + (j + 1) * (j + 1);
  h.idxu_max = h.idxu_block[jd];
  h.idxcg_block.assign(jd * jd * jd, -1);
  h.idxz_block.assign(jd * jd * jd, -1);
  h.idxb_block.assign(jd * jd * jd, -1);
  h.idxb.clear();
  h.zent.clear();
  h.yblk.clear();
  h.cg.clear();
  std::vector<std::array<int, 3>> blocks;
  for (int j1 = 0; j1 <= J; j1++)
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = j1 - j2; j <= std::min(J, j1 + j2); j += 2) {
        h.idxcg_block[b3(j1, j2, j)] = (int)h.cg.size();
        for (int m1 = 0; m1 <= j1; m1++)
          for (int m2 = 0; m2 <= j2; m2++) h.cg.push_back(clebsch_gordan2(j1, j2, j, 2 * m1 - j1, 2 * m2 - j2));
        if (j >= j1) {
          h.idxb_block[b3(j1, j2, j)] = (int)h.idxb.size();
          h.idxb.push_back({j1, j2, j});
        }
        blocks.push_back({j1, j2, j});
      }
  h.idxcg_max = (int)h.cg.size();
  h.idxb_max = (int)h.idxb.size();
  h.chem = p->chem_flag != 0;
  h.nel = h.chem ? p->nspecies : 1;  // prebuilt_arrays.jl:65-69
  h.generic = h.chem || J > SNAP_JMAX;
  h.ncoeff = h.idxb_max * h.nel * h.nel * h.nel;  // get_num_snap_coeffs (utl.jl:8-25)
  h.nblk = (int)blocks.size();
  for (int bk = 0; bk < h.nblk; bk++) {
    const int j1 = blocks[bk][0], j2 = blocks[bk][1], j = blocks[bk][2];
    h.idxz_block[b3(j1, j2, j)] = (int)h.zent.size();
    for (int mb = 0; 2 * mb <= j; mb++)
      for (int ma = 0; ma <= j; ma++) {
        ZEnt z;
        z.j1 = j1;
        z.j2 = j2;
        z.j = j;
        z.ma1min = std::max(0, (2 * ma - j - j2 + j1) / 2);
        z.ma2max = (2 * ma - j - (2 * z.ma1min - j1) + j2) / 2;
        z.na = std::min(j1, (2 * ma - j + j2 + j1) / 2) - z.ma1min + 1;
        z.mb1min = std::max(0, (2 * mb - j - j2 + j1) / 2);
        z.mb2max = (2 * mb - j - (2 * z.mb1min - j1) + j2) / 2;
        z.nb = std::min(j1, (2 * mb - j + j2 + j1) / 2) - z.mb1min + 1;
        z.jju = h.idxu_block[j] + (j + 1) * mb + ma;
        z.cgoff = h.idxcg_block[b3(j1, j2, j)];
        z.blk = bk;
        h.zent.push_back(z);
      }
  }
  h.idxz_max = (int)h.zent.size();
  // Y coefficients: the (triple, role) pairs whose dB/dr half-sum uses the
  // Z block (j1, j2, j) at level j.
  const double one = 1.0;
  for (int bk = 0; bk < h.nblk; bk++) {
    const int j1 = blocks[bk][0], j2 = blocks[bk][1], j = blocks[bk][2];
    YBlk y{};
    y.n = 0;
    if (j >= j1) {  // role j of (j1, j2, j)
      y.jjb[y.n] = h.idxb_block[b3(j1, j2, j)];
      y.fac[y.n++] = one;
    }
    if (j1 >= j && j >= j2) {  // role j1 of (j, j2, j1)
      y.jjb[y.n] = h.idxb_block[b3(j, j2, j1)];
      y.fac[y.n++] = h.bnorm ? one : (j1 + 1) / (j + 1.0);
    }
    if (j1 >= j2 && j2 >= j) {  // role j2 of (j2, j, j1)
      y.jjb[y.n] = h.idxb_block[b3(j2, j, j1)];
      y.fac[y.n++] = h.bnorm ? one : (j1 + 1) / (j + 1.0);
    }
    for (int k = 0; k < y.n; k++) IPG_REQUIRE(y.jjb[k] >= 0, "internal: bad Y block");
    h.yblk.push_back(y);
  }
  // dB/dr half-sum terms per triple (compute_b.jl:145-337)
  h.bterm.clear();
  for (int jjb = 0; jjb < h.idxb_max; jjb++) {
    const int j1 = h.idxb[jjb][0], j2 = h.idxb[jjb][1], j = h.idxb[jjb][2];
    BTerm t;
    t.j = j;
    t.l[0] = j;
    t.jjz0[0] = h.idxz_block[b3(j1, j2, j)];
    t.fac[0] = 1.0;
    t.l[1] = j1;
    t.jjz0[1] = h.idxz_block[b3(j, j2, j1)];
    t.fac[1] = h.bnorm ? 1.0 : (j + 1) / (j1 + 1.0);
    t.l[2] = j2;
    t.jjz0[2] = h.idxz_block[b3(j, j1, j2)];
    t.fac[2] = h.bnorm ? 1.0 : (j + 1) / (j2 + 1.0);
    IPG_REQUIRE(t.jjz0[0] >= 0 && t.jjz0[1] >= 0 && t.jjz0[2] >= 0, "internal: bad dB block");
    h.bterm.push_back(t);
  }
  // left-half positions (mb <= j/2) and the Z entries mapping to each
  h.half_jju.clear();
  h.yoff.assign(1, 0);
  h.yjjz.clear();
  std::vector<std::vector<int>> byjju(h.idxu_max);
  for (int k = 0; k < h.idxz_max; k++) byjju[h.zent[k].jju].push_back(k);
  for (int j = 0; j <= J; j++)
    for (int mb = 0; 2 * mb <= j; mb++)
      for (int ma = 0; ma <= j; ma++) {
        const int jju = h.idxu_block[j] + (j + 1) * mb + ma;
        h.half_jju.push_back(jju);
        for (int k : byjju[jju]) h.yjjz.push_back(k);
        h.yoff.push_back((int)h.yjjz.size());
      }
  h.nhalf = (int)h.half_jju.size();
  h.yjblk.clear();
  for (int k : h.yjjz) h.yjblk.push_back(h.zent[k].blk);
  // compact Z loop plans (see k_snap_zby): u1/u2 start, absolute CG offsets,
  // na, nb, row strides and j+1
  h.zplan.clear();
  for (const ZEnt& z : h.zent) {
    const uint32_t u1b = h.idxu_block[z.j1] + (z.j1 + 1) * z.mb1min + z.ma1min;
    const uint32_t u2b = h.idxu_block[z.j2] + (z.j2 + 1) * z.mb2max + z.ma2max;
    const uint32_t cga0 = z.cgoff + z.ma1min * (z.j2 + 1) + z.ma2max;
    const uint32_t cgb0 = z.cgoff + z.mb1min * (z.j2 + 1) + z.mb2max;
    IPG_REQUIRE(u1b < 65536 && u2b < 65536 && cga0 < 65536 && cgb0 < 65536 && z.na < 16 && z.nb < 16,
                "internal: Z plan field overflow");
    h.zplan.push_back(u1b | (u2b << 16));
    h.zplan.push_back(cga0 | (cgb0 << 16));
    h.zplan.push_back((uint32_t)z.na | ((uint32_t)z.nb << 4) | ((uint32_t)(z.j1 + 1) << 8) | ((uint32_t)z.j2 << 12) |
                      ((uint32_t)(z.j + 1) << 16));
  }
  // runtime-table path: sqrt(p/q) (prebuilt_arrays.jl:314-321) and the
  // (j, mb, ma) of every U entry
  h.rootpq.assign((size_t)jd * jd, 0.0);
  for (int pp = 1; pp <= J; pp++)
    for (int qq = 1; qq <= J; qq++) h.rootpq[pp * jd + qq] = std::sqrt((double)pp / (double)qq);
  h.jmm.clear();
  for (int j = 0; j <= J; j++)
    for (int mb = 0; mb <= j; mb++)
      for (int ma = 0; ma <= j; ma++) h.jmm.push_back(j | (mb << 8) | (ma << 16));
  h.wself = 1.0;
  h.bzero_j.assign(jd, 0.0);
  if (h.bzero)
    for (int j = 0; j <= J; j++) h.bzero_j[j] = h.bnorm ? 1.0 : (double)(j + 1);  // wself^3 = 1
}

static void snap_upload(ipg_snap* sn) {
  const SnapHost& h = sn->host;
  SnapDev& d = sn->dev;
  d.J = h.J;
  d.nspecies = h.nspecies;
  d.idxu_max = h.idxu_max;
  d.idxz_max = h.idxz_max;
  d.idxb_max = h.idxb_max;
  d.ncoeff = h.ncoeff;
  d.nblk = h.nblk;
  d.nhalf = h.nhalf;
  d.bnorm = h.bnorm;
  d.bzero = h.bzero;
  d.switch_flag = h.switch_flag;
  d.rcutfac = h.rcutfac;
  d.rmin0 = h.rmin0;
  d.rfac0 = h.rfac0;
  d.wself = h.wself;
  for (int s = 0; s < 8; s++) {
    d.radii[s] = s < h.nspecies ? h.radii[s] : 0.0;
    d.weight[s] = s < h.nspecies ? h.weight[s] : 0.0;
  }
  for (int j = 0; j <= SNAP_JMAX_ALL; j++) d.bzero_j[j] = j <= h.J ? h.bzero_j[j] : 0.0;
  for (int j = 0; j <= SNAP_JMAX_ALL + 1; j++) d.idxu_block[j] = j <= h.J + 1 ? h.idxu_block[j] : 0;
  d.hoff[0] = 0;
  for (int j = 0; j <= SNAP_JMAX_ALL; j++) d.hoff[j + 1] = d.hoff[j] + (j / 2 + 1) * (j + 1);
  d.generic = h.generic;
  d.chem = h.chem;
  d.nel = h.nel;
  d.wselfall = h.wselfall;
  std::vector<int> idxb3;
  for (const auto& b : h.idxb) idxb3.insert(idxb3.end(), b.begin(), b.end());
  // one allocation for every table
  std::vector<float> cgf(h.cg.begin(), h.cg.end());
  size_t off = 0;
  auto place = [&](size_t bytes) {
    size_t o = (off + 255) & ~(size_t)255;
    off = o + bytes;
    return o;
  };
  size_t o_z = place(sizeof(ZEnt) * h.zent.size());
  size_t o_y = place(sizeof(YBlk) * h.yblk.size());
  size_t o_b = place(sizeof(BTerm) * h.bterm.size());
  size_t o_cd = place(sizeof(double) * h.cg.size());
  size_t o_cf = place(sizeof(float) * cgf.size());
  size_t o_hj = place(sizeof(int) * h.half_jju.size());
  size_t o_yo = place(sizeof(int) * h.yoff.size());
  size_t o_yj = place(sizeof(int) * std::max<size_t>(1, h.yjjz.size()));
  size_t o_yb = place(sizeof(int) * std::max<size_t>(1, h.yjblk.size()));
  size_t o_zp = place(sizeof(uint32_t) * std::max<size_t>(1, h.zplan.size()));
  size_t o_rp = place(sizeof(double) * h.rootpq.size());
  size_t o_jm = place(sizeof(int) * h.jmm.size());
  size_t o_b3 = place(sizeof(int) * idxb3.size());
  size_t o_zb = place(sizeof(int) * h.idxz_block.size());
  if (hipMalloc(&sn->dmem, off) != hipSuccess) {
    (void)hipGetLastError();
    sn->dmem = nullptr;
    fail(IPG_ERR_NOMEM, "hipMalloc of SNAP tables failed");
  }
  char* base = (char*)sn->dmem;
  ipg_ctx* ctx = sn->ctx;
  h2d(ctx, base + o_z, h.zent.data(), sizeof(ZEnt) * h.zent.size());
  h2d(ctx, base + o_y, h.yblk.data(), sizeof(YBlk) * h.yblk.size());
  h2d(ctx, base + o_b, h.bterm.data(), sizeof(BTerm) * h.bterm.size());
  h2d(ctx, base + o_cd, h.cg.data(), sizeof(double) * h.cg.size());
  h2d(ctx, base + o_cf, cgf.data(), sizeof(float) * cgf.size());
  h2d(ctx, base + o_hj, h.half_jju.data(), sizeof(int) * h.half_jju.size());
  h2d(ctx, base + o_yo, h.yoff.data(), sizeof(int) * h.yoff.size());
  if (!h.yjjz.empty()) h2d(ctx, base + o_yj, h.yjjz.data(), sizeof(int) * h.yjjz.size());
  if (!h.yjblk.empty()) h2d(ctx, base + o_yb, h.yjblk.data(), sizeof(int) * h.yjblk.size());
  h2d(ctx, base + o_zp, h.zplan.data(), sizeof(uint32_t) * h.zplan.size());
  h2d(ctx, base + o_rp, h.rootpq.data(), sizeof(double) * h.rootpq.size());
  h2d(ctx, base + o_jm, h.jmm.data(), sizeof(int) * h.jmm.size());
  h2d(ctx, base + o_b3, idxb3.data(), sizeof(int) * idxb3.size());
  h2d(ctx, base + o_zb, h.idxz_block.data(), sizeof(int) * h.idxz_block.size());
  sync(ctx);
  d.zent = (const ZEnt*)(base + o_z);
  d.yblk = (const YBlk*)(base + o_y);
  d.bterm = (const BTerm*)(base + o_b);
  d.cg_d = (const double*)(base + o_cd);
  d.cg_f = (const float*)(base + o_cf);
  d.half_jju = (const int*)(base + o_hj);
  d.yoff = (const int*)(base + o_yo);
  d.yjjz = (const int*)(base + o_yj);
  d.yjblk = (const int*)(base + o_yb);
  d.zplan = (const uint32_t*)(base + o_zp);
  d.idxcg_max = h.idxcg_max;
  d.rootpq = (const double*)(base + o_rp);
  d.jmm = (const int*)(base + o_jm);
  d.idxb3 = (const int*)(base + o_b3);
  d.idxz_block = (const int*)(base + o_zb);
}

// ---------------------------------------------------------------------------
// gathers (deterministic: fixed pair order per atom, no atomics)
// ---------------------------------------------------------------------------
// F_i = sum_(i->j) dE_i/dr_ij - sum_(j->i) dE_j/dr_ji (snap.jl:264-265),
// 16 lanes per atom.  Lane l takes the l-th, (l+16)-th, ... entry of the
// atom's own row, then of its reverse row, and the lanes are summed in a
// fixed tree.  SYM: the reverse row is rev[] of the own row; otherwise
// roff/rpair.  For a symmetric list both give the same entries in the same
// order, so the two variants agree bit for bit.
constexpr int FG_LANES = 16;
template <bool SYM>
__global__ __launch_bounds__(256) void k_force_gather16(int64_t n, const int64_t* __restrict__ off,
                                                        const int32_t* __restrict__ rev,
                                                        const int64_t* __restrict__ roff,
                                                        const int32_t* __restrict__ rpair,
                                                        const double* __restrict__ dedr, double* __restrict__ F) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / FG_LANES;
  const int l = (int)(t % FG_LANES);
  if (i >= n) return;  // whole 16-lane groups leave together
  double fx = 0, fy = 0, fz = 0;
  const int64_t a = off[i], b = off[i + 1];
  for (int64_t q = a + l; q < b; q += FG_LANES) {
    fx += dedr[3 * q];
    fy += dedr[3 * q + 1];
    fz += dedr[3 * q + 2];
  }
  if constexpr (SYM) {
    for (int64_t q = a + l; q < b; q += FG_LANES) {
      const int64_t r = rev[q];
      fx -= dedr[3 * r];
      fy -= dedr[3 * r + 1];
      fz -= dedr[3 * r + 2];
    }
  } else {
    for (int64_t k = roff[i] + l; k < roff[i + 1]; k += FG_LANES) {
      const int64_t r = rpair[k];
      fx -= dedr[3 * r];
      fy -= dedr[3 * r + 1];
      fz -= dedr[3 * r + 2];
    }
  }
  fx = group_sum<FG_LANES>(fx);
  fy = group_sum<FG_LANES>(fy);
  fz = group_sum<FG_LANES>(fz);
  if (l == 0) {
    F[3 * i] = fx;
    F[3 * i + 1] = fy;
    F[3 * i + 2] = fz;
  }
}

// The last two kernels of the adjoint step over a symmetric CUTOFF list.
// k_step_forces: the forces, the own row of dE_i/dr_ij minus the reverse
// pairs rev[q], 16 lanes per atom in a fixed tree (as k_force_gather16<true>,
// bit for bit), and each block's sum of its atoms' E_i in atom order.
// k_step_close (one block): the energy beta0 + the block sums in block
// order (so the sum does not depend on timing) and the step's flag word
// copied to host-visible memory, so the host needs no device-to-host copy
// after it.  The kernel boundary orders the two; no device-wide fence.
constexpr int EPI_ATOMS = 256 / FG_LANES;
__global__ __launch_bounds__(256) void k_step_forces(int64_t n, const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ rev,
                                                     const double* __restrict__ dedr, double* __restrict__ F,
                                                     const double* __restrict__ eatom, double* __restrict__ part) {
  __shared__ double es[EPI_ATOMS];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / FG_LANES;
  const int l = (int)(t % FG_LANES);
  double fx = 0, fy = 0, fz = 0;
  if (i < n) {
    const int64_t a = off[i], b = off[i + 1];
    for (int64_t q = a + l; q < b; q += FG_LANES) {
      fx += dedr[3 * q];
      fy += dedr[3 * q + 1];
      fz += dedr[3 * q + 2];
    }
    for (int64_t q = a + l; q < b; q += FG_LANES) {
      const int64_t r = rev[q];
      fx -= dedr[3 * r];
      fy -= dedr[3 * r + 1];
      fz -= dedr[3 * r + 2];
    }
  }
  fx = group_sum<FG_LANES>(fx);
  fy = group_sum<FG_LANES>(fy);
  fz = group_sum<FG_LANES>(fz);
  if (l == 0) {
    if (i < n) {
      F[3 * i] = fx;
      F[3 * i + 1] = fy;
      F[3 * i + 2] = fz;
    }
    es[threadIdx.x / FG_LANES] = i < n ? eatom[i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < EPI_ATOMS; k++) s += es[k];
    part[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(256) void k_step_close(int64_t nblocks, const double* __restrict__ part, double beta0,
                                                    double* __restrict__ energy, const int32_t* __restrict__ flag,
                                                    int32_t* host_flag) {
  __shared__ double ws[4];
  double s = 0.0;
  for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x) s += part[b];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double tot = ((ws[0] + ws[1]) + ws[2]) + ws[3];
    if (energy) *energy = beta0 + tot;
    *host_flag = *flag;
  }
}

static void force_gather(ipg_ctx* ctx, const SnapRun& R, const double* dedr, double* F) {
  if (R.n == 0) return;
  const unsigned g = grid_for(R.n * FG_LANES, 256);
  launch(ctx, "snap_force_gather", [&] {
    if (R.rev)
      k_force_gather16<true><<<g, 256, 0, ctx->stream>>>(R.n, R.off, R.rev, nullptr, nullptr, dedr, F);
    else
      k_force_gather16<false><<<g, 256, 0, ctx->stream>>>(R.n, R.off, nullptr, R.roff, R.rpair, dedr, F);
  });
}

// dB[a] (M x 3, blocks by the species of the pair's center atom,
// snap.jl:264-265): block s(a) gets the atom's own pairs, block s gets minus
// the reverse pairs whose center has species s.  One pass over the own row
// and one over the reverse row, each in ascending pair order.
template <typename T>
__global__ void k_db_gather(SnapDev P, SnapRun R, const T* __restrict__ dblist, double* __restrict__ dB) {
  const int64_t a = blockIdx.x;
  const int M = P.ncoeff * P.nspecies;
  const int sa = snap_species(R.species, a, P.nspecies);
  const int64_t q0 = R.off[a], q1 = R.off[a + 1], r0 = R.roff[a], r1 = R.roff[a + 1];
  for (int t = threadIdx.x; t < 3 * P.ncoeff; t += blockDim.x) {
    double acc[8];
#pragma unroll
    for (int s = 0; s < 8; s++) acc[s] = 0.0;
    double own = 0.0;
    for (int64_t q = q0; q < q1; q++) own += (double)dblist[q * 3 * P.ncoeff + t];
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (s == sa) acc[s] = own;
    for (int64_t r = r0; r < r1; r++) {
      const int64_t q = R.rpair[r];
      const int sc = snap_species(R.species, R.center[q], P.nspecies);
      const double v = (double)dblist[q * 3 * P.ncoeff + t];
#pragma unroll
      for (int s = 0; s < 8; s++)
        if (s == sc) acc[s] -= v;
    }
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (s < P.nspecies) dB[(a * M + s * P.ncoeff) * 3 + t] = acc[s];
  }
}

// The same gather over a symmetric CUTOFF list: the pairs whose neighbor is
// atom a are the reverse pairs rev[q] of a's own pairs q, met in the same
// (ascending) order as in the general reverse rows, so the result equals
// k_db_gather's bit for bit, without the reverse-row indirection.  With Wa,
// the own-row pass also forms the per-atom virial rows of the own species
// block (snap.jl:267-281, raw positions): Wa[a][k][v] for v = (x x, y y,
// z z, y z, x z, x y), written by the lane whose first index matches.
template <typename T>
__global__ void k_db_gather_sym(SnapDev P, SnapRun R, const T* __restrict__ dblist, double* __restrict__ dB,
                                const double* __restrict__ pos, double* __restrict__ Wa) {
  const int64_t a = blockIdx.x;
  const int M = P.ncoeff * P.nspecies, nc3 = 3 * P.ncoeff;
  const int sa = snap_species(R.species, a, P.nspecies);
  const int64_t q0 = R.off[a], q1 = R.off[a + 1];
  const double xa = Wa ? pos[3 * a] : 0.0, ya = Wa ? pos[3 * a + 1] : 0.0, za = Wa ? pos[3 * a + 2] : 0.0;
  for (int t = threadIdx.x; t < nc3; t += blockDim.x) {
    double own = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
    if (Wa) {
#pragma unroll 8
      for (int64_t q = q0; q < q1; q++) {
        const int64_t j = R.nbr[q];
        const double d = (double)dblist[q * nc3 + t];
        own += d;
        wx += d * xa - d * pos[3 * j];
        wy += d * ya - d * pos[3 * j + 1];
        wz += d * za - d * pos[3 * j + 2];
      }
      const int k = t / 3, d1 = t % 3;
      double* w = Wa + (a * P.ncoeff + k) * 6;
      if (d1 == 0) {
        w[0] = wx;
        w[5] = wy;
        w[4] = wz;
      } else if (d1 == 1) {
        w[1] = wy;
        w[3] = wz;
      } else {
        w[2] = wz;
      }
    } else {
#pragma unroll 8
      for (int64_t q = q0; q < q1; q++) own += (double)dblist[q * nc3 + t];
    }
    if (!dB) continue;
    double acc[8];
#pragma unroll
    for (int s = 0; s < 8; s++) acc[s] = s == sa ? own : 0.0;
#pragma unroll 8
    for (int64_t q = q0; q < q1; q++) {
      const int sc = snap_species(R.species, R.nbr[q], P.nspecies);
      const double v = (double)dblist[(int64_t)R.rev[q] * nc3 + t];
#pragma unroll
      for (int s = 0; s < 8; s++)
        if (s == sc) acc[s] -= v;
    }
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (s < P.nspecies) dB[(a * M + s * P.ncoeff) * 3 + t] = acc[s];
  }
}

__global__ void k_force_from_db(int64_t n, int M, const double* __restrict__ dB, const double* __restrict__ beta,
                                double* __restrict__ F) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int64_t a = t / 3;
  const int k = (int)(t % 3);
  double f = 0.0;
  for (int m = 0; m < M; m++) f += beta[m] * dB[(a * M + m) * 3 + k];  // linear_potential.jl:28-35
  F[t] = f;
}

// Per-center-atom virial rows (snap.jl:267-281): sum over its pairs of
// [dBx x, dBy y, dBz z, dBy z, dBx z, dBx y] with x = x_i - x_j (raw).
__global__ void k_w_atom(SnapDev P, SnapRun R, const double* __restrict__ pos, const double* __restrict__ dblist,
                         double* __restrict__ Wa) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int M = P.ncoeff * P.nspecies;
  if (t >= R.n * (int64_t)M) return;
  const int64_t a = t / M;
  const int m = (int)(t % M);
  const int sa = snap_species(R.species, a, P.nspecies);
  double w[6] = {0, 0, 0, 0, 0, 0};
  if (m / P.ncoeff == sa) {
    const int c = m % P.ncoeff;
    for (int64_t q = R.off[a]; q < R.off[a + 1]; q++) {
      const int32_t jn = R.nbr[q];
      const double* db = dblist + (q * P.ncoeff + c) * 3;
      const double xi = pos[3 * a], yi = pos[3 * a + 1], zi = pos[3 * a + 2];
      const double xj = pos[3 * jn], yj = pos[3 * jn + 1], zj = pos[3 * jn + 2];
      w[0] += db[0] * xi - db[0] * xj;
      w[1] += db[1] * yi - db[1] * yj;
      w[2] += db[2] * zi - db[2] * zj;
      w[3] += db[1] * zi - db[1] * zj;
      w[4] += db[0] * zi - db[0] * zj;
      w[5] += db[0] * yi - db[0] * yj;
    }
  }
  for (int v = 0; v < 6; v++) Wa[t * 6 + v] = w[v];
}

// Per-center-atom virial of the adjoint route: beta_s(i) . W over the
// atom's own pairs is sum_q dE_i/dr_ij (x) (x_i - x_j) with raw positions
// (snap.jl:267-281 contracted with beta, linear_potential.jl:37-42), rows
// [xx, yy, zz, zy, zx, yx] with the first index the dE/dr component.
__global__ void k_virial_atoms(int64_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ nbr,
                               const double* __restrict__ pos, const double* __restrict__ dedr,
                               double* __restrict__ va) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const double xi = pos[3 * a], yi = pos[3 * a + 1], zi = pos[3 * a + 2];
  double w[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t q = off[a]; q < off[a + 1]; q++) {
    const int32_t j = nbr[q];
    const double dx = xi - pos[3 * j], dy = yi - pos[3 * j + 1], dz = zi - pos[3 * j + 2];
    const double gx = dedr[3 * q], gy = dedr[3 * q + 1], gz = dedr[3 * q + 2];
    w[0] += gx * dx;
    w[1] += gy * dy;
    w[2] += gz * dz;
    w[3] += gy * dz;
    w[4] += gx * dz;
    w[5] += gx * dy;
  }
  for (int v = 0; v < 6; v++) va[a * 6 + v] = w[v];
}

// virial[v] = sum_m W[m][v] beta[m] (linear_potential.jl:37-42, W * beta),
// one lane per row, terms in m order.
__global__ void k_w_beta(int M, const double* __restrict__ W, const double* __restrict__ beta,
                         double* __restrict__ virial) {
  const int v = threadIdx.x;
  if (v >= 6) return;
  double s = 0.0;
  for (int m = 0; m < M; m++) s += W[m * 6 + v] * beta[m];
  virial[v] = s;
}

__global__ void k_finish_energy(const double* __restrict__ sum, double beta0, double* __restrict__ out) {
  if (threadIdx.x == 0) out[0] = beta0 + sum[0];
}

// ---------------------------------------------------------------------------
// drivers
// ---------------------------------------------------------------------------
struct SnapCall {
  SnapListDev L;
  SnapRun R;
  int32_t* flag = nullptr;  // device bits: 1, 2, 8 capacity overflow; 4 missing reverse pair
};

SnapCut snap_make_cut(const SnapHost& h, int mode) {
  SnapCut c{};
  c.ns = h.nspecies;
  double cutmax = 0.0;
  for (int i = 0; i < h.nspecies; i++) {  // utl.jl:42-56
    const double cut = 2.0 * h.radii[i] * h.rcutfac;
    cutmax = std::max(cutmax, cut);
    c.cutsq[i * h.nspecies + i] = cut * cut;
    for (int j = i + 1; j < h.nspecies; j++) {
      const double cj = (h.radii[i] + h.radii[j]) * h.rcutfac;
      c.cutsq[i * h.nspecies + j] = cj * cj;
      c.cutsq[j * h.nspecies + i] = cj * cj;
    }
  }
  c.reference = mode == IPG_SNAP_LIST_REFERENCE ? 1 : 0;
  c.query = c.reference ? std::sqrt(cutmax) : cutmax;  // utl.jl:81
  return c;
}

// need: SnapListNeed bits.  The call's flag word is cleared first.
static SnapCall snap_prepare(ipg_snap* sn, const DevSystem& s, int mode, const int32_t* d_appear, int need,
                             int64_t capA = 0, int64_t capB = 0, int ellK = 0) {
  IPG_REQUIRE(mode == IPG_SNAP_LIST_REFERENCE || mode == IPG_SNAP_LIST_CUTOFF, "unknown SNAP list mode");
  SnapCall c;
  c.flag = scratch_t<int32_t>(sn->ctx, "snap_flag", 1);
  SnapCut cut = snap_make_cut(sn->host, mode);
  // CUTOFF mode: the list build clears the flag in its first fill launch
  if (cut.reference) zero(sn->ctx, c.flag, sizeof(int32_t));
  const int32_t* ap = d_appear ? d_appear : s.species;
  if (cut.reference && !ap) {
    // all atoms are species 0: element index 0 everywhere
    int32_t* z = scratch_t<int32_t>(sn->ctx, "snap_appear0", s.n > 0 ? s.n : 1);
    zero(sn->ctx, z, sizeof(int32_t) * (s.n > 0 ? s.n : 1));
    ap = z;
  }
  c.L = build_snap_list(sn->ctx, s, ap, cut, need, "snap_list", capA, capB, c.flag, ellK, !cut.reference);
  c.R.n = s.n;
  c.R.npairs = c.L.list.npairs;
  c.R.d_np = c.L.list.exact ? nullptr : c.L.list.off + s.n;
  c.R.off = c.L.list.off;
  c.R.nbr = c.L.list.nbr;
  c.R.center = c.L.center;
  c.R.rvec = c.L.rvec;
  c.R.species = s.species;
  c.R.roff = c.L.roff;
  c.R.rpair = c.L.rpair;
  c.R.rev = c.L.rev;
  c.R.flag = c.flag;
  return c;
}

// The runtime-table configurations (chem_flag, 2J > 8) run the reference's
// own route, forces = beta . dB (linear_potential.jl:28-35): the adjoint
// and FP32 kernels are compiled for 2J <= 8 without chem_flag, and with
// chem_flag the reference's dB is not the gradient of its B (compute_zi
// keeps one Z block), so an adjoint would not reproduce its forces.
static void require_path(const ipg_snap* sn, int path) {
  IPG_REQUIRE(path == IPG_SNAP_PLAIN_FP64 || path == IPG_SNAP_ADJOINT_FP64 || path == IPG_SNAP_ADJOINT_FP32,
              "unknown SNAP path");
  if (sn->host.generic && path != IPG_SNAP_PLAIN_FP64)
    fail(IPG_ERR_UNSUPPORTED, "chem_flag = true or twojmax > 8: only IPG_SNAP_PLAIN_FP64 (beta . dB) is available");
}

static void check_species_host(const ipg_snap* sn, const ipg_system* sys, const int32_t* appear) {
  for (int64_t i = 0; i < sys->natoms; i++) {
    if (sys->species)
      IPG_REQUIRE(sys->species[i] >= 0 && sys->species[i] < sn->host.nspecies,
                  "species index out of range of snap.species");
    if (appear)
      IPG_REQUIRE(appear[i] >= 0 && appear[i] < sn->host.nspecies,
                  "element index (first appearance order) out of range of snap.species");
  }
}

static const int32_t* upload_appear(ipg_ctx* ctx, const int32_t* appear, int64_t n) {
  if (!appear) return nullptr;
  int32_t* d = scratch_t<int32_t>(ctx, "snap_appear", n > 0 ? n : 1);
  h2d(ctx, d, appear, sizeof(int32_t) * n);
  return d;
}

static const double* expand_beta(ipg_snap* sn, const double* beta, int64_t nbeta) {
  const int nc = sn->host.ncoeff, ns = sn->host.nspecies;
  IPG_REQUIRE(beta != nullptr, "beta is NULL");
  IPG_REQUIRE(nbeta == nc || nbeta == (int64_t)nc * ns, "beta must have ncoeff or ncoeff*nspecies entries");
  std::vector<double> b((size_t)nc * ns);
  for (int s = 0; s < ns; s++)
    for (int c = 0; c < nc; c++) b[(size_t)s * nc + c] = nbeta == nc ? beta[c] : beta[(size_t)s * nc + c];
  for (double v : b) IPG_REQUIRE(std::isfinite(v), "beta must be finite");
  double* d = scratch_t<double>(sn->ctx, "snap_beta", b.size());
  h2d(sn->ctx, d, b.data(), sizeof(double) * b.size());
  return d;
}

// U -> Z/B (and optionally Z to global) in FP64.  Returns the Z buffer.
static double* run_descriptors_fp64(ipg_snap* sn, const SnapCall& c, double* dB_out, double* B_out, double* eatom,
                                    const double* beta, bool want_z) {
  ipg_ctx* ctx = sn->ctx;
  const SnapDev& P = sn->dev;
  const int64_t n = c.R.n;
  if (P.generic) {  // runtime tables: U per element, Z0, B (always formed: E_i needs it)
    const size_t n1 = (size_t)std::max<int64_t>(n, 1);
    double* U = scratch_t<double>(ctx, "snap_U64", 2 * (size_t)P.nel * P.idxu_max * n1);
    double* Z = scratch_t<double>(ctx, "snap_Z64", 2 * (size_t)P.idxz_max * n1);
    double* B = B_out ? B_out : scratch_t<double>(ctx, "snap_Bgen", (size_t)P.ncoeff * n1);
    snap_gen_descriptors(ctx, P, c.R, U, Z, B, beta, eatom);
    (void)dB_out;
    return Z;
  }
  double* U = scratch_t<double>(ctx, "snap_U64", 2 * (size_t)P.idxu_max * std::max<int64_t>(n, 1));
  double* Z = want_z ? scratch_t<double>(ctx, "snap_Z64", 2 * (size_t)P.idxz_max * std::max<int64_t>(n, 1)) : nullptr;
  snap_launch_ui<double>(ctx, P, c.R, U);
  snap_launch_zby<double>(ctx, P, c.R, U, beta, nullptr, B_out, eatom, nullptr, Z, nullptr);
  (void)dB_out;
  return Z;
}

static double* run_dblist(ipg_snap* sn, const SnapCall& c, const double* Z) {
  const SnapDev& P = sn->dev;
  double* dbl = scratch_t<double>(sn->ctx, "snap_dblist", 3 * (size_t)P.ncoeff * std::max<int64_t>(c.R.npairs, 1));
  if (P.generic) snap_gen_dbdr(sn->ctx, P, c.R, Z, dbl);
  else snap_launch_dbdr<double>(sn->ctx, P, c.R, Z, dbl);
  return dbl;
}

template <typename T>
void snap_gather_db(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const T* dbl, double* dB, const double* pos,
                    double* Wa) {
  if (R.n == 0) return;
  IPG_REQUIRE(R.rev || !Wa, "internal: the fused virial rows need the reverse-pair index");
  // one lane per (coefficient, component), in whole waves
  const int threads = std::min(1024, (3 * P.ncoeff + 63) / 64 * 64);
  if (R.rev)
    launch(ctx, "snap_db_gather", [&] {
      k_db_gather_sym<T><<<(unsigned)R.n, threads, 0, ctx->stream>>>(P, R, dbl, dB, pos, Wa);
    });
  else
    launch(ctx, "snap_db_gather", [&] {
      k_db_gather<T><<<(unsigned)R.n, threads, 0, ctx->stream>>>(P, R, dbl, dB);
    });
}
template void snap_gather_db<float>(ipg_ctx*, const SnapDev&, const SnapRun&, const float*, double*, const double*,
                                    double*);
template void snap_gather_db<double>(ipg_ctx*, const SnapDev&, const SnapRun&, const double*, double*,
                                     const double*, double*);

static void gather_db(ipg_snap* sn, const SnapCall& c, const double* dbl, double* dB) {
  snap_gather_db<double>(sn->ctx, sn->dev, c.R, dbl, dB);
}

// Energy and forces on the device; d_energy/d_forces are device pointers.
// d_virial (optional, device, 6): beta . W, the virial stress of the
// linear potential, from the same per-pair derivatives; pos holds the raw
// positions it needs.  Returns true when the step's last kernel copied the
// flag word to sn->host_flag (the adjoint paths over a CUTOFF list), so the
// caller reads it there after synchronizing instead of copying it back.
static bool energy_forces_dev(ipg_snap* sn, const SnapCall& c, const double* d_beta, double beta0, int path,
                              double* d_energy, double* d_forces, double* d_virial = nullptr,
                              const double* pos = nullptr) {
  ipg_ctx* ctx = sn->ctx;
  const SnapDev& P = sn->dev;
  const int64_t n = c.R.n;
  const int M = P.ncoeff * P.nspecies;
  double* eatom = scratch_t<double>(ctx, "snap_eatom", std::max<int64_t>(n, 1));
  double* esum = scratch_t<double>(ctx, "snap_esum", 2);
  double* F = d_forces ? d_forces : scratch_t<double>(ctx, "snap_Fscratch", 3 * std::max<int64_t>(n, 1));
  if (n == 0) {
    zero(ctx, esum, sizeof(double) * 2);
    if (d_virial) zero(ctx, d_virial, sizeof(double) * 6);
    if (d_energy)
      launch(ctx, "snap_finish_energy", [&] { k_finish_energy<<<1, 64, 0, ctx->stream>>>(esum, beta0, d_energy); });
    return false;
  } else if (path == IPG_SNAP_PLAIN_FP64) {
    double* Z = run_descriptors_fp64(sn, c, nullptr, nullptr, eatom, d_beta, true);
    double* dbl = run_dblist(sn, c, Z);
    double* dB = scratch_t<double>(ctx, "snap_dB", 3 * (size_t)M * n);
    gather_db(sn, c, dbl, dB);
    launch(ctx, "snap_force_from_db",
           [&] { k_force_from_db<<<grid_for(3 * n, 256), 256, 0, ctx->stream>>>(n, M, dB, d_beta, F); });
    if (d_energy) reduce_rows(ctx, eatom, n, 1, d_energy, false, beta0);  // E = beta0 + sum_i E_i
    if (d_virial) {
      // the reference's route: W of compute_all_descriptors, then W * beta
      double* Wa = scratch_t<double>(ctx, "snap_Wa", 6 * (size_t)M * n);
      double* Wd = scratch_t<double>(ctx, "snap_W", 6 * (size_t)M);
      launch(ctx, "snap_w_atom", [&] {
        k_w_atom<<<grid_for(n * (int64_t)M, 256), 256, 0, ctx->stream>>>(P, c.R, pos, dbl, Wa);
      });
      reduce_rows(ctx, Wa, n, 6 * M, Wd, false);
      launch(ctx, "snap_w_beta", [&] { k_w_beta<<<1, 64, 0, ctx->stream>>>(M, Wd, d_beta, d_virial); });
    }
    return false;
  } else {
    double* dedr = scratch_t<double>(ctx, "snap_dedr", 3 * std::max<int64_t>(c.R.npairs, 1));
    if (path == IPG_SNAP_ADJOINT_FP64) {
      double* U = scratch_t<double>(ctx, "snap_U64", 2 * (size_t)P.idxu_max * n);
      double* Y = scratch_t<double>(ctx, "snap_Y64", 2 * (size_t)P.idxu_max * n);
      double* yc = scratch_t<double>(ctx, "snap_yc64", (size_t)P.nspecies * P.nblk);
      snap_launch_ycoef<double>(ctx, P, d_beta, yc);
      snap_launch_ui<double>(ctx, P, c.R, U);
      snap_launch_zby<double>(ctx, P, c.R, U, d_beta, yc, nullptr, eatom, Y, nullptr, nullptr);
      snap_launch_dedr_rev<double>(ctx, P, c.R, Y, dedr, nullptr);
    } else if (path == IPG_SNAP_ADJOINT_FP32) {
      const size_t nblk32 = (size_t)((n + 31) / 32) * 32;  // atom-blocked layout, padded to whole batches
      float* U = scratch_t<float>(ctx, "snap_U32", 2 * (size_t)P.idxu_max * nblk32);
      float* Y = scratch_t<float>(ctx, "snap_Y32", 2 * (size_t)P.idxu_max * nblk32);
      snap_launch_ui_rows32(ctx, P, c.R, U);
      snap_launch_zy32(ctx, P, c.R, U, d_beta, Y, eatom);
      snap_launch_dedr_rev32(ctx, P, c.R, Y, dedr);
    } else {
      fail(IPG_ERR_INVALID, "unknown SNAP path");
    }
    if (c.R.rev && !d_virial && sn->host_flag) {
      // forces, energy and the flag word in two kernels
      const unsigned gE = grid_for(n * FG_LANES, 256);
      double* part = scratch_t<double>(ctx, "snap_epart", gE);
      launch(ctx, "snap_force_gather", [&] {
        k_step_forces<<<gE, 256, 0, ctx->stream>>>(n, c.R.off, c.R.rev, dedr, F, eatom, part);
      });
      launch(ctx, "snap_step_close", [&] {
        k_step_close<<<1, 256, 0, ctx->stream>>>(gE, part, beta0, d_energy, c.flag, sn->host_flag);
      });
      return true;
    }
    force_gather(ctx, c.R, dedr, F);
    if (d_energy) reduce_rows(ctx, eatom, n, 1, d_energy, false, beta0);  // E = beta0 + sum_i E_i
    if (d_virial) {
      double* va = scratch_t<double>(ctx, "snap_vatom", 6 * (size_t)n);
      launch(ctx, "snap_virial_atoms", [&] {
        k_virial_atoms<<<grid_for(n, 256), 256, 0, ctx->stream>>>(n, c.R.off, c.R.nbr, pos, dedr, va);
      });
      reduce_rows(ctx, va, n, 6, d_virial, false);
    }
    return false;
  }
  return false;
}

// A CUTOFF-mode list with a pair whose reverse is missing: rebuild the
// general reverse rows and redo the adjoint force gather from the stored
// dE_i/dr_ij of the same call.
static void fix_asymmetric(ipg_snap* sn, SnapCall& c, int path, const double* d_beta, double* d_forces) {
  ipg_ctx* ctx = sn->ctx;
  const int64_t n = c.R.n, np = c.L.list.npairs, npa = np > 0 ? np : 1;
  c.L.roff = scratch_t<int64_t>(ctx, "snap_list_roff", n + 1);
  c.L.rpair = scratch_t<int32_t>(ctx, "snap_list_rpair", npa);
  build_center_and_reverse(ctx, n, np, c.L.list.off, c.L.list.nbr, nullptr, c.L.roff, c.L.rpair, "snap_list");
  c.R.roff = c.L.roff;
  c.R.rpair = c.L.rpair;
  c.R.rev = nullptr;
  double* F = d_forces ? d_forces : scratch_t<double>(ctx, "snap_Fscratch", 3 * std::max<int64_t>(n, 1));
  if (path == IPG_SNAP_PLAIN_FP64) {
    // the dB gather again from the same per-pair dB, then F = dB . beta
    const SnapDev& P = sn->dev;
    const int M = P.ncoeff * P.nspecies;
    const double* dbl = scratch_t<double>(ctx, "snap_dblist", 3 * (size_t)P.ncoeff * npa);
    double* dB = scratch_t<double>(ctx, "snap_dB", 3 * (size_t)M * n);
    snap_gather_db<double>(ctx, P, c.R, dbl, dB);
    launch(ctx, "snap_force_from_db",
           [&] { k_force_from_db<<<grid_for(3 * n, 256), 256, 0, ctx->stream>>>(n, M, dB, d_beta, F); });
    return;
  }
  double* dedr = scratch_t<double>(ctx, "snap_dedr", 3 * std::max<int64_t>(np, 1));
  force_gather(ctx, c.R, dedr, F);
}

// ---------------------------------------------------------------------------
// captured step: energy + forces of one system, replayed from a HIP graph
// ---------------------------------------------------------------------------
void graph_release(ipg_ctx* ctx, SnapGraph& g) {
  if (g.exec) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
  }
  for (auto& p : g.events) {
    ctx->free_events.push_back(p.a);
    ctx->free_events.push_back(p.b);
  }
  g.events.clear();
}

static bool graph_matches(const SnapGraph& g, const ipg_ctx* ctx, const DevSystem& s, int mode,
                          const int32_t* d_appear, const double* d_beta, double beta0, int path, double* d_energy,
                          double* d_forces) {
  if (!g.exec || g.n != s.n || g.pos != s.pos || g.species != s.species || g.appear != d_appear) return false;
  for (int k = 0; k < 3; k++)
    if (g.box.per[k] != s.box.per[k] || g.box.L[k] != s.box.L[k]) return false;
  return g.mode == mode && g.path == path && g.beta == d_beta && g.beta0 == beta0 && g.energy == d_energy &&
         g.forces == d_forces && g.prof == ctx->prof && g.prof_gen == ctx->prof_gen && g.gen == ctx->scratch_gen;
}

// Allocation pass (launches skipped) at the step's sizes, then capture of
// the step into g.exec.  On failure the graph stays off for this profiling
// state, with the reason kept.  Sets g.events and g.gen on success.
bool graph_capture(ipg_ctx* ctx, SnapGraph& g, const std::function<void()>& step) {
  ctx->dry = true;
  try {
    step();
  } catch (...) {
    ctx->dry = false;
    throw;
  }
  ctx->dry = false;
  const uint64_t gen = ctx->scratch_gen;
  if (ctx->prof)  // events for the graph's nodes are created before the capture starts
    while (ctx->free_events.size() < 256) {
      hipEvent_t e = nullptr;
      IPG_HIP(hipEventCreate(&e));
      ctx->free_events.push_back(e);
    }
  std::vector<ipg_ctx::Pending> ev;
  std::string why;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    ctx->capture_events = &ev;
    try {
      step();
    } catch (const std::exception& x) {
      why = x.what();
    }
    ctx->capture_events = nullptr;
    hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
    if (why.empty() && e2 != hipSuccess) why = std::string("hipStreamEndCapture: ") + hipGetErrorString(e2);
    if (why.empty() && ctx->scratch_gen != gen) why = "scratch reallocated during capture";
    if (why.empty()) {
      e2 = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      if (e2 != hipSuccess) {
        g.exec = nullptr;
        why = std::string("hipGraphInstantiate: ") + hipGetErrorString(e2);
      }
    }
    if (graph) (void)hipGraphDestroy(graph);
  } else {
    why = std::string("hipStreamBeginCapture: ") + hipGetErrorString(e);
  }
  if (!why.empty()) {
    (void)hipGetLastError();
    for (auto& p : ev) {
      ctx->free_events.push_back(p.a);
      ctx->free_events.push_back(p.b);
    }
    g.disabled[ctx->prof ? 1 : 0] = true;
    g.reason = why;
    return false;
  }
  g.events = std::move(ev);
  g.gen = gen;
  return true;
}

// Energy and forces into device buffers.  For a periodic system with graphs
// enabled, the first call (and any call whose inputs moved to other buffers
// or whose pair counts outgrew the capacities) runs with exact list sizes,
// then sizes capacities from the counts, allocates for them in a dry pass
// and captures the whole step; later calls replay the graph and read back
// one overflow flag.  An overflowing replay is recomputed exactly, so every
// call returns the exact-list result.
static void snap_energy_forces_call(ipg_snap* sn, const DevSystem& s, int mode, const int32_t* d_appear,
                                    const double* d_beta, double beta0, int path, double* d_energy,
                                    double* d_forces) {
  ipg_ctx* ctx = sn->ctx;
  SnapGraph& g = sn->graph;
  require_path(sn, path);
  const bool periodic = s.box.per[0] && s.box.per[1] && s.box.per[2];
  const bool use_graph = ctx->graphs && periodic && s.n > 0 && !g.disabled[ctx->prof ? 1 : 0];
  bool keep_graph = false;
  if (use_graph && graph_matches(g, ctx, s, mode, d_appear, d_beta, beta0, path, d_energy, d_forces)) {
    IPG_HIP(hipGraphLaunch(g.exec, ctx->stream));
    int32_t over = 0;
    if (g.pinned) {
      sync(ctx);
      over = *(volatile int32_t*)sn->host_flag;
    } else {
      d2h(ctx, &over, g.flag, sizeof(int32_t));
      sync(ctx);
    }
    g.replays++;
    if (ctx->prof) add_graph_prof(ctx, g.events);
    if (over & 16) fail(IPG_ERR_INVALID, "species index out of range of snap.species");
    if (over == 0) return;
    if (over & 11) g.overflows++;   // a list or a neighbor row outgrew its capacity
    keep_graph = (over & 11) == 0;  // only reverse pairs were missing: the capacities hold
  }
  // exact run: the result of this call
  SnapCall c = snap_prepare(sn, s, mode, d_appear, SNAPL_SYM);
  const bool pinned = energy_forces_dev(sn, c, d_beta, beta0, path, d_energy, d_forces);
  {
    int32_t f = 0;
    if (pinned) {
      sync(ctx);
      f = *(volatile int32_t*)sn->host_flag;
    } else {
      d2h(ctx, &f, c.flag, sizeof(int32_t));
      sync(ctx);
    }
    if (f & 16) fail(IPG_ERR_INVALID, "species index out of range of snap.species");
    if (c.R.rev && (f & 4)) fix_asymmetric(sn, c, path, d_beta, d_forces);
  }
  if (!use_graph || keep_graph) return;
  graph_release(ctx, g);
  const int64_t na = c.L.npairs_a, nb = c.L.list.npairs;
  g.capA = na + na / 8 + 256;
  g.capB = nb + nb / 8 + 256;
  g.avg_nbr = (double)nb / (double)s.n;
  g.ellK = c.L.ellK;
  g.flag = c.flag;  // cleared by the first launch of every step
  bool step_pinned = false;
  auto step = [&] {
    SnapCall cc = snap_prepare(sn, s, mode, d_appear, SNAPL_SYM, g.capA, g.capB, g.ellK);
    cc.R.avg_nbr = g.avg_nbr;
    step_pinned = energy_forces_dev(sn, cc, d_beta, beta0, path, d_energy, d_forces);
  };
  if (!graph_capture(ctx, g, step)) return;  // this call's result came from the exact run
  g.pinned = step_pinned;
  g.n = s.n;
  g.pos = s.pos;
  g.species = s.species;
  g.appear = d_appear;
  g.box = s.box;
  g.mode = mode;
  g.path = path;
  g.beta = d_beta;
  g.beta0 = beta0;
  g.energy = d_energy;
  g.forces = d_forces;
  g.prof = ctx->prof;
  g.prof_gen = ctx->prof_gen;
  g.captures++;
}

}  // namespace ipg

using namespace ipg;

extern "C" {

int ipg_snap_create(ipg_ctx* ctx, const ipg_snap_params* params, ipg_snap** out) {
  return guard([&] {
    IPG_REQUIRE(ctx && params && out, "NULL argument");
    IPG_HIP(hipSetDevice(ctx->device));
    auto* sn = new ipg_snap();
    sn->ctx = ctx;
    try {
      snap_build_host(sn->host, params);
      snap_upload(sn);
      if (hipHostMalloc((void**)&sn->host_flag, 64, hipHostMallocCoherent) != hipSuccess) {
        (void)hipGetLastError();
        sn->host_flag = nullptr;
        fail(IPG_ERR_NOMEM, "hipHostMalloc of the SNAP status word failed");
      }
      *sn->host_flag = 0;
    } catch (...) {
      if (sn->dmem) (void)hipFree(sn->dmem);
      if (sn->host_flag) (void)hipHostFree(sn->host_flag);
      delete sn;
      throw;
    }
    *out = sn;
  });
}

int ipg_snap_destroy(ipg_snap* sn) {
  return guard([&] {
    if (!sn) return;
    if (sn->ctx) {
      (void)hipStreamSynchronize(sn->ctx->stream);
      graph_release(sn->ctx, sn->graph);
    }
    if (sn->dmem) (void)hipFree(sn->dmem);
    if (sn->beta_dev) (void)hipFree(sn->beta_dev);
    if (sn->host_flag) (void)hipHostFree(sn->host_flag);
    delete sn;
  });
}

int ipg_snap_graph_info(const ipg_snap* sn, int32_t* state, int64_t* captures, int64_t* replays,
                        int64_t* overflows, char* reason, int64_t reason_len) {
  return guard([&] {
    IPG_REQUIRE(sn, "snap is NULL");
    const SnapGraph& g = sn->graph;
    const bool off = g.disabled[sn->ctx && sn->ctx->prof ? 1 : 0];
    if (state) *state = off ? 2 : (g.exec ? 1 : 0);
    if (captures) *captures = g.captures;
    if (replays) *replays = g.replays;
    if (overflows) *overflows = g.overflows;
    if (reason && reason_len > 0) {
      const size_t m = std::min<size_t>(g.reason.size(), (size_t)reason_len - 1);
      std::memcpy(reason, g.reason.data(), m);
      reason[m] = 0;
    }
  });
}

int ipg_snap_info(const ipg_snap* sn, int32_t* ncoeff, int32_t* idxu_max, int32_t* idxz_max, int32_t* idxb_max) {
  return guard([&] {
    IPG_REQUIRE(sn, "snap is NULL");
    if (ncoeff) *ncoeff = sn->host.ncoeff;
    if (idxu_max) *idxu_max = sn->host.idxu_max;
    if (idxz_max) *idxz_max = sn->host.idxz_max;
    if (idxb_max) *idxb_max = sn->host.idxb_max;
  });
}

int ipg_snap_neighborlist(ipg_snap* sn, const ipg_system* sys, int mode, const int32_t* appear, int64_t* offsets,
                          int32_t* nbr, double* R, double* r, int64_t cap, int64_t* npairs) {
  return guard([&] {
    IPG_REQUIRE(sn && sys && offsets && npairs, "NULL argument");
    ipg_ctx* ctx = sn->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    check_species_host(sn, sys, appear);
    DevSystem s = upload_system(ctx, sys, "snap_api");
    SnapCall c = snap_prepare(sn, s, mode, upload_appear(ctx, appear, s.n), SNAPL_R);
    d2h(ctx, offsets, c.L.list.off, sizeof(int64_t) * (s.n + 1));
    *npairs = c.L.list.npairs;
    if (c.L.list.npairs > cap) {
      sync(ctx);
      fail(IPG_ERR_CAPACITY, "capacity " + std::to_string(cap) + " < pairs " + std::to_string(c.L.list.npairs));
    }
    const int64_t np = c.L.list.npairs;
    if (nbr) d2h(ctx, nbr, c.L.list.nbr, sizeof(int32_t) * np);
    if (R) d2h(ctx, R, c.L.R, sizeof(double) * np);
    if (r) d2h(ctx, r, c.L.rvec, sizeof(double) * 3 * np);
    sync(ctx);
  });
}

int ipg_snap_local_descriptors(ipg_snap* sn, const ipg_system* sys, int mode, const int32_t* appear, double* B) {
  return guard([&] {
    IPG_REQUIRE(sn && sys && B, "NULL argument");
    ipg_ctx* ctx = sn->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    check_species_host(sn, sys, appear);
    DevSystem s = upload_system(ctx, sys, "snap_api");
    SnapCall c = snap_prepare(sn, s, mode, upload_appear(ctx, appear, s.n), 0);
    if (s.n == 0) return;
    double* dB = scratch_t<double>(ctx, "snap_Bout", (size_t)sn->host.ncoeff * s.n);
    run_descriptors_fp64(sn, c, nullptr, dB, nullptr, nullptr, false);
    d2h(ctx, B, dB, sizeof(double) * sn->host.ncoeff * s.n);
    sync(ctx);
  });
}

int ipg_snap_all_descriptors(ipg_snap* sn, const ipg_system* sys, int mode, const int32_t* appear, double* B,
                             double* dB, double* W) {
  return guard([&] {
    IPG_REQUIRE(sn && sys, "NULL argument");
    ipg_ctx* ctx = sn->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    check_species_host(sn, sys, appear);
    DevSystem s = upload_system(ctx, sys, "snap_api");
    SnapCall c = snap_prepare(sn, s, mode, upload_appear(ctx, appear, s.n), SNAPL_REVERSE);
    const int nc = sn->host.ncoeff, M = nc * sn->host.nspecies;
    if (s.n == 0) {
      if (W) std::fill(W, W + 6 * M, 0.0);
      return;
    }
    double* dBo = scratch_t<double>(ctx, "snap_Bout", (size_t)nc * s.n);
    double* Z = run_descriptors_fp64(sn, c, nullptr, dBo, nullptr, nullptr, dB || W);
    if (B) d2h(ctx, B, dBo, sizeof(double) * nc * s.n);
    if (dB || W) {
      double* dbl = run_dblist(sn, c, Z);
      if (dB) {
        double* d = scratch_t<double>(ctx, "snap_dB", 3 * (size_t)M * s.n);
        gather_db(sn, c, dbl, d);
        d2h(ctx, dB, d, sizeof(double) * 3 * M * s.n);
      }
      if (W) {
        double* Wa = scratch_t<double>(ctx, "snap_Wa", 6 * (size_t)M * s.n);
        double* Wd = scratch_t<double>(ctx, "snap_W", 6 * (size_t)M);
        launch(ctx, "snap_w_atom", [&] {
          k_w_atom<<<grid_for(s.n * (int64_t)M, 256), 256, 0, ctx->stream>>>(sn->dev, c.R, s.pos, dbl, Wa);
        });
        reduce_rows(ctx, Wa, s.n, 6 * M, Wd, false);
        d2h(ctx, W, Wd, sizeof(double) * 6 * M);
      }
    }
    sync(ctx);
  });
}

int ipg_snap_force_descriptors(ipg_snap* sn, const ipg_system* sys, int mode, const int32_t* appear, double* dB) {
  if (!dB) return guard([] { fail(IPG_ERR_INVALID, "dB is NULL"); });
  return ipg_snap_all_descriptors(sn, sys, mode, appear, nullptr, dB, nullptr);
}

int ipg_snap_energy_forces(ipg_snap* sn, const ipg_system* sys, int mode, const int32_t* appear, const double* beta,
                           int64_t nbeta, double beta0, int path, double* energy, double* forces) {
  return guard([&] {
    IPG_REQUIRE(sn && sys, "NULL argument");
    IPG_REQUIRE(std::isfinite(beta0), "beta0 must be finite");
    ipg_ctx* ctx = sn->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    check_species_host(sn, sys, appear);
    const double* db = expand_beta(sn, beta, nbeta);
    DevSystem s = upload_system(ctx, sys, "snap_api");
    const int32_t* d_appear = upload_appear(ctx, appear, s.n);
    double* out = scratch_t<double>(ctx, "snap_api_out", 1 + 3 * std::max<int64_t>(s.n, 1));
    snap_energy_forces_call(sn, s, mode, d_appear, db, beta0, path, out, forces ? out + 1 : nullptr);
    if (energy) d2h(ctx, energy, out, sizeof(double));
    if (forces) d2h(ctx, forces, out + 1, sizeof(double) * 3 * s.n);
    sync(ctx);
  });
}

int ipg_snap_energy_forces_virial(ipg_snap* sn, const ipg_system* sys, int mode, const int32_t* appear,
                                  const double* beta, int64_t nbeta, double beta0, int path, double* energy,
                                  double* forces, double* virial) {
  return guard([&] {
    IPG_REQUIRE(sn && sys, "NULL argument");
    IPG_REQUIRE(std::isfinite(beta0), "beta0 must be finite");
    require_path(sn, path);
    ipg_ctx* ctx = sn->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    check_species_host(sn, sys, appear);
    const double* db = expand_beta(sn, beta, nbeta);
    DevSystem s = upload_system(ctx, sys, "snap_api");
    SnapCall c = snap_prepare(sn, s, mode, upload_appear(ctx, appear, s.n), SNAPL_SYM);
    double* out = scratch_t<double>(ctx, "snap_api_vout", 8 + 3 * std::max<int64_t>(s.n, 1));
    energy_forces_dev(sn, c, db, beta0, path, out, out + 8, out + 1, s.pos);
    int32_t f = 0;
    d2h(ctx, &f, c.flag, sizeof(int32_t));
    sync(ctx);
    // the virial sums own rows only; a missing reverse pair changes forces
    if (c.R.rev && (f & 4)) fix_asymmetric(sn, c, path, db, out + 8);
    double h[7];
    d2h(ctx, h, out, sizeof(double) * 7);
    if (forces) d2h(ctx, forces, out + 8, sizeof(double) * 3 * s.n);
    sync(ctx);
    if (energy) *energy = h[0];
    if (virial)
      for (int k = 0; k < 6; k++) virial[k] = h[1 + k];
  });
}

int ipg_snap_virial_stress(ipg_snap* sn, const ipg_system* sys, int mode, const int32_t* appear, const double* beta,
                           int64_t nbeta, int path, double* virial) {
  if (!virial) return guard([] { fail(IPG_ERR_INVALID, "virial is NULL"); });
  return ipg_snap_energy_forces_virial(sn, sys, mode, appear, beta, nbeta, 0.0, path, nullptr, nullptr, virial);
}

int ipg_snap_energy_forces_dev(ipg_snap* sn, const ipg_system* sys, int mode, const double* beta, double beta0,
                               int path, double* energy, double* forces) {
  return guard([&] {
    IPG_REQUIRE(sn && sys && beta, "NULL argument");
    ipg_ctx* ctx = sn->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    DevSystem s = make_dev_system(sys);
    snap_energy_forces_call(sn, s, mode, nullptr, beta, beta0, path, energy, forces);
  });
}

int ipg_snap_force_vjp(ipg_snap* sn, const ipg_system* sys, int mode, const int32_t* appear, const double* beta,
                       int64_t nbeta, const double* gF, int precision, double* gbeta, double* gpos) {
  return guard([&] {
    IPG_REQUIRE(sn && sys && gF, "NULL argument");
    IPG_REQUIRE(precision == IPG_FP64 || precision == IPG_FP32, "precision must be IPG_FP64 or IPG_FP32");
    if (sn->host.generic)
      fail(IPG_ERR_UNSUPPORTED, "chem_flag = true or twojmax > 8: the force pullback is compiled for 2J <= 8");
    ipg_ctx* ctx = sn->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    check_species_host(sn, sys, appear);
    for (int64_t i = 0; i < 3 * sys->natoms; i++) IPG_REQUIRE(std::isfinite(gF[i]), "gF must be finite");
    const double* db = expand_beta(sn, beta, nbeta);
    DevSystem s = upload_system(ctx, sys, "snap_api");
    SnapCall c = snap_prepare(sn, s, mode, upload_appear(ctx, appear, s.n), SNAPL_REVERSE);
    const SnapDev& P = sn->dev;
    const int64_t n = s.n;
    const int nc = P.ncoeff, M = nc * P.nspecies;
    if (n == 0) {
      if (gbeta) std::fill(gbeta, gbeta + nbeta, 0.0);
      return;
    }
    double* tg = scratch_t<double>(ctx, "snap_vjp_tang", 3 * n);
    h2d(ctx, tg, gF, sizeof(double) * 3 * n);
    c.R.tang = tg;
    double* dedr = scratch_t<double>(ctx, "snap_dedr", 3 * std::max<int64_t>(c.R.npairs, 1));
    double* dedr_t = scratch_t<double>(ctx, "snap_dedr_t", 3 * std::max<int64_t>(c.R.npairs, 1));
    double* tanB = scratch_t<double>(ctx, "snap_tanB", (size_t)M * n);
    if (precision == IPG_FP64) {
      using S = Dual<double>;
      S* U = scratch_t<S>(ctx, "snap_Ud64", 2 * (size_t)P.idxu_max * n);
      S* Y = scratch_t<S>(ctx, "snap_Yd64", 2 * (size_t)P.idxu_max * n);
      double* yc = scratch_t<double>(ctx, "snap_yc64", (size_t)P.nspecies * P.nblk);
      snap_launch_ycoef<double>(ctx, P, db, yc);
      snap_launch_ui<S>(ctx, P, c.R, U);
      snap_launch_zby<S>(ctx, P, c.R, U, nullptr, yc, nullptr, nullptr, Y, nullptr, tanB);
      snap_launch_dedr_rev<S>(ctx, P, c.R, Y, dedr, dedr_t);
    } else {
      using S = Dual<float>;
      S* U = scratch_t<S>(ctx, "snap_Ud32", 2 * (size_t)P.idxu_max * n);
      S* Y = scratch_t<S>(ctx, "snap_Yd32", 2 * (size_t)P.idxu_max * n);
      float* yc = scratch_t<float>(ctx, "snap_yc32", (size_t)P.nspecies * P.nblk);
      snap_launch_ycoef<float>(ctx, P, db, yc);
      snap_launch_ui<S>(ctx, P, c.R, U);
      snap_launch_zby<S>(ctx, P, c.R, U, nullptr, yc, nullptr, nullptr, Y, nullptr, tanB);
      snap_launch_dedr_rev<S>(ctx, P, c.R, Y, dedr, dedr_t);
    }
    double* gp = scratch_t<double>(ctx, "snap_vjp_gpos", 3 * n);
    force_gather(ctx, c.R, dedr_t, gp);
    double* gb = scratch_t<double>(ctx, "snap_vjp_gbeta", M);
    reduce_rows(ctx, tanB, n, M, gb, false);
    std::vector<double> hgb(M);
    d2h(ctx, hgb.data(), gb, sizeof(double) * M);
    if (gpos) d2h(ctx, gpos, gp, sizeof(double) * 3 * n);
    sync(ctx);
    if (gbeta) {
      // gbeta_c = sum_k gF_k . dB[k][c] = -d/de sum_(i in block) B_i[c](x + e gF)
      if (nbeta == nc) {
        for (int cc = 0; cc < nc; cc++) {
          double v = 0.0;
          for (int sp = 0; sp < P.nspecies; sp++) v += hgb[sp * nc + cc];
          gbeta[cc] = -v;
        }
      } else {
        for (int m = 0; m < M; m++) gbeta[m] = -hgb[m];
      }
    }
  });
}

}  // extern "C"
