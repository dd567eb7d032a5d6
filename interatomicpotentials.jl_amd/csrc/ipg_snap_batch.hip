// ipg_snap_batch.hip -- SNAP descriptors of many configurations at once, for
// fitting: the per-atom B (compute_local_descriptors, snap.jl:103-124), the
// per-atom dB (compute_force_descriptors done as compute_all_descriptors
// does it, snap.jl:226-285) and the per-configuration energy rows
// sum_i B_i in species blocks (the rows linear_potential.jl:20-26 contracts
// with beta).
//
// All configurations are packed into one atom range; one set of launches
// covers the whole batch: the neighbor rows (each atom scans its own
// configuration with its own box), U, Z and B, per-pair dB and the gathers.
// Repeated calls with the same buffers replay the whole pipeline from one
// captured HIP graph.  On the first call of a precision, an autotuner times
// the launch shapes that matter (lanes per atom of the neighbor scan, atoms
// per wave of the U kernel, block size of the dB kernel) on the batch
// itself and keeps the fastest.
#include "ipg_snap_driver.h"

#include <algorithm>

namespace ipg {

// out[c][s][e] = sum over the atoms a of configuration c with species s of
// rows[a][e], e < width: the energy rows (rows = B, width ncoeff) and the
// virial rows (rows = per-atom W, width 6 ncoeff).  One block per
// (configuration, chunk of 64 entries): lane e of each of the four waves
// sums every fourth atom, then the four partial sums are added in wave
// order (a fixed order, so repeated calls give identical rows).
constexpr int ROWS_WAVES = 4;
__global__ __launch_bounds__(64 * ROWS_WAVES) void k_batch_rows(const int64_t* __restrict__ start,
                                                               const int32_t* __restrict__ species,
                                                               const double* __restrict__ rows, int width,
                                                               int nspecies, double* __restrict__ out) {
  __shared__ double part[ROWS_WAVES][8][64];
  const int64_t c = blockIdx.x;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t a0 = start[c], a1 = start[c + 1];
  const int e = (int)blockIdx.y * 64 + l;
  double acc[8];
#pragma unroll
  for (int s = 0; s < 8; s++) acc[s] = 0.0;
  if (e < width)
    for (int64_t a = a0 + w; a < a1; a += ROWS_WAVES) {
      const int sa = snap_species(species, a, nspecies);
      const double v = rows[a * width + e];
#pragma unroll
      for (int s = 0; s < 8; s++)
        if (s == sa) acc[s] += v;
    }
#pragma unroll
  for (int s = 0; s < 8; s++) part[w][s][l] = acc[s];
  __syncthreads();
  if (w == 0 && e < width)
    for (int s = 0; s < nspecies; s++) {
      double v = part[0][s][l];
#pragma unroll
      for (int q = 1; q < ROWS_WAVES; q++) v += part[q][s][l];
      out[(c * nspecies + s) * width + e] = v;
    }
}

// Per-atom virial rows of compute_all_descriptors (snap.jl:267-281), own
// species block only: Wa[a][k][v] = sum over the pairs q of atom a of
// dB_q[k][d1(v)] x_a[d2(v)] - dB_q[k][d1(v)] x_j[d2(v)] with raw (unwrapped)
// positions, v = (x x, y y, z z, y z, x z, x y).  One block per atom; lane
// (k, d1) accumulates the three position components, then writes the
// entries whose first index is d1.
template <typename T>
__global__ void k_batch_w_atom(int ncoeff, SnapRun R, const double* __restrict__ pos, const T* __restrict__ dbl,
                               double* __restrict__ Wa) {
  const int64_t a = blockIdx.x;
  const int64_t q0 = R.off[a], q1 = R.off[a + 1];
  const double xa = pos[3 * a], ya = pos[3 * a + 1], za = pos[3 * a + 2];
  for (int t = threadIdx.x; t < 3 * ncoeff; t += blockDim.x) {
    double wx = 0.0, wy = 0.0, wz = 0.0;
    for (int64_t q = q0; q < q1; q++) {
      const int64_t j = R.nbr[q];
      const double d = (double)dbl[q * 3 * ncoeff + t];
      wx += d * xa - d * pos[3 * j];
      wy += d * ya - d * pos[3 * j + 1];
      wz += d * za - d * pos[3 * j + 2];
    }
    const int k = t / 3, d1 = t % 3;
    double* w = Wa + (a * ncoeff + k) * 6;
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
  }
}

// Elapsed milliseconds of the launches named `name` made by f(), timed with
// the context's HIP events; the context's own statistics are left as they
// were.
template <typename F>
static double time_named(ipg_ctx* ctx, const char* name, F&& f) {
  flush_prof(ctx);
  const bool was = ctx->prof;
  std::vector<std::string> only;
  only.swap(ctx->prof_only);
  const auto before = ctx->stats[name];
  ctx->prof = true;
  f();
  flush_prof(ctx);
  const double ms = ctx->stats[name].second - before.second;
  ctx->stats[name] = before;
  ctx->prof = was;
  ctx->prof_only.swap(only);
  return ms;
}

namespace {
struct BatchOut {
  double* B = nullptr;
  double* dB = nullptr;
  double* erow = nullptr;
  double* wrow = nullptr;
};
}  // namespace

// Per-pair dB from Z, then the outputs built from it: the per-atom dB
// gather and the per-configuration virial rows.  Over a symmetric list
// (R.rev) the gather forms the per-atom virial rows in the same pass.
template <typename T>
static void per_pair(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const BatchSys& b, const T* Z,
                     const BatchOut& o, const char* name) {
  if (!o.dB && !o.wrow) return;
  T* dbl = scratch_t<T>(ctx, name, 3 * (size_t)P.ncoeff * std::max<int64_t>(R.npairs, 1));
  if constexpr (std::is_same<T, double>::value) {
    if (P.generic) snap_gen_dbdr(ctx, P, R, Z, dbl);
    else snap_launch_dbdr<T>(ctx, P, R, Z, dbl);
  } else {
    snap_launch_dbdr<T>(ctx, P, R, Z, dbl);
  }
  double* Wa = o.wrow ? scratch_t<double>(ctx, "batch_Wa", 6 * (size_t)P.ncoeff * std::max<int64_t>(b.n, 1)) : nullptr;
  if (R.rev) {
    snap_gather_db<T>(ctx, P, R, dbl, o.dB, b.pos, Wa);
  } else {
    if (o.dB) snap_gather_db<T>(ctx, P, R, dbl, o.dB);
    if (Wa && b.n > 0)
      launch(ctx, "batch_w_atom", [&] {
        k_batch_w_atom<T><<<(unsigned)b.n, 64, 0, ctx->stream>>>(P.ncoeff, R, b.pos, dbl, Wa);
      });
  }
  if (o.wrow) {
    launch(ctx, "batch_wrow", [&] {
      k_batch_rows<<<dim3((unsigned)b.nconf, (6 * P.ncoeff + 63) / 64), 64 * ROWS_WAVES, 0, ctx->stream>>>(
          b.start, b.species, Wa, 6 * P.ncoeff, P.nspecies, o.wrow);
    });
  }
}

// One pass of the pipeline.  cap/ellK > 0: capacity mode (for capture).
// general: general reverse rows instead of the reverse-pair index (for a
// list that a pair within an ulp of the cutoff made asymmetric).  Returns
// the list (pair count and row width, for sizing the capture).
static SnapListDev batch_run(ipg_snap* sn, const BatchSys& b, int fp32, const BatchOut& o, int64_t cap, int ellK,
                             const BatchTune& tn, int32_t* flag, bool general) {
  ipg_ctx* ctx = sn->ctx;
  const SnapDev& P = sn->dev;
  const SnapCut cut = snap_make_cut(sn->host, IPG_SNAP_LIST_CUTOFF);
  zero(ctx, flag, sizeof(int32_t));
  SnapListDev L = build_snap_list_batch(ctx, b, cut, general ? SNAPL_REVERSE : SNAPL_SYM, "snap_batch", cap, flag,
                                        ellK, tn.scan_lp);
  SnapRun R;
  R.n = b.n;
  R.npairs = L.list.npairs;
  R.d_np = L.list.exact ? nullptr : L.list.off + b.n;
  R.off = L.list.off;
  R.nbr = L.list.nbr;
  R.center = L.center;
  R.rvec = L.rvec;
  R.species = b.species;
  R.roff = L.roff;
  R.rpair = L.rpair;
  R.rev = L.rev;
  R.apw = tn.ui_apw;
  R.dbdr_threads = tn.dbdr_threads;
  R.flag = flag;
  const int64_t n = std::max<int64_t>(b.n, 1);
  double* Bv = o.B ? o.B : scratch_t<double>(ctx, "batch_B", (size_t)n * P.ncoeff);
  if (fp32) {
    const size_t nb32 = (size_t)((b.n + 31) / 32) * 32;
    float* U = scratch_t<float>(ctx, "batch_U32", 2 * (size_t)P.idxu_max * std::max<size_t>(nb32, 32));
    float* Z = scratch_t<float>(ctx, "batch_Z32", 2 * (size_t)P.idxz_max * n);
    snap_launch_ui<float>(ctx, P, R, U);
    snap_launch_zb32(ctx, P, R, U, Z, Bv);
    per_pair<float>(ctx, P, R, b, Z, o, "batch_dbl32");
  } else if (P.generic) {  // runtime tables (chem_flag, 2J > 8)
    double* U = scratch_t<double>(ctx, "batch_U64", 2 * (size_t)P.nel * P.idxu_max * n);
    double* Z = scratch_t<double>(ctx, "batch_Z64", 2 * (size_t)P.idxz_max * n);
    snap_gen_descriptors(ctx, P, R, U, Z, Bv, nullptr, nullptr);
    per_pair<double>(ctx, P, R, b, Z, o, "batch_dbl64");
  } else {
    double* U = scratch_t<double>(ctx, "batch_U64", 2 * (size_t)P.idxu_max * n);
    double* Z = scratch_t<double>(ctx, "batch_Z64", 2 * (size_t)P.idxz_max * n);
    snap_launch_ui<double>(ctx, P, R, U);
    snap_launch_zby<double>(ctx, P, R, U, nullptr, nullptr, Bv, nullptr, nullptr, Z, nullptr);
    per_pair<double>(ctx, P, R, b, Z, o, "batch_dbl64");
  }
  if (o.erow)
    launch(ctx, "batch_erow", [&] {
      k_batch_rows<<<dim3((unsigned)b.nconf, (P.ncoeff + 63) / 64), 64 * ROWS_WAVES, 0, ctx->stream>>>(
          b.start, b.species, Bv, P.ncoeff, P.nspecies, o.erow);
    });
  return L;
}

// Autotuner: times the candidate launch shapes on this batch (exact
// sizing) and keeps the fastest of each.
static void batch_tune(ipg_snap* sn, const BatchSys& b, int fp32, BatchTune& tn, int32_t* flag) {
  ipg_ctx* ctx = sn->ctx;
  const SnapDev& P = sn->dev;
  const SnapCut cut = snap_make_cut(sn->host, IPG_SNAP_LIST_CUTOFF);
  const int lps[3] = {16, 32, 64};
  int best = 0;
  for (int k = 0; k < 3; k++) {
    zero(ctx, flag, sizeof(int32_t));
    tn.ms_scan[k] = time_named(ctx, "batch_scan", [&] {
      build_snap_list_batch(ctx, b, cut, SNAPL_SYM, "snap_batch", 0, flag, 0, lps[k]);
    });
    if (tn.ms_scan[k] < tn.ms_scan[best]) best = k;
  }
  tn.scan_lp = lps[best];
  if (P.generic) {  // the runtime-table kernels have one launch shape
    tn.ui_apw = 1;
    tn.dbdr_threads = 64;
    return;
  }
  zero(ctx, flag, sizeof(int32_t));
  SnapListDev L = build_snap_list_batch(ctx, b, cut, SNAPL_SYM, "snap_batch", 0, flag, 0, tn.scan_lp);
  SnapRun R;
  R.n = b.n;
  R.npairs = L.list.npairs;
  R.off = L.list.off;
  R.nbr = L.list.nbr;
  R.center = L.center;
  R.rvec = L.rvec;
  R.species = b.species;
  R.roff = L.roff;
  R.rpair = L.rpair;
  const int64_t n = std::max<int64_t>(b.n, 1), npa = std::max<int64_t>(R.npairs, 1);
  const char* nui = fp32 ? "snap_ui_fp32" : "snap_ui_fp64";
  const char* ndb = fp32 ? "snap_dbdr_fp32" : "snap_dbdr_fp64";
  double* Bv = scratch_t<double>(ctx, "batch_B", (size_t)n * P.ncoeff);
  if (fp32) {
    const size_t nb32 = (size_t)((b.n + 31) / 32) * 32;
    float* U = scratch_t<float>(ctx, "batch_U32", 2 * (size_t)P.idxu_max * std::max<size_t>(nb32, 32));
    float* Z = scratch_t<float>(ctx, "batch_Z32", 2 * (size_t)P.idxz_max * n);
    float* dbl = scratch_t<float>(ctx, "batch_dbl32", 3 * (size_t)P.ncoeff * npa);
    for (int a = 1; a <= 2; a++) {
      R.apw = a;
      tn.ms_ui[a - 1] = time_named(ctx, nui, [&] { snap_launch_ui<float>(ctx, P, R, U); });
    }
    R.apw = tn.ms_ui[0] <= tn.ms_ui[1] ? 1 : 2;
    snap_launch_ui<float>(ctx, P, R, U);
    snap_launch_zb32(ctx, P, R, U, Z, Bv);
    for (int k = 0; k < 2; k++) {
      R.dbdr_threads = k == 0 ? 64 : 128;
      tn.ms_dbdr[k] = time_named(ctx, ndb, [&] { snap_launch_dbdr<float>(ctx, P, R, Z, dbl); });
    }
  } else {
    double* U = scratch_t<double>(ctx, "batch_U64", 2 * (size_t)P.idxu_max * n);
    double* Z = scratch_t<double>(ctx, "batch_Z64", 2 * (size_t)P.idxz_max * n);
    double* dbl = scratch_t<double>(ctx, "batch_dbl64", 3 * (size_t)P.ncoeff * npa);
    for (int a = 1; a <= 2; a++) {
      R.apw = a;
      tn.ms_ui[a - 1] = time_named(ctx, nui, [&] { snap_launch_ui<double>(ctx, P, R, U); });
    }
    R.apw = tn.ms_ui[0] <= tn.ms_ui[1] ? 1 : 2;
    snap_launch_ui<double>(ctx, P, R, U);
    snap_launch_zby<double>(ctx, P, R, U, nullptr, nullptr, Bv, nullptr, nullptr, Z, nullptr);
    for (int k = 0; k < 2; k++) {
      R.dbdr_threads = k == 0 ? 64 : 128;
      tn.ms_dbdr[k] = time_named(ctx, ndb, [&] { snap_launch_dbdr<double>(ctx, P, R, Z, dbl); });
    }
  }
  tn.ui_apw = tn.ms_ui[0] <= tn.ms_ui[1] ? 1 : 2;
  tn.dbdr_threads = tn.ms_dbdr[0] <= tn.ms_dbdr[1] ? 64 : 128;
}

// Device-side batch layout (configuration of every atom, configuration
// starts and boxes) in context scratch, uploaded only when it changed.
static BatchSys batch_layout(ipg_snap* sn, int64_t nconf, const int64_t* natoms, const double* box,
                             const int32_t* periodic, const double* d_pos, const int32_t* d_species) {
  ipg_ctx* ctx = sn->ctx;
  IPG_REQUIRE(nconf >= 1, "nconf must be >= 1");
  IPG_REQUIRE(natoms && box && periodic, "NULL batch layout");
  BatchSys b;
  b.nconf = nconf;
  int64_t n = 0;
  double expect = 0.0;
  const double rc = snap_make_cut(sn->host, IPG_SNAP_LIST_CUTOFF).query;
  for (int64_t c = 0; c < nconf; c++) {
    IPG_REQUIRE(natoms[c] >= 0, "natoms must be >= 0");
    double vol = 1.0;
    for (int d = 0; d < 3; d++) {
      if (periodic[3 * c + d])
        IPG_REQUIRE(std::isfinite(box[3 * c + d]) && box[3 * c + d] > 0.0,
                    "box length of a periodic dimension must be finite and > 0");
      vol *= periodic[3 * c + d] ? box[3 * c + d] : 0.0;
    }
    const double dens = vol > 0.0 ? (double)natoms[c] / vol : 0.0;
    expect = std::max(expect, vol > 0.0 ? std::min<double>((double)natoms[c], dens * 4.18879020478639 * rc * rc * rc)
                                        : (double)natoms[c]);
    n += natoms[c];
  }
  IPG_REQUIRE(n < (int64_t)INT32_MAX, "too many atoms in the batch");
  b.n = n;
  b.expect = expect;
  b.pos = d_pos;
  b.species = d_species;
  int32_t* cfg = scratch_t<int32_t>(ctx, "batch_cfg", std::max<int64_t>(n, 1));
  int64_t* start = scratch_t<int64_t>(ctx, "batch_start", nconf + 1);
  IpgBox* boxes = scratch_t<IpgBox>(ctx, "batch_boxes", nconf);
  // the device copy is still valid when no scratch buffer moved since the
  // call that uploaded it and the layout is the same
  BatchLayout& hl = sn->blayout;
  const bool same = hl.gen == ctx->scratch_gen && hl.natoms.size() == (size_t)nconf &&
                    std::equal(natoms, natoms + nconf, hl.natoms.begin()) &&
                    std::equal(box, box + 3 * nconf, hl.box.begin()) &&
                    std::equal(periodic, periodic + 3 * nconf, hl.per.begin());
  if (!same) {
    std::vector<int32_t> hcfg((size_t)n);
    std::vector<int64_t> hst((size_t)nconf + 1, 0);
    std::vector<IpgBox> hb((size_t)nconf);
    for (int64_t c = 0; c < nconf; c++) {
      hst[c + 1] = hst[c] + natoms[c];
      for (int64_t a = hst[c]; a < hst[c + 1]; a++) hcfg[a] = (int32_t)c;
      for (int d = 0; d < 3; d++) {
        hb[c].per[d] = periodic[3 * c + d] ? 1 : 0;
        hb[c].L[d] = periodic[3 * c + d] ? box[3 * c + d] : INFINITY;
      }
    }
    h2d(ctx, cfg, hcfg.data(), sizeof(int32_t) * n);
    h2d(ctx, start, hst.data(), sizeof(int64_t) * (nconf + 1));
    h2d(ctx, boxes, hb.data(), sizeof(IpgBox) * nconf);
    sync(ctx);  // the host vectors go out of scope
    hl.natoms.assign(natoms, natoms + nconf);
    hl.box.assign(box, box + 3 * nconf);
    hl.per.assign(periodic, periodic + 3 * nconf);
  }
  b.cfg = cfg;
  b.start = start;
  b.boxes = boxes;
  return b;
}

static void batch_descriptors(ipg_snap* sn, const BatchSys& b, int precision, const BatchOut& o) {
  ipg_ctx* ctx = sn->ctx;
  IPG_REQUIRE(precision == IPG_FP32 || precision == IPG_FP64, "precision must be IPG_FP64 or IPG_FP32");
  // chem_flag (and with it bnorm_flag) and 2J > 8 run the runtime-table FP64 kernels
  if (sn->host.generic && precision != IPG_FP64)
    fail(IPG_ERR_UNSUPPORTED, "chem_flag = true or twojmax > 8: the batched descriptors run in FP64 only");
  const int fp32 = precision == IPG_FP32 ? 1 : 0;
  int32_t* flag = scratch_t<int32_t>(ctx, "batch_flag", 1);
  BatchTune& tn = sn->btune[fp32];
  SnapGraph& g = sn->bgraph;
  const bool use_graph = ctx->graphs && b.n > 0 && !g.disabled[ctx->prof ? 1 : 0];
  const bool match = g.exec && sn->bprec == precision && g.n == b.n && g.pos == b.pos && g.species == b.species &&
                     g.energy == o.B && g.forces == o.dB && g.beta == o.erow && g.aux == o.wrow &&
                     g.mode == (int)b.nconf &&
                     g.prof == ctx->prof && g.prof_gen == ctx->prof_gen && g.gen == ctx->scratch_gen;
  if (use_graph && match) {
    IPG_HIP(hipGraphLaunch(g.exec, ctx->stream));
    int32_t over = 0;
    d2h(ctx, &over, g.flag, sizeof(int32_t));
    sync(ctx);
    g.replays++;
    if (ctx->prof) add_graph_prof(ctx, g.events);
    if (over & 16) fail(IPG_ERR_INVALID, "species index out of range of snap.species");
    if (over == 0) return;
    g.overflows++;
  }
  if (!tn.scan_lp && b.n > 0) batch_tune(sn, b, fp32, tn, flag);
  // exact run: the result of this call (again with the general reverse rows
  // when a pair within an ulp of the cutoff left the list asymmetric)
  bool general = false;
  SnapListDev L = batch_run(sn, b, fp32, o, 0, 0, tn, flag, false);
  int32_t f = 0;
  d2h(ctx, &f, flag, sizeof(int32_t));
  sync(ctx);
  if (f & 16) fail(IPG_ERR_INVALID, "species index out of range of snap.species");
  if (f & 4) {
    general = true;
    L = batch_run(sn, b, fp32, o, 0, 0, tn, flag, true);
  }
  sn->bpairs = L.list.npairs;
  if (!use_graph) return;
  graph_release(ctx, g);
  const int64_t np = L.list.npairs;
  g.capB = np + np / 8 + 256;
  g.ellK = L.ellK;
  g.flag = flag;
  auto step = [&] { batch_run(sn, b, fp32, o, g.capB, g.ellK, tn, flag, general); };
  if (!graph_capture(ctx, g, step)) return;
  sn->bprec = precision;
  g.n = b.n;
  g.pos = b.pos;
  g.species = b.species;
  g.energy = o.B;
  g.forces = o.dB;
  g.beta = o.erow;
  g.aux = o.wrow;
  g.mode = (int)b.nconf;
  g.prof = ctx->prof;
  g.prof_gen = ctx->prof_gen;
  g.captures++;
}

}  // namespace ipg

using namespace ipg;

extern "C" {

int ipg_snap_batch_descriptors_dev(ipg_snap* sn, int64_t nconf, const int64_t* natoms, const double* box,
                                   const int32_t* periodic, const double* pos, const int32_t* species, int precision,
                                   double* B, double* dB, double* erow, double* wrow) {
  return guard([&] {
    IPG_REQUIRE(sn, "snap is NULL");
    ipg_ctx* ctx = sn->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    BatchSys b = batch_layout(sn, nconf, natoms, box, periodic, pos, species);
    IPG_REQUIRE(b.n == 0 || pos, "positions are NULL");
    BatchOut o{B, dB, erow, wrow};
    batch_descriptors(sn, b, precision, o);
    sn->blayout.gen = ctx->scratch_gen;
  });
}

int ipg_snap_batch_descriptors(ipg_snap* sn, int64_t nconf, const ipg_system* systems, int precision, double* B,
                               double* dB, double* erow, double* wrow) {
  return guard([&] {
    IPG_REQUIRE(sn && systems, "NULL argument");
    IPG_REQUIRE(nconf >= 1, "nconf must be >= 1");
    ipg_ctx* ctx = sn->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    std::vector<int64_t> na((size_t)nconf);
    std::vector<double> bx((size_t)nconf * 3);
    std::vector<int32_t> per((size_t)nconf * 3);
    int64_t n = 0;
    for (int64_t c = 0; c < nconf; c++) {
      check_system_host(systems + c);
      if (systems[c].species)
        for (int64_t a = 0; a < systems[c].natoms; a++)
          IPG_REQUIRE(systems[c].species[a] >= 0 && systems[c].species[a] < sn->host.nspecies,
                      "species index out of range of snap.species");
      na[c] = systems[c].natoms;
      for (int d = 0; d < 3; d++) {
        bx[3 * c + d] = systems[c].box[d];
        per[3 * c + d] = systems[c].periodic[d];
      }
      n += na[c];
    }
    std::vector<double> hp((size_t)std::max<int64_t>(n, 1) * 3);
    std::vector<int32_t> hs((size_t)std::max<int64_t>(n, 1), 0);
    int64_t o = 0;
    for (int64_t c = 0; c < nconf; c++) {
      std::copy(systems[c].positions, systems[c].positions + 3 * na[c], hp.begin() + 3 * o);
      if (systems[c].species) std::copy(systems[c].species, systems[c].species + na[c], hs.begin() + o);
      o += na[c];
    }
    double* dpos = scratch_t<double>(ctx, "batch_api_pos", hp.size());
    int32_t* dsp = scratch_t<int32_t>(ctx, "batch_api_species", hs.size());
    h2d(ctx, dpos, hp.data(), sizeof(double) * 3 * n);
    h2d(ctx, dsp, hs.data(), sizeof(int32_t) * n);
    const int M = sn->host.ncoeff * sn->host.nspecies;
    double* dBo = B ? scratch_t<double>(ctx, "batch_api_B", (size_t)std::max<int64_t>(n, 1) * sn->host.ncoeff)
                    : nullptr;
    double* ddB = dB ? scratch_t<double>(ctx, "batch_api_dB", (size_t)std::max<int64_t>(n, 1) * M * 3) : nullptr;
    double* der = erow ? scratch_t<double>(ctx, "batch_api_erow", (size_t)nconf * M) : nullptr;
    double* dwr = wrow ? scratch_t<double>(ctx, "batch_api_wrow", (size_t)nconf * M * 6) : nullptr;
    BatchSys b = batch_layout(sn, nconf, na.data(), bx.data(), per.data(), dpos, dsp);
    batch_descriptors(sn, b, precision, BatchOut{dBo, ddB, der, dwr});
    sn->blayout.gen = ctx->scratch_gen;
    if (B) d2h(ctx, B, dBo, sizeof(double) * n * sn->host.ncoeff);
    if (dB) d2h(ctx, dB, ddB, sizeof(double) * n * M * 3);
    if (erow) d2h(ctx, erow, der, sizeof(double) * nconf * M);
    if (wrow) d2h(ctx, wrow, dwr, sizeof(double) * nconf * M * 6);
    sync(ctx);
  });
}

int ipg_snap_batch_info(const ipg_snap* sn, int precision, int32_t* scan_lanes, int32_t* ui_atoms_per_wave,
                        int32_t* dbdr_threads, int32_t* graph_state, int64_t* replays, int64_t* npairs) {
  return guard([&] {
    IPG_REQUIRE(sn, "snap is NULL");
    IPG_REQUIRE(precision == IPG_FP32 || precision == IPG_FP64, "precision must be IPG_FP64 or IPG_FP32");
    const BatchTune& tn = sn->btune[precision == IPG_FP32 ? 1 : 0];
    if (scan_lanes) *scan_lanes = tn.scan_lp;
    if (ui_atoms_per_wave) *ui_atoms_per_wave = tn.ui_apw;
    if (dbdr_threads) *dbdr_threads = tn.dbdr_threads;
    const SnapGraph& g = sn->bgraph;
    if (graph_state) *graph_state = g.disabled[sn->ctx && sn->ctx->prof ? 1 : 0] ? 2 : (g.exec ? 1 : 0);
    if (replays) *replays = g.replays;
    if (npairs) *npairs = sn->bpairs;
  });
}

}  // extern "C"
