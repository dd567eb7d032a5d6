// ipg_nlist.hip -- cell-binned neighbor lists with the reference's exact
// membership (src/nnlist.jl:23-46, src/BasisSystems/SNAP/utilities/utl.jl:38-101).
//
// Pipeline: bin atoms on wrapped coordinates into cells at least rcut wide
// -> counting sort by cell (atomics for slots, then each cell's atoms are
// put in index order, so the layout is deterministic) -> per-atom count over
// the 27 surrounding cells -> exclusive scan -> per-atom fill -> optional
// per-atom sort of j and displacement pass.  Membership is decided on the
// original coordinates with the PeriodicEuclidean formula and the
// sqrt-free threshold s <= smax, where smax is the largest double whose
// correctly rounded sqrt is <= rcut; this equals sqrt(s) <= rcut exactly.
#include "ipg_nlist.h"

#include <algorithm>

namespace ipg {

struct Grid {
  IpgBox box;
  double lo[3], w[3];
  int nc[3];
  int full[3];  // periodic dimension with < 3 cells: visit all of them once
  int64_t ncell;
};

__device__ __forceinline__ int cell_coord(const Grid& g, int d, double x) {
  double u;
  if (g.box.per[d]) u = x - floor(x / g.box.L[d]) * g.box.L[d];
  else u = x - g.lo[d];
  int c = (int)floor(u / g.w[d]);
  c = c < 0 ? 0 : c;
  return c >= g.nc[d] ? g.nc[d] - 1 : c;
}

__global__ void k_cell_assign(const double* __restrict__ pos, int64_t n, Grid g, int32_t* __restrict__ cellof,
                              int32_t* __restrict__ ccount) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c0 = cell_coord(g, 0, pos[3 * i]);
  int c1 = cell_coord(g, 1, pos[3 * i + 1]);
  int c2 = cell_coord(g, 2, pos[3 * i + 2]);
  int32_t c = (int32_t)(((int64_t)c2 * g.nc[1] + c1) * g.nc[0] + c0);
  cellof[i] = c;
  atomicAdd(&ccount[c], 1);
}

__global__ void k_cell_place(int64_t n, const int32_t* __restrict__ cellof, const int64_t* __restrict__ cstart,
                             int32_t* __restrict__ cfill, int32_t* __restrict__ ctmp) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t c = cellof[i];
  int32_t slot = atomicAdd(&cfill[c], 1);
  ctmp[cstart[c] + slot] = (int32_t)i;
}

// Order each cell's atoms by index (rank of every entry within its cell, so
// the layout does not depend on the atomic slot order) and gather positions
// in cell order: exact doubles for the membership test and FP32 wrapped
// coordinates for the prefilter.
__global__ void k_cell_rank(int64_t n, const int32_t* __restrict__ ctmp, const int32_t* __restrict__ cellof,
                            const int64_t* __restrict__ cstart, const double* __restrict__ pos, Grid g,
                            int32_t* __restrict__ catom, double4* __restrict__ spos, float4* __restrict__ sposf) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t i = ctmp[k];
  const int32_t c = cellof[i];
  const int64_t a = cstart[c], b = cstart[c + 1];
  int64_t r = 0;
  for (int64_t m = a; m < b; m++) r += ctmp[m] < i;
  const int64_t dst = a + r;
  catom[dst] = i;
  const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
  spos[dst] = make_double4(x, y, z, 0.0);
  double w[3] = {x, y, z};
  for (int d = 0; d < 3; d++) w[d] = g.box.per[d] ? w[d] - floor(w[d] / g.box.L[d]) * g.box.L[d] : w[d] - g.lo[d];
  sposf[dst] = make_float4((float)w[0], (float)w[1], (float)w[2], 0.f);
}

template <typename F>
__device__ __forceinline__ void for_each_cell(const Grid& g, int32_t cell, const int64_t* __restrict__ cstart,
                                              F&& f) {
  int cc[3];
  cc[0] = cell % g.nc[0];
  cc[1] = (cell / g.nc[0]) % g.nc[1];
  cc[2] = cell / (g.nc[0] * g.nc[1]);
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (g.full[d]) {
      lo[d] = 0;
      hi[d] = g.nc[d] - 1;
    } else {
      lo[d] = cc[d] - 1;
      hi[d] = cc[d] + 1;
    }
  }
  for (int z = lo[2]; z <= hi[2]; z++) {
    int zz = z;
    if (!g.full[2]) {
      if (g.box.per[2]) zz = (z + g.nc[2]) % g.nc[2];
      else if (z < 0 || z >= g.nc[2]) continue;
    }
    for (int y = lo[1]; y <= hi[1]; y++) {
      int yy = y;
      if (!g.full[1]) {
        if (g.box.per[1]) yy = (y + g.nc[1]) % g.nc[1];
        else if (y < 0 || y >= g.nc[1]) continue;
      }
      for (int x = lo[0]; x <= hi[0]; x++) {
        int xx = x;
        if (!g.full[0]) {
          if (g.box.per[0]) xx = (x + g.nc[0]) % g.nc[0];
          else if (x < 0 || x >= g.nc[0]) continue;
        }
        const int64_t cid = ((int64_t)zz * g.nc[1] + yy) * g.nc[0] + xx;
        f(cstart[cid], cstart[cid + 1]);
      }
    }
  }
}

__device__ __forceinline__ bool keep_pair(int kind, int32_t i, int32_t j) {
  if (kind == KIND_HALF) return j > i;
  if (kind == KIND_FULL) return j != i;
  return true;  // KIND_FULL_SELF
}

struct Prefilter {
  float L[3], invL[3];
  float r2;  // FP32 squared-distance bound that no member can exceed
};

// LP lanes per atom scan the surrounding cells; each cell's candidates are
// taken LP at a time.  A cheap FP32 minimum-image distance rejects far
// candidates; the rest get the exact FP64 PeriodicEuclidean test.  Hits are
// placed with a ballot, so the output order is fixed (cell order, then slot).
template <int LP, bool FILL>
__global__ __launch_bounds__(256) void k_list_scan(int64_t n, Grid g, Prefilter pf, const int32_t* __restrict__ catom,
                                                   const double4* __restrict__ spos, const float4* __restrict__ sposf,
                                                   const int32_t* __restrict__ cellof,
                                                   const int64_t* __restrict__ cstart, double smax, int kind,
                                                   int32_t* __restrict__ cnt, const int64_t* __restrict__ off,
                                                   int32_t* __restrict__ nbr, int64_t cap) {
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t k = gt / LP;
  const int l = (int)(gt % LP);
  const int lane = threadIdx.x & 63;
  if (k >= n) return;  // whole groups leave together
  const uint64_t gmask = (LP == 64 ? ~0ull : ((1ull << LP) - 1ull)) << (lane - l);
  const uint64_t below = (1ull << lane) - 1ull;
  const int32_t i = catom[k];
  const double4 pi = spos[k];
  const float4 qi = sposf[k];
  int64_t w = FILL ? off[i] : 0;
  int count = 0;
  for_each_cell(g, cellof[i], cstart, [&](int64_t s0, int64_t s1) {
    for (int64_t s = s0; s < s1; s += LP) {
      const int64_t sl = s + l;
      bool hit = false;
      int32_t j = -1;
      if (sl < s1) {
        j = catom[sl];
        if (keep_pair(kind, i, j)) {
          const float4 qj = sposf[sl];
          float dx = qj.x - qi.x, dy = qj.y - qi.y, dz = qj.z - qi.z;
          if (g.box.per[0]) dx -= pf.L[0] * rintf(dx * pf.invL[0]);
          if (g.box.per[1]) dy -= pf.L[1] * rintf(dy * pf.invL[1]);
          if (g.box.per[2]) dz -= pf.L[2] * rintf(dz * pf.invL[2]);
          if (dx * dx + dy * dy + dz * dz <= pf.r2) {
            const double4 pj = spos[sl];
            hit = ipg_pe_sumsq(g.box, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z) <= smax;
          }
        }
      }
      const uint64_t m = __ballot(hit) & gmask;
      if (FILL) {
        if (hit && w + __popcll(m & below) < cap) nbr[w + __popcll(m & below)] = j;
        w += __popcll(m);
      } else {
        count += __popcll(m);
      }
    }
  });
  if (!FILL && l == 0) cnt[i] = count;
}

// In-segment rank sort: out[off + rank] = in[q], rank = number of smaller
// entries of the same segment (values within a segment are distinct).
__global__ void k_seg_rank_sort(int64_t nseg, const int64_t* __restrict__ off, const int32_t* __restrict__ in,
                                int32_t* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= off[nseg]) return;  // off[nseg] is the device total
  int64_t lo = 0, hi = nseg;  // find the segment with off[seg] <= q < off[seg+1]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= q) lo = mid;
    else hi = mid;
  }
  const int32_t v = in[q];
  int64_t r = 0;
  for (int64_t m = off[lo]; m < off[lo + 1]; m++) r += in[m] < v;
  out[off[lo] + r] = v;
}

__global__ void k_list_disp(int64_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ nbr,
                            const double* __restrict__ pos, IpgBox box, double* __restrict__ disp) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  for (int64_t k = off[i]; k < off[i + 1]; k++) {
    const int32_t j = nbr[k];
    disp[3 * k] = ipg_disp1(xi, pos[3 * j], box.L[0], box.per[0]);
    disp[3 * k + 1] = ipg_disp1(yi, pos[3 * j + 1], box.L[1], box.per[1]);
    disp[3 * k + 2] = ipg_disp1(zi, pos[3 * j + 2], box.L[2], box.per[2]);
  }
}

// per-dimension min/max of positions (for non-periodic extents)
__global__ void k_minmax(const double* __restrict__ pos, int64_t n, double* __restrict__ part) {
  __shared__ double smin[3][256], smax[3][256];
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    for (int d = 0; d < 3; d++) {
      double v = pos[3 * i + d];
      mn[d] = fmin(mn[d], v);
      mx[d] = fmax(mx[d], v);
    }
  for (int d = 0; d < 3; d++) {
    smin[d][threadIdx.x] = mn[d];
    smax[d][threadIdx.x] = mx[d];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int d = 0; d < 3; d++) {
        smin[d][threadIdx.x] = fmin(smin[d][threadIdx.x], smin[d][threadIdx.x + s]);
        smax[d][threadIdx.x] = fmax(smax[d][threadIdx.x], smax[d][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; d++) {
      part[blockIdx.x * 6 + d] = smin[d][0];
      part[blockIdx.x * 6 + 3 + d] = smax[d][0];
    }
}

static void extents(ipg_ctx* ctx, const DevSystem& s, double lo[3], double hi[3]) {
  const int nb = 64;
  double* part = scratch_t<double>(ctx, "minmax_part", nb * 6);
  launch(ctx, "minmax", [&] { k_minmax<<<nb, 256, 0, ctx->stream>>>(s.pos, s.n, part); });
  std::vector<double> h(nb * 6);
  d2h(ctx, h.data(), part, sizeof(double) * nb * 6);
  sync(ctx);
  for (int d = 0; d < 3; d++) {
    lo[d] = INFINITY;
    hi[d] = -INFINITY;
    for (int b = 0; b < nb; b++) {
      lo[d] = std::fmin(lo[d], h[b * 6 + d]);
      hi[d] = std::fmax(hi[d], h[b * 6 + 3 + d]);
    }
  }
}

static Grid make_grid(ipg_ctx* ctx, const DevSystem& s, double rcut) {
  Grid g{};
  g.box = s.box;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  bool need = !(s.box.per[0] && s.box.per[1] && s.box.per[2]);
  if (need && s.n > 0) extents(ctx, s, lo, hi);
  double ext[3];
  for (int d = 0; d < 3; d++) {
    if (s.box.per[d]) {
      g.lo[d] = 0.0;
      ext[d] = s.box.L[d];
    } else {
      IPG_REQUIRE(std::isfinite(lo[d]) && std::isfinite(hi[d]), "positions must be finite");
      g.lo[d] = lo[d];
      ext[d] = hi[d] - lo[d];
    }
  }
  double cw = rcut * (1.0 + 1e-9);
  int64_t cap = 4 * s.n > 64 ? 4 * s.n : 64;
  for (;;) {
    int64_t prod = 1;
    for (int d = 0; d < 3; d++) {
      double q = std::floor(ext[d] / cw);
      int c = q < 1 ? 1 : (q > 4096 ? 4096 : (int)q);
      g.nc[d] = c;
      prod *= c;
    }
    if (prod <= cap) break;
    cw *= 1.25;
  }
  g.ncell = 1;
  for (int d = 0; d < 3; d++) {
    g.w[d] = g.nc[d] == 1 ? (ext[d] > 0 ? ext[d] * (1.0 + 1e-12) + 1e-300 : 1.0) : ext[d] / g.nc[d];
    g.full[d] = (s.box.per[d] && g.nc[d] < 3) ? 1 : 0;
    g.ncell *= g.nc[d];
  }
  return g;
}

DevSystem make_dev_system(const ipg_system* s) {
  check_system_shape(s);
  DevSystem d;
  d.n = s->natoms;
  d.pos = s->positions;
  d.species = s->species;
  for (int k = 0; k < 3; k++) {
    d.box.per[k] = s->periodic[k] ? 1 : 0;
    d.box.L[k] = s->periodic[k] ? s->box[k] : INFINITY;
  }
  return d;
}

DevSystem upload_system(ipg_ctx* ctx, const ipg_system* s, const char* tag) {
  check_system_host(s);
  DevSystem d = make_dev_system(s);
  std::string t(tag);
  double* p = scratch_t<double>(ctx, (t + "_pos").c_str(), 3 * (d.n > 0 ? d.n : 1));
  h2d(ctx, p, s->positions, sizeof(double) * 3 * d.n);
  d.pos = p;
  if (s->species) {
    int32_t* sp = scratch_t<int32_t>(ctx, (t + "_species").c_str(), d.n > 0 ? d.n : 1);
    h2d(ctx, sp, s->species, sizeof(int32_t) * d.n);
    d.species = sp;
  } else {
    d.species = nullptr;
  }
  return d;
}

// Capacity mode: the scanned offsets are clamped to the capacity so every
// later reader stays inside the buffers, and `bit` is raised in *flag when
// the true total exceeded it (the results of that call are then discarded).
__global__ void k_cap_clamp(int64_t n, int64_t* __restrict__ off, int64_t cap, int32_t* __restrict__ flag,
                            int32_t bit) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int64_t v = off[i];
  if (v > cap) {
    off[i] = cap;
    if (i == n) atomicOr(flag, bit);
  }
}

static void cap_clamp(ipg_ctx* ctx, int64_t n, int64_t* off, int64_t cap, int32_t* flag, int32_t bit) {
  IPG_REQUIRE(flag != nullptr, "internal: capacity mode without an overflow flag");
  launch(ctx, "cap_clamp", [&] { k_cap_clamp<<<grid_for(n + 1, 256), 256, 0, ctx->stream>>>(n, off, cap, flag, bit); });
}

// Common list builder; alloc(name, bytes) supplies device storage.
template <typename Alloc>
static DevList build_list_impl(ipg_ctx* ctx, const DevSystem& s, double rcut, int kind, bool sorted, bool want_disp,
                               int64_t cap, int32_t* capflag, Alloc&& alloc) {
  IPG_REQUIRE(std::isfinite(rcut) && rcut > 0, "rcutoff must be finite and > 0");
  DevList L;
  L.n = s.n;
  L.off = (int64_t*)alloc("off", sizeof(int64_t) * (s.n + 1));
  if (s.n == 0) {
    zero(ctx, L.off, sizeof(int64_t));
    L.nbr = (int32_t*)alloc("nbr", 4);
    L.disp = want_disp ? (double*)alloc("disp", 8) : nullptr;
    return L;
  }
  Grid g = make_grid(ctx, s, rcut);
  const double smax = sqrt_le_threshold(rcut);
  Prefilter pf;
  double ext = 0.0;
  for (int d = 0; d < 3; d++) {
    pf.L[d] = g.box.per[d] ? (float)g.box.L[d] : 0.f;
    pf.invL[d] = g.box.per[d] ? (float)(1.0 / g.box.L[d]) : 0.f;
    ext = std::max(ext, g.w[d] * g.nc[d]);
  }
  // FP32 coordinate and arithmetic error bound per component, then a
  // squared-distance bound no true member can exceed
  const double err = 8.0 * ext * std::ldexp(1.0, -23) + 1e-6 * rcut;
  const double rr = rcut + 2.0 * std::sqrt(3.0) * err;
  pf.r2 = (float)(rr * rr * (1.0 + 1e-5));
  int32_t* cellof = scratch_t<int32_t>(ctx, "nl_cellof", s.n);
  int32_t* ccount = scratch_t<int32_t>(ctx, "nl_ccount", 2 * g.ncell + 2);
  int32_t* cfill = ccount + g.ncell + 1;
  int64_t* cstart = scratch_t<int64_t>(ctx, "nl_cstart", g.ncell + 1);
  int32_t* ctmp = scratch_t<int32_t>(ctx, "nl_ctmp", s.n);
  int32_t* catom = scratch_t<int32_t>(ctx, "nl_catom", s.n);
  double4* spos = scratch_t<double4>(ctx, "nl_spos", s.n);
  float4* sposf = scratch_t<float4>(ctx, "nl_sposf", s.n);
  int32_t* cnt = scratch_t<int32_t>(ctx, "nl_cnt", s.n);
  zero(ctx, ccount, sizeof(int32_t) * (2 * g.ncell + 2));
  const unsigned gb = grid_for(s.n, 256);
  launch(ctx, "cell_assign", [&] { k_cell_assign<<<gb, 256, 0, ctx->stream>>>(s.pos, s.n, g, cellof, ccount); });
  exclusive_scan(ctx, ccount, g.ncell, cstart, false);
  launch(ctx, "cell_place", [&] { k_cell_place<<<gb, 256, 0, ctx->stream>>>(s.n, cellof, cstart, cfill, ctmp); });
  launch(ctx, "cell_rank", [&] {
    k_cell_rank<<<gb, 256, 0, ctx->stream>>>(s.n, ctmp, cellof, cstart, s.pos, g, catom, spos, sposf);
  });
  // lanes per atom from the expected candidate count
  int ncellv = 1;
  for (int d = 0; d < 3; d++) ncellv *= g.full[d] ? g.nc[d] : std::min(3, g.nc[d]);
  const double cand = (double)s.n / (double)g.ncell * ncellv;
  const int LP = cand >= 192 ? 16 : (cand >= 48 ? 8 : 4);
  const unsigned gl = grid_for(s.n * LP, 256);
  auto scan = [&](auto lp, auto fill, int64_t* offp, int32_t* nbrp) {
    constexpr int LPc = decltype(lp)::value;
    constexpr bool F = decltype(fill)::value;
    k_list_scan<LPc, F><<<gl, 256, 0, ctx->stream>>>(s.n, g, pf, catom, spos, sposf, cellof, cstart, smax, kind, cnt,
                                                   offp, nbrp, L.npairs);
  };
  auto dispatch = [&](auto fill, int64_t* offp, int32_t* nbrp) {
    if (LP == 16) scan(std::integral_constant<int, 16>(), fill, offp, nbrp);
    else if (LP == 8) scan(std::integral_constant<int, 8>(), fill, offp, nbrp);
    else scan(std::integral_constant<int, 4>(), fill, offp, nbrp);
  };
  launch(ctx, "nlist_count", [&] { dispatch(std::false_type(), nullptr, nullptr); });
  // exact size (reads the total back) or a capacity given by the caller
  // (no host synchronization; the device total is L.off[n], clamped to the
  // capacity, and capflag records an overflow)
  const int64_t np = cap > 0 ? cap : exclusive_scan(ctx, cnt, s.n, L.off, true);
  if (cap > 0) {
    exclusive_scan(ctx, cnt, s.n, L.off, false);
    cap_clamp(ctx, s.n, L.off, cap, capflag, 1);
  }
  L.npairs = np;
  L.exact = cap <= 0;
  L.nbr = (int32_t*)alloc("nbr", sizeof(int32_t) * (np > 0 ? np : 1));
  L.disp = want_disp ? (double*)alloc("disp", sizeof(double) * 3 * (np > 0 ? np : 1)) : nullptr;
  int32_t* fill_to = L.nbr;
  if (sorted && np > 0) fill_to = scratch_t<int32_t>(ctx, "nl_unsorted", np);
  launch(ctx, "nlist_fill", [&] { dispatch(std::true_type(), L.off, fill_to); });
  if (sorted && np > 0)
    launch(ctx, "nlist_sort", [&] {
      k_seg_rank_sort<<<grid_for(np, 256), 256, 0, ctx->stream>>>(s.n, L.off, fill_to, L.nbr);
    });
  if (want_disp && np > 0)
    launch(ctx, "nlist_disp", [&] { k_list_disp<<<gb, 256, 0, ctx->stream>>>(s.n, L.off, L.nbr, s.pos, s.box, L.disp); });
  return L;
}

DevList build_list(ipg_ctx* ctx, const DevSystem& s, double rcut, int kind, bool sorted, bool want_disp,
                   const char* tag, int64_t cap, int32_t* capflag) {
  std::string t(tag);
  return build_list_impl(ctx, s, rcut, kind, sorted, want_disp, cap, capflag,
                         [&](const char* what, size_t bytes) { return scratch(ctx, (t + "_" + what).c_str(), bytes); });
}

DevList build_list_owned(ipg_ctx* ctx, const DevSystem& s, double rcut, int kind, bool sorted, bool want_disp) {
  std::vector<void*> got;
  try {
    return build_list_impl(ctx, s, rcut, kind, sorted, want_disp, 0, nullptr, [&](const char*, size_t bytes) {
      void* p = nullptr;
      if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        fail(IPG_ERR_NOMEM, "hipMalloc failed for a neighbor list");
      }
      got.push_back(p);
      return p;
    });
  } catch (...) {
    for (void* p : got) (void)hipFree(p);
    throw;
  }
}

// ---------------------------------------------------------------------------
// SNAP list, stage B (utl.jl:69-90) over the sorted stage-A inrange list
// (radius `query`, self included, so the enumerate index m is the rank in
// that list).
// ---------------------------------------------------------------------------
struct SnapCutDev {
  int ns;
  int reference;
  double cutsq[64];
};

template <bool FILL>
__global__ void k_snap_stage_b(int64_t n, const int64_t* __restrict__ offA, const int32_t* __restrict__ nbrA,
                               const double* __restrict__ pos, IpgBox box, const int32_t* __restrict__ species,
                               const int32_t* __restrict__ appear, SnapCutDev cut, int32_t* __restrict__ cnt,
                               const int64_t* __restrict__ off, int32_t* __restrict__ nbr, double* __restrict__ R,
                               double* __restrict__ rvec, int64_t cap) {
#pragma clang fp contract(off)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int ne = cut.reference ? appear[i] : (species ? species[i] : 0);
  double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  int64_t a = offA[i], b = offA[i + 1];
  int count = 0;
  int64_t w = FILL ? off[i] : 0;
  for (int64_t q = a; q < b; q++) {
    int32_t j = nbrA[q];
    int64_t m = q - a + 1;  // 1-based enumerate index (utl.jl:75)
    int me = cut.reference ? appear[m - 1] : (species ? species[j] : 0);
    double rx = -1.0 * ipg_disp1(xi, pos[3 * j], box.L[0], box.per[0]);
    double ry = -1.0 * ipg_disp1(yi, pos[3 * j + 1], box.L[1], box.per[1]);
    double rz = -1.0 * ipg_disp1(zi, pos[3 * j + 2], box.L[2], box.per[2]);
    double rsq = rx * rx + ry * ry + rz * rz;
    if (j != i && rsq <= cut.cutsq[ne * cut.ns + me]) {
      if (FILL) {
        if (w >= cap) break;  // capacity mode overflow (flagged by the clamp)
        nbr[w] = j;
        R[w] = sqrt(rsq);
        rvec[3 * w] = rx;
        rvec[3 * w + 1] = ry;
        rvec[3 * w + 2] = rz;
        w++;
      } else {
        count++;
      }
    }
  }
  if (!FILL) cnt[i] = count;
}

__global__ void k_pair_center(int64_t n, const int64_t* __restrict__ off, int32_t* __restrict__ center) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int64_t q = off[i]; q < off[i + 1]; q++) center[q] = (int32_t)i;
}

// The pair count is read on the device (*np = off[n] of the list), so a
// capacity-sized grid needs no host readback.
__global__ void k_rev_count(const int64_t* __restrict__ np, const int32_t* __restrict__ nbr,
                            int32_t* __restrict__ rcnt) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= *np) return;
  atomicAdd(&rcnt[nbr[q]], 1);
}

__global__ void k_rev_place(const int64_t* __restrict__ np, const int32_t* __restrict__ nbr,
                            const int64_t* __restrict__ roff, int32_t* __restrict__ rfill,
                            int32_t* __restrict__ rtmp) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= *np) return;
  int32_t j = nbr[q];
  int32_t slot = atomicAdd(&rfill[j], 1);
  rtmp[roff[j] + slot] = (int32_t)q;
}

void build_center_and_reverse(ipg_ctx* ctx, int64_t n, int64_t npairs, const int64_t* off, const int32_t* nbr,
                              int32_t* center, int64_t* roff, int32_t* rpair, const char* tag) {
  std::string t(tag);
  unsigned gn = grid_for(n > 0 ? n : 1, 256), gp = grid_for(npairs > 0 ? npairs : 1, 256);
  if (center) launch(ctx, "pair_center", [&] { k_pair_center<<<gn, 256, 0, ctx->stream>>>(n, off, center); });
  if (!roff) return;
  int32_t* rc = scratch_t<int32_t>(ctx, (t + "_rcnt").c_str(), 2 * n + 2);
  int32_t* rf = rc + n + 1;
  zero(ctx, rc, sizeof(int32_t) * (2 * n + 2));
  const int64_t* np = off + n;  // device pair count (<= npairs, which may be a capacity)
  if (npairs > 0) launch(ctx, "rev_count", [&] { k_rev_count<<<gp, 256, 0, ctx->stream>>>(np, nbr, rc); });
  exclusive_scan(ctx, rc, n, roff, false);
  if (npairs > 0) {
    int32_t* rtmp = scratch_t<int32_t>(ctx, (t + "_rtmp").c_str(), npairs);
    launch(ctx, "rev_place", [&] { k_rev_place<<<gp, 256, 0, ctx->stream>>>(np, nbr, roff, rf, rtmp); });
    launch(ctx, "rev_sort", [&] { k_seg_rank_sort<<<gp, 256, 0, ctx->stream>>>(n, roff, rtmp, rpair); });
  }
}

SnapListDev build_snap_list(ipg_ctx* ctx, const DevSystem& s, const int32_t* d_appear, const SnapCut& cut,
                            bool want_reverse, const char* tag, int64_t capA, int64_t capB, int32_t* capflag) {
  std::string t(tag);
  SnapListDev out;
  DevList A = build_list(ctx, s, cut.query, KIND_FULL_SELF, true, false, (t + "_A").c_str(), capA, capflag);
  out.npairs_a = A.npairs;
  SnapCutDev cd{};
  cd.ns = cut.ns;
  cd.reference = cut.reference;
  for (int k = 0; k < cut.ns * cut.ns; k++) cd.cutsq[k] = cut.cutsq[k];
  int64_t n = s.n;
  DevList& L = out.list;
  L.n = n;
  L.off = scratch_t<int64_t>(ctx, (t + "_off").c_str(), n + 1);
  int32_t* cnt = scratch_t<int32_t>(ctx, (t + "_cnt").c_str(), n > 0 ? n : 1);
  unsigned gn = grid_for(n > 0 ? n : 1, 256);
  if (n > 0)
    launch(ctx, "snap_list_count", [&] {
      k_snap_stage_b<false><<<gn, 256, 0, ctx->stream>>>(n, A.off, A.nbr, s.pos, s.box, s.species, d_appear, cd,
                                                         cnt, nullptr, nullptr, nullptr, nullptr, 0);
    });
  int64_t np;
  if (capB > 0) {
    np = capB;
    exclusive_scan(ctx, cnt, n, L.off, false);
    cap_clamp(ctx, n, L.off, capB, capflag, 2);
    L.exact = false;
  } else {
    np = exclusive_scan(ctx, cnt, n, L.off, true);
  }
  L.npairs = np;
  int64_t npa = np > 0 ? np : 1;
  L.nbr = scratch_t<int32_t>(ctx, (t + "_nbr").c_str(), npa);
  out.R = scratch_t<double>(ctx, (t + "_R").c_str(), npa);
  out.rvec = scratch_t<double>(ctx, (t + "_rvec").c_str(), 3 * npa);
  if (n > 0 && np > 0)
    launch(ctx, "snap_list_fill", [&] {
      k_snap_stage_b<true><<<gn, 256, 0, ctx->stream>>>(n, A.off, A.nbr, s.pos, s.box, s.species, d_appear, cd,
                                                        nullptr, L.off, L.nbr, out.R, out.rvec, np);
    });
  if (want_reverse) {
    out.center = scratch_t<int32_t>(ctx, (t + "_center").c_str(), npa);
    out.roff = scratch_t<int64_t>(ctx, (t + "_roff").c_str(), n + 1);
    out.rpair = scratch_t<int32_t>(ctx, (t + "_rpair").c_str(), npa);
    build_center_and_reverse(ctx, n, np, L.off, L.nbr, out.center, out.roff, out.rpair, t.c_str());
  }
  return out;
}

}  // namespace ipg

using namespace ipg;

extern "C" {

int ipg_neighborlist(ipg_ctx* ctx, const ipg_system* sys, double rcut, int kind, int64_t* offsets, int32_t* nbr,
                     double* disp, int64_t cap, int64_t* npairs) {
  return guard([&] {
    IPG_REQUIRE(ctx && sys && offsets && npairs, "NULL argument");
    IPG_REQUIRE(kind == IPG_LIST_HALF || kind == IPG_LIST_FULL, "kind must be IPG_LIST_HALF or IPG_LIST_FULL");
    IPG_HIP(hipSetDevice(ctx->device));
    DevSystem s = upload_system(ctx, sys, "nl_api");
    DevList L = build_list(ctx, s, rcut, kind == IPG_LIST_HALF ? KIND_HALF : KIND_FULL, true, disp != nullptr,
                           "nl_api");
    d2h(ctx, offsets, L.off, sizeof(int64_t) * (s.n + 1));
    *npairs = L.npairs;
    if (L.npairs > cap) {
      sync(ctx);
      fail(IPG_ERR_CAPACITY, "capacity " + std::to_string(cap) + " < pairs " + std::to_string(L.npairs));
    }
    IPG_REQUIRE(L.npairs == 0 || nbr, "nbr is NULL");
    d2h(ctx, nbr, L.nbr, sizeof(int32_t) * L.npairs);
    if (disp) d2h(ctx, disp, L.disp, sizeof(double) * 3 * L.npairs);
    sync(ctx);
  });
}

int ipg_nlist_build_dev(ipg_ctx* ctx, const ipg_system* sys, double rcut, int kind, ipg_nlist** out) {
  return guard([&] {
    IPG_REQUIRE(ctx && sys && out, "NULL argument");
    IPG_REQUIRE(kind == IPG_LIST_HALF || kind == IPG_LIST_FULL, "kind must be IPG_LIST_HALF or IPG_LIST_FULL");
    IPG_HIP(hipSetDevice(ctx->device));
    DevSystem s = make_dev_system(sys);
    auto* h = new ipg_nlist();
    try {
      h->list = build_list_owned(ctx, s, rcut, kind == IPG_LIST_HALF ? KIND_HALF : KIND_FULL, false, false);
    } catch (...) {
      delete h;
      throw;
    }
    h->ctx = ctx;
    h->rcut = rcut;
    h->kind = kind;
    *out = h;
  });
}

int ipg_nlist_info(const ipg_nlist* list, int64_t* natoms, int64_t* npairs, const int64_t** d_offsets,
                   const int32_t** d_nbr) {
  return guard([&] {
    IPG_REQUIRE(list, "list is NULL");
    if (natoms) *natoms = list->list.n;
    if (npairs) *npairs = list->list.npairs;
    if (d_offsets) *d_offsets = list->list.off;
    if (d_nbr) *d_nbr = list->list.nbr;
  });
}

int ipg_nlist_destroy(ipg_nlist* list) {
  return guard([&] {
    if (!list) return;
    if (list->ctx) (void)hipStreamSynchronize(list->ctx->stream);
    if (list->list.off) (void)hipFree(list->list.off);
    if (list->list.nbr) (void)hipFree(list->list.nbr);
    if (list->list.disp) (void)hipFree(list->list.disp);
    delete list;
  });
}

}  // extern "C"
