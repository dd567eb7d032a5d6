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
  int full[3];  // periodic dimension with fewer than 2 hw + 1 cells: visit all of them once
  int hw;       // cells are at least rcut / hw wide; a cell's candidates lie within hw cells of it
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
                            const int64_t* __restrict__ cstart, const double* __restrict__ pos,
                            const int32_t* __restrict__ species, Grid g, int32_t* __restrict__ catom,
                            double4* __restrict__ spos, float4* __restrict__ sposf) {
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
  spos[dst] = make_double4(x, y, z, species ? (double)species[i] : 0.0);  // w: species index
  double w[3] = {x, y, z};
  for (int d = 0; d < 3; d++) w[d] = g.box.per[d] ? w[d] - floor(w[d] / g.box.L[d]) * g.box.L[d] : w[d] - g.lo[d];
  sposf[dst] = make_float4((float)w[0], (float)w[1], (float)w[2], 0.f);
}

// The same binning in two launches, for small systems whose counts and slot
// list fit in LDS: a thread per atom computes its cell and its double and FP32
// coordinates (and clears the caller's flag word); then one block counts
// the cells with LDS atomics, scans the counts, claims slots and ranks
// every entry by atom index within its cell.  The outputs (cellof, cstart,
// catom, spos, sposf) equal those of the four kernels above; a captured
// step then waits on the host for two short launches instead of five.
constexpr int CELLS_BLOCK = 1024;
__global__ void k_cell_ids(int64_t n, const double* __restrict__ pos, const int32_t* __restrict__ species, Grid g,
                           int32_t* __restrict__ cellof, double4* __restrict__ pw, float4* __restrict__ pf,
                           int32_t* __restrict__ clear) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (clear && i == 0) *clear = 0;
  if (i >= n) return;
  const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
  const int c0 = cell_coord(g, 0, x), c1 = cell_coord(g, 1, y), c2 = cell_coord(g, 2, z);
  cellof[i] = (int32_t)(((int64_t)c2 * g.nc[1] + c1) * g.nc[0] + c0);
  if (!pw) return;  // cells only
  pw[i] = make_double4(x, y, z, species ? (double)species[i] : 0.0);  // w: species index
  double w[3] = {x, y, z};
  for (int d = 0; d < 3; d++) w[d] = g.box.per[d] ? w[d] - floor(w[d] / g.box.L[d]) * g.box.L[d] : w[d] - g.lo[d];
  pf[i] = make_float4((float)w[0], (float)w[1], (float)w[2], 0.f);
}

// Mid-size systems: the block only counts, scans and claims slots (ctmp, in
// global memory); the parallel k_cell_rank then orders and gathers.
__global__ __launch_bounds__(CELLS_BLOCK) void k_cells_slots(int32_t n, int32_t nc, const int32_t* __restrict__ cellof,
                                                             int64_t* __restrict__ cstart, int32_t* __restrict__ ctmp) {
  extern __shared__ int32_t lds[];
  __shared__ int32_t wsum[CELLS_BLOCK / 64];
  int32_t* cur = lds;
  const int tid = threadIdx.x;
  for (int c = tid; c < nc; c += CELLS_BLOCK) cur[c] = 0;
  __syncthreads();
  for (int32_t i = tid; i < n; i += CELLS_BLOCK) atomicAdd(&cur[cellof[i]], 1);
  __syncthreads();
  const int per = (nc + CELLS_BLOCK - 1) / CELLS_BLOCK;
  const int c0 = min(tid * per, nc), c1 = min(c0 + per, nc);
  int32_t own = 0;
  for (int c = c0; c < c1; c++) own += cur[c];
  const int lane = tid & 63, wv = tid >> 6;
  int32_t x = own;
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  int32_t run = x - own;
  for (int w = 0; w < wv; w++) run += wsum[w];
  for (int c = c0; c < c1; c++) {
    const int32_t k = cur[c];
    cstart[c] = run;
    cur[c] = run;
    run += k;
  }
  if (tid == 0) cstart[nc] = n;
  __syncthreads();
  for (int32_t i = tid; i < n; i += CELLS_BLOCK) ctmp[atomicAdd(&cur[cellof[i]], 1)] = i;
}

__global__ __launch_bounds__(CELLS_BLOCK) void k_cells_order(int32_t n, int32_t nc, const int32_t* __restrict__ cellof,
                                                             const double4* __restrict__ pw,
                                                             const float4* __restrict__ pf,
                                                             int64_t* __restrict__ cstart, int32_t* __restrict__ catom,
                                                             double4* __restrict__ spos, float4* __restrict__ sposf) {
  extern __shared__ int32_t lds[];
  __shared__ int32_t wsum[CELLS_BLOCK / 64];
  int32_t* cur = lds;            // counts, then the end of every cell's slots
  int32_t* slot = lds + nc + 1;  // atoms in slot order
  const int tid = threadIdx.x;
  for (int c = tid; c < nc; c += CELLS_BLOCK) cur[c] = 0;
  __syncthreads();
  for (int32_t i = tid; i < n; i += CELLS_BLOCK) atomicAdd(&cur[cellof[i]], 1);
  __syncthreads();
  // exclusive scan of the counts, each thread a contiguous run of cells
  const int per = (nc + CELLS_BLOCK - 1) / CELLS_BLOCK;
  const int c0 = min(tid * per, nc), c1 = min(c0 + per, nc);
  int32_t own = 0;
  for (int c = c0; c < c1; c++) own += cur[c];
  const int lane = tid & 63, wv = tid >> 6;
  int32_t x = own;
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  int32_t run = x - own;
  for (int w = 0; w < wv; w++) run += wsum[w];
  for (int c = c0; c < c1; c++) {
    const int32_t k = cur[c];
    cstart[c] = run;
    cur[c] = run;  // fill cursor: the cell's first slot
    run += k;
  }
  if (tid == 0) cstart[nc] = n;
  __syncthreads();
  for (int32_t i = tid; i < n; i += CELLS_BLOCK) slot[atomicAdd(&cur[cellof[i]], 1)] = i;
  __syncthreads();
  // cur[c] is now the end of cell c and the start of cell c + 1
  for (int32_t k = tid; k < n; k += CELLS_BLOCK) {
    const int32_t i = slot[k];
    const int32_t c = cellof[i];
    const int32_t a = c > 0 ? cur[c - 1] : 0, b = cur[c];
    int32_t r = 0;
    for (int32_t m = a; m < b; m++) r += slot[m] < i;
    catom[a + r] = i;
    spos[a + r] = pw[i];
    sposf[a + r] = pf[i];
  }
}

// Candidate atoms of a cell: the cells within hw of it in each dimension
// (all cells of a periodic dimension with fewer than 2 hw + 1 cells,
// clipped at open boundaries), each visited once.  Cells are numbered x
// fastest and atoms are sorted by cell, so the cells x-hw .. x+hw of one
// (y, z) row hold one contiguous range of atoms (two ranges where a
// periodic row wraps).  Range r (0 .. 2 (2 hw + 1)^2 - 1) is run r%2 of row
// r/2; empty ranges come back as s0 == s1.
__device__ __forceinline__ void cell_range(const Grid& g, int32_t cell, const int64_t* __restrict__ cstart, int r,
                                           int32_t& s0, int32_t& s1) {
  s0 = s1 = 0;
  const int nx = g.nc[0], hw = g.hw, span = 2 * g.hw + 1;
  const int cx = cell % nx, cy = (cell / nx) % g.nc[1], cz = cell / (nx * g.nc[1]);
  const int nyr = g.full[1] ? g.nc[1] : span, nzr = g.full[2] ? g.nc[2] : span;
  const int row = r >> 1, run = r & 1;
  if (row >= nyr * nzr) return;
  int yy = row % nyr, zz = row / nyr;
  if (!g.full[1]) {
    yy = cy + yy - hw;
    if (g.box.per[1]) yy = (yy + g.nc[1]) % g.nc[1];
    else if (yy < 0 || yy >= g.nc[1]) return;
  }
  if (!g.full[2]) {
    zz = cz + zz - hw;
    if (g.box.per[2]) zz = (zz + g.nc[2]) % g.nc[2];
    else if (zz < 0 || zz >= g.nc[2]) return;
  }
  const int64_t rb = ((int64_t)zz * g.nc[1] + yy) * nx;
  int64_t a = 0, b = 0;
  if (g.full[0]) {
    if (run == 0) a = rb, b = rb + nx;
  } else if (g.box.per[0]) {  // nx >= 2 hw + 1 here
    if (cx - hw < 0) {
      if (run == 0) a = rb + nx + cx - hw, b = rb + nx;
      else a = rb, b = rb + cx + hw + 1;
    } else if (cx + hw >= nx) {
      if (run == 0) a = rb + cx - hw, b = rb + nx;
      else a = rb, b = rb + cx + hw + 1 - nx;
    } else if (run == 0) {
      a = rb + cx - hw, b = rb + cx + hw + 1;
    }
  } else if (run == 0) {
    a = rb + (cx - hw > 0 ? cx - hw : 0);
    b = rb + (cx + hw < nx ? cx + hw : nx - 1) + 1;
  }
  if (b > a) {
    s0 = (int32_t)cstart[a];
    s1 = (int32_t)cstart[b];
  }
}

__device__ __forceinline__ bool keep_pair(int kind, int32_t i, int32_t j) {
  if (kind == KIND_HALF) return j > i;
  if (kind == KIND_FULL) return j != i;
  return true;  // KIND_FULL_SELF
}

// Species-pair cutoffs of the SNAP list (utl.jl:42-56).  reference = 1
// selects the element lookup of utl.jl:70,86 (two-stage builder only).
struct SnapCutDev {
  int ns;
  int reference;
  double cutsq[64];
};

// Second membership test of the SNAP list (utl.jl:84-88): r = x_j - x_i by
// get_displacement, r.r (left to right, unfused) against the cutoff of the
// pair's species (w components of the cell-ordered coordinates).
__device__ __forceinline__ bool snap_cut_ok(const IpgBox& b, const double4& pi, const double4& pj,
                                            const SnapCutDev& sc) {
#pragma clang fp contract(off)
  const double rx = -1.0 * ipg_disp1(pi.x, pj.x, b.L[0], b.per[0]);
  const double ry = -1.0 * ipg_disp1(pi.y, pj.y, b.L[1], b.per[1]);
  const double rz = -1.0 * ipg_disp1(pi.z, pj.z, b.L[2], b.per[2]);
  const double rsq = rx * rx + ry * ry + rz * rz;
  // species indices come from the w components; one outside [0, ns) reads
  // species 0 (the SNAP kernels report it)
  const int si = (int)pi.w, sj = (int)pj.w;
  return rsq <= sc.cutsq[((unsigned)si < (unsigned)sc.ns ? si : 0) * sc.ns + ((unsigned)sj < (unsigned)sc.ns ? sj : 0)];
}

struct Prefilter {
  float L[3], invL[3];
  float r2;  // FP32 squared-distance bound that no member can exceed
};

// LP lanes per atom scan the surrounding cells; each cell's candidates are
// taken LP at a time.  A cheap FP32 minimum-image distance rejects far
// candidates; the survivors are queued in LDS in candidate order, and the
// exact FP64 PeriodicEuclidean test runs over the queue LP entries at a time
// (so the FP64 code runs on full lanes, not once per candidate chunk).  Hits
// are placed with a ballot, so the output order is fixed (cell order, then
// slot).
// MODE: SCAN_COUNT (cnt[i]), SCAN_FILL (rows at off[i], writes stop at cap)
// or SCAN_ELL (the first K hits of atom i in ell[i*K ...] and their count
// in cnt[i]; more than K hits raise bit 8 of *flag).
enum { SCAN_COUNT = 0, SCAN_FILL = 1, SCAN_ELL = 2 };
constexpr int SCAN_QCAP = 4;  // queue capacity per group, in chunks of LP
constexpr int64_t HALF_CELL_MIN_N = 32768;  // atoms from which lists use half-cutoff cells
template <int LP, int MODE, bool SC, int HW>
__global__ __launch_bounds__(256) void k_list_scan(int64_t n, Grid g, Prefilter pf, const int32_t* __restrict__ catom,
                                                   const double4* __restrict__ spos, const float4* __restrict__ sposf,
                                                   const int32_t* __restrict__ cellof,
                                                   const int64_t* __restrict__ cstart, double smax, int kind,
                                                   int32_t* __restrict__ cnt, const int64_t* __restrict__ off,
                                                   int32_t* __restrict__ nbr, int64_t cap, SnapCutDev sc,
                                                   int ellK, int32_t* __restrict__ flag) {
  constexpr bool FILL = MODE == SCAN_FILL;
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t k = gt / LP;
  const int l = (int)(gt % LP);
  const int lane = threadIdx.x & 63;
  if (k >= n) return;  // whole groups leave together
  const uint64_t gmask = (LP == 64 ? ~0ull : ((1ull << (LP % 64)) - 1ull)) << (lane - l);
  const uint64_t below = (1ull << lane) - 1ull;
  const int32_t i = catom[k];
  const double4 pi = spos[k];
  const float4 qi = sposf[k];
  int64_t w = FILL ? off[i] : 0;
  int count = 0;
  // the 2 (2 HW + 1)^2 candidate ranges, computed up front by the group's
  // lanes (all cstart loads in flight together)
  constexpr int NR = 2 * (2 * HW + 1) * (2 * HW + 1), NS = (NR + LP - 1) / LP;
  const int32_t cell = cellof[i];
  int32_t rs[NS], re[NS];
#pragma unroll
  for (int t = 0; t < NS; t++) cell_range(g, cell, cstart, l + t * LP, rs[t], re[t]);
  const int base = lane - l;
  // survivors of the prefilter (cell-order slots), one queue per group
  __shared__ int32_t queue[256 / LP][SCAN_QCAP * LP];
  int32_t* qg = queue[threadIdx.x / LP];
  int qn = 0;  // the same in every lane of the group
  auto drain = [&] {
    __builtin_amdgcn_wave_barrier();  // queue stores before the loads below
    for (int t = 0; t < qn; t += LP) {
      bool hit = false;
      int32_t j = -1;
      if (t + l < qn) {
        const int32_t sl = qg[t + l];
        j = catom[sl];
        const double4 pj = spos[sl];
        hit = ipg_pe_sumsq(g.box, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z) <= smax;
        if constexpr (SC) hit = hit && snap_cut_ok(g.box, pi, pj, sc);
      }
      const uint64_t m = __ballot(hit) & gmask;
      if constexpr (MODE == SCAN_FILL) {
        if (hit && w + __popcll(m & below) < cap) nbr[w + __popcll(m & below)] = j;
        w += __popcll(m);
      } else {
        if constexpr (MODE == SCAN_ELL) {
          const int r = count + __popcll(m & below);
          if (hit && r < ellK) nbr[(int64_t)i * ellK + r] = j;
        }
        count += __popcll(m);
      }
    }
    __builtin_amdgcn_wave_barrier();  // loads done before the queue is refilled
    qn = 0;
  };
  if constexpr (HW == 2) {
    // half-cutoff cells: 50 short ranges, walked as one concatenated list
    // so that every chunk of LP candidates is full; per range its first
    // slot and its offset in the list (prefix of the lengths in range order)
    __shared__ int32_t rfirst[256 / LP][NR], roff[256 / LP][NR + 1];
    const int grp = threadIdx.x / LP;
    int basec = 0;
#pragma unroll
    for (int t = 0; t < NS; t++) {
      const int r = l + t * LP;
      const int len = r < NR ? re[t] - rs[t] : 0;
      int incl = len;
      for (int o = 1; o < LP; o <<= 1) {
        const int y = __shfl_up(incl, o, LP);
        if (l >= o) incl += y;
      }
      if (r < NR) {
        rfirst[grp][r] = rs[t];
        roff[grp][r] = basec + incl - len;
      }
      basec += __shfl(incl, LP - 1, LP);
    }
    if (l == 0) roff[grp][NR] = basec;
    __builtin_amdgcn_wave_barrier();  // table stores before the loads below
    const int total = basec;
    int r = 0;
    for (int c0 = 0; c0 < total; c0 += LP) {
      const int c = c0 + l;
      bool pass = false;
      int32_t sl = 0;
      if (c < total) {
        while (roff[grp][r + 1] <= c) r++;
        sl = rfirst[grp][r] + (c - roff[grp][r]);
        const int32_t j = catom[sl];
        const float4 qj = sposf[sl];
        if (keep_pair(kind, i, j)) {
          float dx = qj.x - qi.x, dy = qj.y - qi.y, dz = qj.z - qi.z;
          if (g.box.per[0]) dx -= pf.L[0] * rintf(dx * pf.invL[0]);
          if (g.box.per[1]) dy -= pf.L[1] * rintf(dy * pf.invL[1]);
          if (g.box.per[2]) dz -= pf.L[2] * rintf(dz * pf.invL[2]);
          pass = dx * dx + dy * dy + dz * dz <= pf.r2;
        }
      }
      const uint64_t m = __ballot(pass) & gmask;
      if (pass) qg[qn + __popcll(m & below)] = sl;
      qn += __popcll(m);
      if (qn > (SCAN_QCAP - 1) * LP) drain();  // room for one more chunk
    }
  } else
  for (int r = 0; r < NR; r++) {
    const int t = r / LP;
    int32_t a0 = rs[0], a1 = re[0];
#pragma unroll
    for (int u = 1; u < NS; u++)
      if (t == u) a0 = rs[u], a1 = re[u];
    const int32_t s0 = __shfl(a0, base + r % LP, 64), s1 = __shfl(a1, base + r % LP, 64);
    for (int64_t s = s0; s < s1; s += LP) {
      const int64_t sl = s + l;
      bool pass = false;
      if (sl < s1) {
        const int32_t j = catom[sl];
        const float4 qj = sposf[sl];
        if (keep_pair(kind, i, j)) {
          float dx = qj.x - qi.x, dy = qj.y - qi.y, dz = qj.z - qi.z;
          if (g.box.per[0]) dx -= pf.L[0] * rintf(dx * pf.invL[0]);
          if (g.box.per[1]) dy -= pf.L[1] * rintf(dy * pf.invL[1]);
          if (g.box.per[2]) dz -= pf.L[2] * rintf(dz * pf.invL[2]);
          pass = dx * dx + dy * dy + dz * dz <= pf.r2;
        }
      }
      const uint64_t m = __ballot(pass) & gmask;
      if (pass) qg[qn + __popcll(m & below)] = (int32_t)sl;
      qn += __popcll(m);
      if (qn > (SCAN_QCAP - 1) * LP) drain();  // room for one more chunk
    }
  }
  if (qn > 0) drain();
  if (MODE != SCAN_FILL && l == 0) {
    if (MODE == SCAN_ELL && count > ellK) {
      atomicOr(flag, 8);
      count = ellK;  // the row holds what was written, so every list slot gets an entry
    }
    cnt[i] = count;
  }
}

// Cell-pair rows for small periodic SNAP lists (every dimension periodic
// with 4..1023 cells: the 27 cells around a cell are distinct and lie
// within 1.5 cells of its center, less than half a box length).  No sort by
// cell: k_cell_pack writes each atom's cell (16 bits), its coordinates and
// species (exact) and its FP32 wrapped coordinates.  Then one block per
// home cell reads every atom's cell (coalesced; thread t holds atoms
// 2048 u + 8 t + e), tests it against a bitmap of the 27 cells around the
// home cell, and compacts the candidates in atom order (a block prefix sum
// of the per-u counts, packed four to a 64-bit word).  The candidates'
// coordinates are staged in LDS: exact, and FP32 moved to the periodic
// image nearest the home cell's center, so that the prefilter needs no
// wrapping.  Half-waves take the home cell's atoms: the FP32 prefilter
// (the bound of k_list_scan) over every candidate, then the exact tests
// (PeriodicEuclidean and the species-pair cutoff) on the survivors, whose
// order is kept, so each row comes out sorted by j and goes straight to
// ell.  cnt[i] = row length (clamped to K, with bit 8 of *flag on
// overflow); a cell holding more than CR_MAXH atoms, a neighborhood of
// more than CR_MAXC, or more than CR_KMAX prefilter survivors also set
// bit 8, and the caller falls back to the binned scan.
constexpr int CR_MAXC = 640, CR_MAXH = 64, CR_KMAX = 128, CR_UMAX = 8, CR_MAXCELL = 16384;
constexpr int CR_MAXN = CR_UMAX * 8 * 256;  // atoms: 8 cells per 16-byte load, CR_UMAX loads per thread
__global__ void k_cell_pack(int64_t n, const double* __restrict__ pos, const int32_t* __restrict__ species, Grid g,
                            uint16_t* __restrict__ cellc, double4* __restrict__ pw, float4* __restrict__ pf,
                            int32_t* __restrict__ clear) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (clear && i == 0) *clear = 0;
  if (i >= n) return;
  const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
  cellc[i] = (uint16_t)(((int64_t)cell_coord(g, 2, z) * g.nc[1] + cell_coord(g, 1, y)) * g.nc[0] + cell_coord(g, 0, x));
  const double sp = species ? (double)species[i] : 0.0;
  pw[i] = make_double4(x, y, z, sp);
  double w[3] = {x, y, z};
  for (int d = 0; d < 3; d++) w[d] = w[d] - floor(w[d] / g.box.L[d]) * g.box.L[d];
  pf[i] = make_float4((float)w[0], (float)w[1], (float)w[2], (float)sp);  // w: species index (exact)
}

__device__ __forceinline__ uint32_t cr_field(uint64_t lo, uint64_t hi, int u) {  // 16-bit field u of (lo, hi)
  return (uint32_t)(((u < 4 ? lo : hi) >> (16 * (u & 3))) & 0xffffu);
}

__global__ __launch_bounds__(256) void k_snapc_cell_rows(int32_t n, Grid g, Prefilter pfl,
                                                         const int4* __restrict__ cellc8,
                                                         const double4* __restrict__ pw,
                                                         const float4* __restrict__ pf, double smax, SnapCutDev sc,
                                                         int K, int32_t* __restrict__ ell, int32_t* __restrict__ cnt,
                                                         int32_t* __restrict__ flag) {
  __shared__ int32_t cid[CR_MAXC];
  __shared__ float4 cposf[CR_MAXC];  // FP32 image nearest the home cell's center; w: species
  __shared__ double cx[CR_MAXC], cy[CR_MAXC], cz[CR_MAXC];
  __shared__ uint32_t bitmap[CR_MAXCELL / 32];
  __shared__ int32_t hid[CR_MAXH];  // LDS slots of the home cell's atoms
  __shared__ int32_t rows[8][CR_KMAX];
  __shared__ uint64_t wtot[2][4];
  __shared__ int32_t nhome;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int h = blockIdx.x;
  const int hx = h % g.nc[0], hy = (h / g.nc[0]) % g.nc[1], hz = h / (g.nc[0] * g.nc[1]);
  const int nq = (n + 7) >> 3;  // 16-byte loads of 8 cells
  int4 v[CR_UMAX];
#pragma unroll
  for (int u = 0; u < CR_UMAX; u++)
    if (t + 256 * u < nq) v[u] = cellc8[t + 256 * u];
  const int nwords = (int)((g.ncell + 31) >> 5);
  for (int w = t; w < nwords; w += 256) bitmap[w] = 0u;
  if (t == 0) nhome = 0;
  __syncthreads();
  if (t < 27) {
    int x = hx + t % 3 - 1, y = hy + (t / 3) % 3 - 1, z = hz + t / 9 - 1;
    x += x < 0 ? g.nc[0] : (x >= g.nc[0] ? -g.nc[0] : 0);
    y += y < 0 ? g.nc[1] : (y >= g.nc[1] ? -g.nc[1] : 0);
    z += z < 0 ? g.nc[2] : (z >= g.nc[2] ? -g.nc[2] : 0);
    const int c = (z * g.nc[1] + y) * g.nc[0] + x;
    atomicOr(&bitmap[c >> 5], 1u << (c & 31));
  }
  __syncthreads();  // bitmap complete
  // bit 8u+e of the masks: atom 2048 u + 8 t + e
  uint64_t cm = 0, hm = 0;
#pragma unroll
  for (int u = 0; u < CR_UMAX; u++) {
    if (t + 256 * u < nq) {
      const uint32_t vv[4] = {(uint32_t)v[u].x, (uint32_t)v[u].y, (uint32_t)v[u].z, (uint32_t)v[u].w};
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const bool in = 8 * (t + 256 * u) + e < n;
        const int c = in ? (int)((vv[e >> 1] >> (16 * (e & 1))) & 0xffffu) : 0;  // padding after atom n - 1
        const uint64_t bit = 1ull << (8 * u + e);
        if (in && ((bitmap[c >> 5] >> (c & 31)) & 1u)) cm |= bit;
        if (in && c == h) hm |= bit;
      }
    }
  }
  // per-u candidate counts, four 16-bit fields per word; block prefix sums
  uint64_t p0 = 0, p1 = 0;
#pragma unroll
  for (int u = 0; u < CR_UMAX; u++) {
    const uint64_t c = (uint64_t)__popcll((cm >> (8 * u)) & 0xffull);
    if (u < 4) p0 |= c << (16 * u);
    else p1 |= c << (16 * (u - 4));
  }
  uint64_t i0 = p0, i1 = p1;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y0 = __shfl_up(i0, o, 64), y1 = __shfl_up(i1, o, 64);
    if (lane >= o) {
      i0 += y0;
      i1 += y1;
    }
  }
  if (lane == 63) {
    wtot[0][wv] = i0;
    wtot[1][wv] = i1;
  }
  __syncthreads();
  uint64_t e0 = i0 - p0, e1 = i1 - p1, t0 = 0, t1 = 0;
  for (int w = 0; w < 4; w++) {
    if (w < wv) {
      e0 += wtot[0][w];
      e1 += wtot[1][w];
    }
    t0 += wtot[0][w];
    t1 += wtot[1][w];
  }
  int nc = 0;
#pragma unroll
  for (int u = 0; u < CR_UMAX; u++) nc += (int)cr_field(t0, t1, u);
  // a cell that overflows the LDS lists leaves its atoms with empty rows
  // (and bit 8 set), so no later kernel of a captured step reads a stale row
  auto empty_rows = [&] {
#pragma unroll
    for (int u = 0; u < CR_UMAX; u++) {
      uint32_t mh = (uint32_t)((hm >> (8 * u)) & 0xffull);
      while (mh) {
        const int e = __builtin_ctz(mh);
        cnt[8 * (t + 256 * u) + e] = 0;
        mh &= mh - 1;
      }
    }
  };
  if (nc > CR_MAXC) {
    empty_rows();
    if (t == 0) atomicOr(flag, 8);
    return;  // uniform: nc is the block total
  }
  int ou = 0;  // candidates of the slices before u
#pragma unroll
  for (int u = 0; u < CR_UMAX; u++) {
    uint32_t m = (uint32_t)((cm >> (8 * u)) & 0xffull);
    const uint32_t mh = (uint32_t)((hm >> (8 * u)) & 0xffull);
    int base = ou + (int)cr_field(e0, e1, u);
    while (m) {
      const int e = __builtin_ctz(m);
      cid[base] = 8 * (t + 256 * u) + e;
      if ((mh >> e) & 1u) {
        const int sl = atomicAdd(&nhome, 1);
        if (sl < CR_MAXH) hid[sl] = base;
      }
      base++;
      m &= m - 1;
    }
    ou += (int)cr_field(t0, t1, u);
  }
  __syncthreads();
  {
    // FP32 coordinates moved to the image nearest the home cell's center
    const float hc[3] = {(float)((hx + 0.5) * g.w[0]), (float)((hy + 0.5) * g.w[1]), (float)((hz + 0.5) * g.w[2])};
    for (int k = t; k < nc; k += 256) {  // one round of loads for the whole neighborhood
      const int32_t j = cid[k];
      const double4 p = pw[j];
      float4 q = pf[j];
      q.x += pfl.L[0] * rintf((hc[0] - q.x) * pfl.invL[0]);
      q.y += pfl.L[1] * rintf((hc[1] - q.y) * pfl.invL[1]);
      q.z += pfl.L[2] * rintf((hc[2] - q.z) * pfl.invL[2]);
      cposf[k] = q;
      cx[k] = p.x;
      cy[k] = p.y;
      cz[k] = p.z;
    }
  }
  __syncthreads();
  const int nh = nhome;
  if (nh > CR_MAXH) {
    empty_rows();
    if (t == 0) atomicOr(flag, 8);
    return;
  }
  // half-wave hw (0, 1) of wave wv takes home atoms 2 (wv + 4 it) + hw
  const int hw = lane >> 5, hl = lane & 31;
  const uint32_t hbelow = (1u << hl) - 1u;
  int32_t* row = rows[2 * wv + hw];
  for (int p = wv; 2 * p < nh; p += 4) {
    const bool act = 2 * p + hw < nh;
    const int si = act ? hid[2 * p + hw] : 0;
    const int32_t i = cid[si];
    const float4 qi = cposf[si];
    int len1 = 0;  // the same in the 32 lanes of a half
    for (int c0 = 0; c0 < nc; c0 += 32) {
      bool pass = false;
      if (act && c0 + hl < nc && c0 + hl != si) {
        const float4 qj = cposf[c0 + hl];
        const float dx = qj.x - qi.x, dy = qj.y - qi.y, dz = qj.z - qi.z;
        pass = dx * dx + dy * dy + dz * dz <= pfl.r2;
      }
      const uint32_t m = (uint32_t)(__ballot(pass) >> (32 * hw));
      const int r = len1 + __popc(m & hbelow);
      if (pass && r < CR_KMAX) row[r] = c0 + hl;
      len1 += __popc(m);
    }
    if (len1 > CR_KMAX) {
      if (hl == 0) atomicOr(flag, 8);
      len1 = CR_KMAX;
    }
    __builtin_amdgcn_wave_barrier();  // row stores before the loads below
    const int lmax = max(len1, __shfl_xor(len1, 32, 64));
    const double4 pi = make_double4(cx[si], cy[si], cz[si], (double)qi.w);
    int len = 0;
    for (int e0 = 0; e0 < lmax; e0 += 32) {
      bool hit = false;
      int32_t j = -1;
      if (e0 + hl < len1) {
        const int sj = row[e0 + hl];
        j = cid[sj];
        const double4 pj = make_double4(cx[sj], cy[sj], cz[sj], (double)cposf[sj].w);
        hit = ipg_pe_sumsq(g.box, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z) <= smax && snap_cut_ok(g.box, pi, pj, sc);
      }
      const uint32_t m = (uint32_t)(__ballot(hit) >> (32 * hw));
      const int r = len + __popc(m & hbelow);
      if (hit && r < K) ell[(int64_t)i * K + r] = j;
      len += __popc(m);
    }
    if (act && hl == 0) {
      if (len > K) atomicOr(flag, 8);
      cnt[i] = len < K ? len : K;
    }
    __builtin_amdgcn_wave_barrier();  // loads done before the next rows are written
  }
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

// hw = 2: cells of half the cutoff, whose 5x5x5 neighborhood holds 42%
// fewer candidates than the 3x3x3 one of full-cutoff cells.
static Grid make_grid(ipg_ctx* ctx, const DevSystem& s, double rcut, int hw = 1) {
  Grid g{};
  g.box = s.box;
  g.hw = hw;
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
  double cw = rcut * (1.0 + 1e-9) / hw;
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
    g.full[d] = (s.box.per[d] && g.nc[d] < 2 * hw + 1) ? 1 : 0;
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

// Common list builder; alloc(name, bytes) supplies device storage.
// ellK > 0 (with sc, sorted): the first pass also writes each atom's hits
// into rows of width ellK (scratch "nl_ell"); when no row overflowed the
// returned list has L.ellK = ellK and nbr is left for the caller to fill
// from the rows (the fill and sort passes are skipped).  An overflow sets
// bit 8 of *capflag; exact sizing then runs the fill and sort passes.
template <typename Alloc>
static DevList build_list_impl(ipg_ctx* ctx, const DevSystem& s, double rcut, int kind, bool sorted, bool want_disp,
                               int64_t cap, int32_t* capflag, const SnapCutDev* sc, Alloc&& alloc, int ellK = 0,
                               bool clear_flag = false, bool defer = false, bool binned = false) {
  IPG_REQUIRE(std::isfinite(rcut) && rcut > 0, "rcutoff must be finite and > 0");
  IPG_REQUIRE(!clear_flag || capflag, "internal: flag clear without a flag");
  DevList L;
  L.n = s.n;
  L.off = (int64_t*)alloc("off", sizeof(int64_t) * (s.n + 1));
  if (s.n == 0) {
    zero2(ctx, L.off, sizeof(int64_t), capflag, clear_flag ? sizeof(int32_t) : 0);
    L.nbr = (int32_t*)alloc("nbr", 4);
    L.disp = want_disp ? (double*)alloc("disp", 8) : nullptr;
    return L;
  }
  // large systems: half-cutoff cells (the neighbor scan dominates there)
  Grid g = make_grid(ctx, s, rcut, s.n >= HALF_CELL_MIN_N ? 2 : 1);
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
  SnapCutDev scd{};
  if (sc) scd = *sc;
  // expected neighbors per atom at the grid's mean density
  double vol = (double)g.ncell;
  for (int d = 0; d < 3; d++) vol *= g.w[d];
  const double expect = std::min((double)s.n / vol * 4.18879020478639 * rcut * rcut * rcut, (double)s.n);
  if (ellK < 0) ellK = (int)std::min<double>(4096.0, std::ceil((1.5 * expect + 32.0) / 32.0) * 32.0);
  const bool ell_mode = ellK > 0 && (sc ? sorted : !sorted);
  L.ellK_req = ell_mode ? ellK : 0;
  int32_t* ell = ell_mode ? scratch_t<int32_t>(ctx, "nl_ell", (size_t)s.n * ellK) : nullptr;
  int32_t* cnt = scratch_t<int32_t>(ctx, "nl_cnt", s.n);
  const unsigned gb = grid_for(s.n, 256);
  // small periodic SNAP lists: cell-pair rows (every atom's cell scanned by
  // every cell's block: n * ncell bounded; at most 20 atoms per cell on
  // average, so the LDS lists of a cell and its neighborhood rarely fill)
  bool packed = !binned && g.hw == 1 && sc && ell_mode && ellK <= CR_KMAX && s.n <= CR_MAXN &&
                g.ncell <= CR_MAXCELL && s.n * g.ncell <= ((int64_t)1 << 25) && s.n <= 20 * g.ncell;
  for (int d = 0; d < 3; d++) packed = packed && s.box.per[d] && g.nc[d] >= 4 && g.nc[d] <= 1023;
  if (packed) {
    IPG_REQUIRE(capflag != nullptr, "internal: row-buffer scan without a flag");
    uint16_t* cellc = scratch_t<uint16_t>(ctx, "nl_cellc", (s.n + 7) / 8 * 8);  // read 8 at a time
    double4* pw = scratch_t<double4>(ctx, "nl_pw", s.n);
    float4* pfw = scratch_t<float4>(ctx, "nl_pf", s.n);
    launch(ctx, "cell_pack", [&] {
      k_cell_pack<<<gb, 256, 0, ctx->stream>>>(s.n, s.pos, s.species, g, cellc, pw, pfw,
                                               clear_flag ? capflag : nullptr);
    });
    launch(ctx, "nlist_rows", [&] {
      k_snapc_cell_rows<<<(unsigned)g.ncell, 256, 0, ctx->stream>>>((int32_t)s.n, g, pf, (const int4*)cellc, pw,
                                                                     pfw, smax, scd, ellK, ell, cnt, capflag);
    });
  }
  auto rebuild_binned = [&] {
    return build_list_impl(ctx, s, rcut, kind, sorted, want_disp, cap, capflag, sc, alloc, ellK, clear_flag, defer,
                           true);
  };
  int32_t* cellof = packed ? nullptr : scratch_t<int32_t>(ctx, "nl_cellof", s.n);
  int32_t* ccount = packed ? nullptr : scratch_t<int32_t>(ctx, "nl_ccount", 2 * g.ncell + 2);
  int32_t* cfill = packed ? nullptr : ccount + g.ncell + 1;
  int64_t* cstart = packed ? nullptr : scratch_t<int64_t>(ctx, "nl_cstart", g.ncell + 1);
  int32_t* ctmp = packed ? nullptr : scratch_t<int32_t>(ctx, "nl_ctmp", s.n);
  int32_t* catom = packed ? nullptr : scratch_t<int32_t>(ctx, "nl_catom", s.n);
  double4* spos = packed ? nullptr : scratch_t<double4>(ctx, "nl_spos", s.n);
  float4* sposf = packed ? nullptr : scratch_t<float4>(ctx, "nl_sposf", s.n);
  const size_t cells_lds = sizeof(int32_t) * (size_t)(s.n + g.ncell + 1);
  // the ordering block is latency-bound on one CU: 8192 atoms took 27 us
  // there against about 20 us for the four kernels, 1000 atoms a few us
  if (packed) {
    // rows written above
  } else if (s.n <= 4096 && cells_lds <= 48 * 1024) {
    double4* pw = scratch_t<double4>(ctx, "nl_pw", s.n);
    float4* pfw = scratch_t<float4>(ctx, "nl_pf", s.n);
    launch(ctx, "cell_assign", [&] {
      k_cell_ids<<<gb, 256, 0, ctx->stream>>>(s.n, s.pos, s.species, g, cellof, pw, pfw,
                                              clear_flag ? capflag : nullptr);
    });
    launch(ctx, "cell_rank", [&] {
      k_cells_order<<<1, CELLS_BLOCK, cells_lds, ctx->stream>>>((int32_t)s.n, (int32_t)g.ncell, cellof, pw, pfw,
                                                                cstart, catom, spos, sposf);
    });
  } else if (s.n <= 16384 && sizeof(int32_t) * (size_t)(g.ncell + 1) <= 48 * 1024) {
    // three launches: cells (and the flag clear), slots in one block, then
    // the parallel rank and gather
    launch(ctx, "cell_assign", [&] {
      k_cell_ids<<<gb, 256, 0, ctx->stream>>>(s.n, s.pos, s.species, g, cellof, nullptr, nullptr,
                                              clear_flag ? capflag : nullptr);
    });
    launch(ctx, "cell_place", [&] {
      k_cells_slots<<<1, CELLS_BLOCK, sizeof(int32_t) * (size_t)(g.ncell + 1), ctx->stream>>>(
          (int32_t)s.n, (int32_t)g.ncell, cellof, cstart, ctmp);
    });
    launch(ctx, "cell_rank", [&] {
      k_cell_rank<<<gb, 256, 0, ctx->stream>>>(s.n, ctmp, cellof, cstart, s.pos, s.species, g, catom, spos, sposf);
    });
  } else {
    // the cell counters, and the caller's flag word when this build starts
    // its step (one fill launch for both)
    zero2(ctx, ccount, sizeof(int32_t) * (2 * g.ncell + 2), capflag, clear_flag ? sizeof(int32_t) : 0);
    launch(ctx, "cell_assign", [&] { k_cell_assign<<<gb, 256, 0, ctx->stream>>>(s.pos, s.n, g, cellof, ccount); });
    exclusive_scan(ctx, ccount, g.ncell, cstart, false);
    launch(ctx, "cell_place", [&] { k_cell_place<<<gb, 256, 0, ctx->stream>>>(s.n, cellof, cstart, cfill, ctmp); });
    launch(ctx, "cell_rank", [&] {
      k_cell_rank<<<gb, 256, 0, ctx->stream>>>(s.n, ctmp, cellof, cstart, s.pos, s.species, g, catom, spos, sposf);
    });
  }
  // lanes per atom from the expected candidates per contiguous row range
  // lanes per atom: from the expected candidates per contiguous row range,
  // or with half-cutoff cells (ranges walked as one list) per atom
  int LP;
  if (g.hw == 1) {
    const double cand = (double)s.n / (double)g.ncell * (g.full[0] ? g.nc[0] : std::min(3, g.nc[0]));
    LP = cand >= 80 ? 64 : (cand >= 18 ? 32 : (cand >= 9 ? 16 : 8));
  } else {
    double cand = (double)s.n / (double)g.ncell;
    for (int d = 0; d < 3; d++) cand *= g.full[d] ? g.nc[d] : std::min(5, g.nc[d]);
    LP = cand >= 128 ? 32 : (cand >= 48 ? 16 : 8);
  }
  const unsigned gl = grid_for(s.n * LP, 256);
  auto scan = [&](auto lp, auto mode, int64_t* offp, int32_t* nbrp) {
    constexpr int LPc = decltype(lp)::value;
    constexpr int M = decltype(mode)::value;
    auto go = [&](auto sc_c, auto hw_c) {
      constexpr bool SCc = decltype(sc_c)::value;
      constexpr int HWc = decltype(hw_c)::value;
      k_list_scan<LPc, M, SCc, HWc><<<gl, 256, 0, ctx->stream>>>(s.n, g, pf, catom, spos, sposf, cellof, cstart,
                                                                 smax, kind, cnt, offp, nbrp, L.npairs, scd, ellK,
                                                                 capflag);
    };
    if (g.hw == 2) {
      if (sc) go(std::true_type(), std::integral_constant<int, 2>());
      else go(std::false_type(), std::integral_constant<int, 2>());
    } else {
      if (sc) go(std::true_type(), std::integral_constant<int, 1>());
      else go(std::false_type(), std::integral_constant<int, 1>());
    }
  };
  auto dispatch = [&](auto mode, int64_t* offp, int32_t* nbrp) {
    if (LP == 64) scan(std::integral_constant<int, 64>(), mode, offp, nbrp);
    else if (LP == 32) scan(std::integral_constant<int, 32>(), mode, offp, nbrp);
    else if (LP == 16) scan(std::integral_constant<int, 16>(), mode, offp, nbrp);
    else scan(std::integral_constant<int, 8>(), mode, offp, nbrp);
  };
  if (packed) {
    // rows and counts written by the cell-pair kernel
  } else if (ell_mode) {
    IPG_REQUIRE(capflag != nullptr, "internal: row-buffer scan without a flag");
    launch(ctx, "nlist_count", [&] { dispatch(std::integral_constant<int, SCAN_ELL>(), nullptr, ell); });
  } else {
    launch(ctx, "nlist_count", [&] { dispatch(std::integral_constant<int, SCAN_COUNT>(), nullptr, nullptr); });
  }
  if (ell_mode && defer && cap <= 0) {
    L.npairs = (int64_t)(expect * (double)s.n);
    L.exact = false;
    L.ellK = ellK;
    L.ell = ell;
    L.cnt = cnt;
    L.deferred = true;
    return L;
  }
  // exact size (reads the total back) or a capacity given by the caller
  // (no host synchronization; the device total is L.off[n], clamped to the
  // capacity, and capflag records an overflow)
  int64_t np = cap;
  if (cap > 0) exclusive_scan(ctx, cnt, s.n, L.off, false, cap, capflag, 1);
  else if (ell_mode) exclusive_scan(ctx, cnt, s.n, L.off, false);  // total read with the flag below
  else np = exclusive_scan(ctx, cnt, s.n, L.off, true);
  L.exact = cap <= 0;
  if (ell_mode) {
    // The rows are used unless one overflowed.  With a capacity the
    // overflow bit makes the caller recompute; with exact sizing it is read
    // here (with the total: one synchronization), and the list is recounted
    // and filled without the row buffer.
    bool ok = true;
    if (cap <= 0) {
      int32_t f = 0;
      d2h(ctx, &np, L.off + s.n, sizeof(int64_t));
      d2h(ctx, &f, capflag, sizeof(int32_t));
      sync(ctx);
      ok = (f & 8) == 0;
    }
    if (ok) {
      L.npairs = np;
      L.nbr = (int32_t*)alloc("nbr", sizeof(int32_t) * (np > 0 ? np : 1));
      L.disp = want_disp ? (double*)alloc("disp", sizeof(double) * 3 * (np > 0 ? np : 1)) : nullptr;
      L.ellK = ellK;
      L.ell = ell;
      L.cnt = cnt;
      L.ell_sorted = packed;
      return L;
    }
    if (packed) return rebuild_binned();  // a row, a cell or a neighborhood overflowed: the binned scan
    launch(ctx, "nlist_count", [&] { dispatch(std::integral_constant<int, SCAN_COUNT>(), nullptr, nullptr); });
    np = exclusive_scan(ctx, cnt, s.n, L.off, true);
    std::vector<int32_t> hc((size_t)s.n);
    d2h(ctx, hc.data(), cnt, sizeof(int32_t) * s.n);
    sync(ctx);
    const int widest = *std::max_element(hc.begin(), hc.end());
    L.ellK_req = (int)std::min<int64_t>(4096, ((int64_t)widest + 16 + 31) / 32 * 32);  // width for a next call
  }
  L.npairs = np;
  L.nbr = (int32_t*)alloc("nbr", sizeof(int32_t) * (np > 0 ? np : 1));
  L.disp = want_disp ? (double*)alloc("disp", sizeof(double) * 3 * (np > 0 ? np : 1)) : nullptr;
  int32_t* fill_to = L.nbr;
  if (sorted && np > 0) fill_to = scratch_t<int32_t>(ctx, "nl_unsorted", np);
  launch(ctx, "nlist_fill", [&] { dispatch(std::integral_constant<int, SCAN_FILL>(), L.off, fill_to); });
  if (sorted && np > 0)
    launch(ctx, "nlist_sort", [&] {
      k_seg_rank_sort<<<grid_for(np, 256), 256, 0, ctx->stream>>>(s.n, L.off, fill_to, L.nbr);
    });
  if (want_disp && np > 0)
    launch(ctx, "nlist_disp", [&] { k_list_disp<<<gb, 256, 0, ctx->stream>>>(s.n, L.off, L.nbr, s.pos, s.box, L.disp); });
  return L;
}

DevList build_list(ipg_ctx* ctx, const DevSystem& s, double rcut, int kind, bool sorted, bool want_disp,
                   const char* tag, int64_t cap, int32_t* capflag, int ellK, bool defer) {
  std::string t(tag);
  IPG_REQUIRE(ellK == 0 || (!sorted && capflag), "internal: row-buffer lists are unsorted and need a flag");
  IPG_REQUIRE(!defer || (ellK != 0 && !want_disp), "internal: a deferred list has rows only");
  return build_list_impl(
      ctx, s, rcut, kind, sorted, want_disp, cap, capflag, nullptr,
      [&](const char* what, size_t bytes) { return scratch(ctx, (t + "_" + what).c_str(), bytes); }, ellK, false,
      defer);
}

DevList build_list_owned(ipg_ctx* ctx, const DevSystem& s, double rcut, int kind, bool sorted, bool want_disp) {
  std::vector<void*> got;
  try {
    return build_list_impl(ctx, s, rcut, kind, sorted, want_disp, 0, nullptr, nullptr, [&](const char*, size_t bytes) {
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
  ne = (unsigned)ne < (unsigned)cut.ns ? ne : 0;  // an index outside [0, ns) reads species 0
  double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  int64_t a = offA[i], b = offA[i + 1];
  int count = 0;
  int64_t w = FILL ? off[i] : 0;
  for (int64_t q = a; q < b; q++) {
    int32_t j = nbrA[q];
    int64_t m = q - a + 1;  // 1-based enumerate index (utl.jl:75)
    int me = cut.reference ? appear[m - 1] : (species ? species[j] : 0);
    me = (unsigned)me < (unsigned)cut.ns ? me : 0;
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

// Per pair of a sorted CUTOFF-mode SNAP list: center atom, r = x_j - x_i
// (get_displacement, as utl.jl:84), optionally |r| and the index of the
// reverse pair (j -> i) found by binary search in j's sorted row.  A pair
// whose reverse is missing (possible only when r.r lies within an ulp of the
// cutoff under periodic wrapping) raises bit 4 of *flag.
__global__ void k_snapc_pairs(int64_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ nbr,
                              const double* __restrict__ pos, IpgBox box, int32_t* __restrict__ center,
                              double* __restrict__ rvec, double* __restrict__ R, int32_t* __restrict__ rev,
                              int32_t* __restrict__ flag) {
#pragma clang fp contract(off)
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= off[n]) return;  // device pair count (the grid may cover a capacity)
  int64_t lo = 0, hi = n;   // center: off[lo] <= q < off[lo + 1]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= q) lo = mid;
    else hi = mid;
  }
  const int32_t i = (int32_t)lo, j = nbr[q];
  center[q] = i;
  const double rx = -1.0 * ipg_disp1(pos[3 * i], pos[3 * j], box.L[0], box.per[0]);
  const double ry = -1.0 * ipg_disp1(pos[3 * i + 1], pos[3 * j + 1], box.L[1], box.per[1]);
  const double rz = -1.0 * ipg_disp1(pos[3 * i + 2], pos[3 * j + 2], box.L[2], box.per[2]);
  rvec[3 * q] = rx;
  rvec[3 * q + 1] = ry;
  rvec[3 * q + 2] = rz;
  if (R) R[q] = sqrt(rx * rx + ry * ry + rz * rz);
  if (rev) {
    int64_t a = off[j], b = off[j + 1];
    const int64_t end = b;
    while (a < b) {  // first entry >= i in j's ascending row
      const int64_t m = (a + b) >> 1;
      if (nbr[m] < i) a = m + 1;
      else b = m;
    }
    if (a < end && nbr[a] == i) {
      rev[q] = (int32_t)a;
    } else {
      rev[q] = (int32_t)q;
      atomicOr(flag, 4);
    }
  }
}

// Sorted CUTOFF-mode rows from the first pass's row buffer: each group of LP
// lanes stages atom i's hits in LDS, ranks them (the sorted order of the
// inrange list), and writes nbr, center, r = x_j - x_i (get_displacement,
// utl.jl:84) and optionally |r|.
template <int LP>
__global__ __launch_bounds__(256) void k_snapc_finish(int64_t n, int K, const int32_t* __restrict__ ell,
                                                      const int32_t* __restrict__ cnt,
                                                      const int64_t* __restrict__ off, const double* __restrict__ pos,
                                                      IpgBox box1, const int32_t* __restrict__ cfg,
                                                      const IpgBox* __restrict__ boxes, int32_t* __restrict__ nbr,
                                                      int32_t* __restrict__ center, double* __restrict__ rvec,
                                                      double* __restrict__ R) {
#pragma clang fp contract(off)
  extern __shared__ int32_t rowbuf[];  // [256 / LP][K]
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = gt / LP;
  const int l = (int)(gt % LP);
  int32_t* row = rowbuf + (threadIdx.x / LP) * K;
  const bool act = i < n;
  int c = 0;
  int64_t a = 0, b = 0;
  if (act) {
    c = min(cnt[i], K);
    a = off[i];
    b = off[i + 1];
    for (int k = l; k < c; k += LP) row[k] = ell[i * K + k];
  }
  __syncthreads();
  if (!act) return;
  const IpgBox box = cfg ? boxes[cfg[i]] : box1;  // packed configurations: the atom's own box
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  for (int k = l; k < c; k += LP) {
    const int32_t j = row[k];
    int r = 0;
    for (int m = 0; m < c; m++) r += row[m] < j;
    const int64_t q = a + r;
    if (q >= b) continue;  // row cut by a capacity (flagged by the scan)
    nbr[q] = j;
    center[q] = (int32_t)i;
    const double rx = -1.0 * ipg_disp1(xi, pos[3 * j], box.L[0], box.per[0]);
    const double ry = -1.0 * ipg_disp1(yi, pos[3 * j + 1], box.L[1], box.per[1]);
    const double rz = -1.0 * ipg_disp1(zi, pos[3 * j + 2], box.L[2], box.per[2]);
    rvec[3 * q] = rx;
    rvec[3 * q + 1] = ry;
    rvec[3 * q + 2] = rz;
    if (R) R[q] = sqrt(rx * rx + ry * ry + rz * rz);
  }
}

// The same from rows already sorted by j (the cell-pair kernel's): no rank
// pass, and with rev the reverse pair of q = (i -> j) is found by binary
// search in j's row of the row buffer (complete when this kernel starts):
// rev[q] = off[j] + the position of i there.  A pair whose reverse is
// missing (r.r within an ulp of the cutoff under wrapping, or j's row cut
// by a capacity) points at itself and sets bit 4 of *flag.
template <bool REV>
__global__ __launch_bounds__(256) void k_snapc_finish_sorted(int64_t n, int K, const int32_t* __restrict__ ell,
                                                             const int32_t* __restrict__ cnt,
                                                             const int64_t* __restrict__ off,
                                                             const double* __restrict__ pos, IpgBox box,
                                                             int32_t* __restrict__ nbr, int32_t* __restrict__ center,
                                                             double* __restrict__ rvec, double* __restrict__ R,
                                                             int32_t* __restrict__ rev, int32_t* __restrict__ flag) {
#pragma clang fp contract(off)
  constexpr int LP = 32;
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = gt / LP;
  const int l = (int)(gt % LP);
  if (i >= n) return;
  const int c = min(cnt[i], K);
  const int64_t a = off[i], b = off[i + 1];
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  for (int k = l; k < c; k += LP) {
    const int32_t j = ell[i * K + k];
    const int64_t q = a + k;
    if (q >= b) break;  // row cut by a capacity (flagged by the scan)
    nbr[q] = j;
    center[q] = (int32_t)i;
    const double rx = -1.0 * ipg_disp1(xi, pos[3 * j], box.L[0], box.per[0]);
    const double ry = -1.0 * ipg_disp1(yi, pos[3 * j + 1], box.L[1], box.per[1]);
    const double rz = -1.0 * ipg_disp1(zi, pos[3 * j + 2], box.L[2], box.per[2]);
    rvec[3 * q] = rx;
    rvec[3 * q + 1] = ry;
    rvec[3 * q + 2] = rz;
    if (R) R[q] = sqrt(rx * rx + ry * ry + rz * rz);
    if constexpr (REV) {
      const int32_t* rj = ell + (int64_t)j * K;
      const int cj = min(cnt[j], K);
      int lo = 0, hi = cj;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rj[mid] < i) lo = mid + 1;
        else hi = mid;
      }
      if (lo < cj && rj[lo] == i && off[j] + lo < off[j + 1]) {
        rev[q] = (int32_t)(off[j] + lo);
      } else {
        rev[q] = (int32_t)q;
        atomicOr(flag, 4);
      }
    }
  }
}

// rev[q]: the pair j -> i of pair q = (i -> j), by binary search in j's
// sorted row (the rows are complete once the finish kernel has run).  A
// pair whose reverse is missing (r.r within an ulp of the cutoff under
// wrapping, or j's row cut by a capacity) points at itself and sets bit 4.
__global__ void k_rev_search(int64_t np_max, int64_t n, const int64_t* __restrict__ off,
                             const int32_t* __restrict__ nbr, const int32_t* __restrict__ center,
                             int32_t* __restrict__ rev, int32_t* __restrict__ flag) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np_max || q >= off[n]) return;
  const int32_t i = center[q], j = nbr[q];
  const int64_t end = off[j + 1];
  int64_t lo = off[j], hi = end;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (nbr[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  if (lo < end && nbr[lo] == i) {
    rev[q] = (int32_t)lo;
  } else {
    rev[q] = (int32_t)q;
    atomicOr(flag, 4);
  }
}

SnapListDev build_snap_list(ipg_ctx* ctx, const DevSystem& s, const int32_t* d_appear, const SnapCut& cut, int need,
                            const char* tag, int64_t capA, int64_t capB, int32_t* flag, int ellK,
                            bool clear_flag) {
  std::string t(tag);
  SnapListDev out;
  const int64_t n = s.n;
  if (!cut.reference) {
    // CUTOFF mode: both membership tests of utl.jl:69-90 in one species-aware
    // cell scan, rows sorted by j (the order of the stage-A inrange list)
    SnapCutDev cd{};
    cd.ns = cut.ns;
    for (int k = 0; k < cut.ns * cut.ns; k++) cd.cutsq[k] = cut.cutsq[k];
    IPG_REQUIRE(flag != nullptr, "internal: SNAP list without a flag word");
    DevList L = build_list_impl(
        ctx, s, cut.query, KIND_FULL, true, false, capB, flag, &cd,
        [&](const char* what, size_t bytes) { return scratch(ctx, (t + "_" + what).c_str(), bytes); },
        ellK != 0 ? ellK : -1, clear_flag);
    out.list = L;
    out.npairs_a = L.npairs;
    out.ellK = L.ellK_req;
    IPG_REQUIRE(L.npairs < ((int64_t)1 << 31), "too many SNAP pairs for 32-bit pair indices");
    const int64_t npa = L.npairs > 0 ? L.npairs : 1;
    out.center = scratch_t<int32_t>(ctx, (t + "_center").c_str(), npa);
    out.rvec = scratch_t<double>(ctx, (t + "_rvec").c_str(), 3 * npa);
    out.R = (need & SNAPL_R) ? scratch_t<double>(ctx, (t + "_R").c_str(), npa) : nullptr;
    out.rev = (need & SNAPL_SYM) ? scratch_t<int32_t>(ctx, (t + "_rev").c_str(), npa) : nullptr;
    IPG_REQUIRE(!out.rev || flag, "internal: reverse-pair search without a flag");
    if (n > 0 && L.npairs > 0 && L.ellK > 0 && L.ell_sorted) {
      launch(ctx, "snap_list_finish", [&] {
        const unsigned gf = grid_for(n * 32, 256);
        if (out.rev)
          k_snapc_finish_sorted<true><<<gf, 256, 0, ctx->stream>>>(n, L.ellK, L.ell, L.cnt, L.off, s.pos, s.box, L.nbr,
                                                                  out.center, out.rvec, out.R, out.rev, flag);
        else
          k_snapc_finish_sorted<false><<<gf, 256, 0, ctx->stream>>>(n, L.ellK, L.ell, L.cnt, L.off, s.pos, s.box,
                                                                   L.nbr, out.center, out.rvec, out.R, nullptr, flag);
      });
    } else if (n > 0 && L.npairs > 0 && L.ellK > 0) {
      const int K = L.ellK;
      const int LP = K >= 128 ? 64 : 32;
      const size_t sm = sizeof(int32_t) * (size_t)(256 / LP) * K;
      IPG_REQUIRE(sm <= 64 * 1024, "internal: neighbor row buffer too wide");
      launch(ctx, "snap_list_finish", [&] {
        if (LP == 64)
          k_snapc_finish<64><<<grid_for(n * 64, 256), 256, sm, ctx->stream>>>(
              n, K, L.ell, L.cnt, L.off, s.pos, s.box, nullptr, nullptr, L.nbr, out.center, out.rvec, out.R);
        else
          k_snapc_finish<32><<<grid_for(n * 32, 256), 256, sm, ctx->stream>>>(
              n, K, L.ell, L.cnt, L.off, s.pos, s.box, nullptr, nullptr, L.nbr, out.center, out.rvec, out.R);
      });
      if (out.rev)
        launch(ctx, "snap_list_rev", [&] {
          k_rev_search<<<grid_for(L.npairs, 256), 256, 0, ctx->stream>>>(L.npairs, n, L.off, L.nbr, out.center,
                                                                        out.rev, flag);
        });
    } else if (n > 0 && L.npairs > 0) {
      launch(ctx, "snap_pairs", [&] {
        k_snapc_pairs<<<grid_for(L.npairs, 256), 256, 0, ctx->stream>>>(n, L.off, L.nbr, s.pos, s.box, out.center,
                                                                       out.rvec, out.R, out.rev, flag);
      });
    }
    if (need & SNAPL_REVERSE) {
      out.roff = scratch_t<int64_t>(ctx, (t + "_roff").c_str(), n + 1);
      out.rpair = scratch_t<int32_t>(ctx, (t + "_rpair").c_str(), npa);
      build_center_and_reverse(ctx, n, L.npairs, L.off, L.nbr, nullptr, out.roff, out.rpair, t.c_str());
    }
    return out;
  }
  const bool want_reverse = (need & (SNAPL_REVERSE | SNAPL_SYM)) != 0;
  int32_t* capflag = flag;
  DevList A = build_list(ctx, s, cut.query, KIND_FULL_SELF, true, false, (t + "_A").c_str(), capA, capflag);
  out.npairs_a = A.npairs;
  SnapCutDev cd{};
  cd.ns = cut.ns;
  cd.reference = cut.reference;
  for (int k = 0; k < cut.ns * cut.ns; k++) cd.cutsq[k] = cut.cutsq[k];
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
    exclusive_scan(ctx, cnt, n, L.off, false, capB, capflag, 2);
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

// ---------------------------------------------------------------------------
// Packed configurations (batched descriptors): CUTOFF-mode SNAP lists of
// many small systems in one pass.  Atom i of configuration c = cfg[i]
// scans the atoms of c (start[c] .. start[c+1]) in index order with c's own
// box, applying the two membership tests of the single-system list
// (PeriodicEuclidean distance against the query radius, then the
// species-pair cutoff on get_displacement, utl.jl:69-90); its hits land in
// its row of `ell` in ascending j.
// ---------------------------------------------------------------------------
template <int LP>
__global__ __launch_bounds__(256) void k_batch_scan(int64_t n, const int32_t* __restrict__ cfg,
                                                    const int64_t* __restrict__ start,
                                                    const IpgBox* __restrict__ boxes, const double* __restrict__ pos,
                                                    const int32_t* __restrict__ species, double smax, SnapCutDev sc,
                                                    int32_t* __restrict__ cnt, int32_t* __restrict__ ell, int K,
                                                    int32_t* __restrict__ flag) {
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = gt / LP;
  const int l = (int)(gt % LP);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;  // whole groups leave together
  const uint64_t gmask = (LP == 64 ? ~0ull : ((1ull << LP) - 1ull)) << (lane - l);
  const uint64_t below = (1ull << lane) - 1ull;
  const int c = cfg[i];
  const IpgBox box = boxes[c];
  const int64_t lo = start[c], hi = start[c + 1];
  const double4 pi = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], species ? (double)species[i] : 0.0);
  int count = 0;
  for (int64_t s0 = lo; s0 < hi; s0 += LP) {
    const int64_t j = s0 + l;
    bool hit = false;
    if (j < hi && j != i) {
      const double4 pj =
          make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], species ? (double)species[j] : 0.0);
      hit = ipg_pe_sumsq(box, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z) <= smax && snap_cut_ok(box, pi, pj, sc);
    }
    const uint64_t m = __ballot(hit) & gmask;
    const int r = count + __popcll(m & below);
    if (hit && r < K) ell[i * K + r] = (int32_t)j;
    count += __popcll(m);
  }
  if (l == 0) {
    if (count > K) {
      atomicOr(flag, 8);
      count = K;  // the row holds what was written, so every list slot gets an entry
    }
    cnt[i] = count;
  }
}

SnapListDev build_snap_list_batch(ipg_ctx* ctx, const BatchSys& b, const SnapCut& cut, int need, const char* tag,
                                  int64_t cap, int32_t* flag, int ellK, int lp) {
  IPG_REQUIRE(!cut.reference, "batched SNAP lists use CUTOFF mode");
  IPG_REQUIRE(flag != nullptr, "internal: batched list without a flag word");
  std::string t(tag);
  SnapListDev out;
  const int64_t n = b.n;
  SnapCutDev cd{};
  cd.ns = cut.ns;
  for (int k = 0; k < cut.ns * cut.ns; k++) cd.cutsq[k] = cut.cutsq[k];
  const double smax = sqrt_le_threshold(cut.query);
  DevList& L = out.list;
  L.n = n;
  L.off = scratch_t<int64_t>(ctx, (t + "_off").c_str(), n + 1);
  int32_t* cnt = scratch_t<int32_t>(ctx, (t + "_cnt").c_str(), n > 0 ? n : 1);
  int K = ellK > 0 ? ellK : (int)std::min<double>(4096.0, std::ceil((1.5 * b.expect + 32.0) / 32.0) * 32.0);
  const int LP = lp == 16 || lp == 32 || lp == 64 ? lp : 32;
  int64_t np = 0;
  for (;;) {
    int32_t* ell = scratch_t<int32_t>(ctx, (t + "_ell").c_str(), (size_t)std::max<int64_t>(n, 1) * K);
    if (n > 0)
      launch(ctx, "batch_scan", [&] {
        const unsigned g = grid_for(n * LP, 256);
        if (LP == 64)
          k_batch_scan<64><<<g, 256, 0, ctx->stream>>>(n, b.cfg, b.start, b.boxes, b.pos, b.species, smax, cd, cnt,
                                                       ell, K, flag);
        else if (LP == 16)
          k_batch_scan<16><<<g, 256, 0, ctx->stream>>>(n, b.cfg, b.start, b.boxes, b.pos, b.species, smax, cd, cnt,
                                                       ell, K, flag);
        else
          k_batch_scan<32><<<g, 256, 0, ctx->stream>>>(n, b.cfg, b.start, b.boxes, b.pos, b.species, smax, cd, cnt,
                                                       ell, K, flag);
      });
    if (cap > 0) {
      np = cap;
      exclusive_scan(ctx, cnt, n, L.off, false, cap, flag, 1);
      L.exact = false;
      break;
    }
    np = exclusive_scan(ctx, cnt, n, L.off, true);
    int32_t f = 0;
    d2h(ctx, &f, flag, sizeof(int32_t));
    sync(ctx);
    if (!(f & 8) || K >= 4096) {
      IPG_REQUIRE(!(f & 8), "a neighbor row exceeds 4096 entries");
      break;
    }
    K *= 2;  // a row overflowed: scan again with wider rows
    zero(ctx, flag, sizeof(int32_t));
  }
  L.npairs = np;
  L.ellK = K;
  L.ellK_req = K;
  out.ellK = K;
  out.npairs_a = np;
  IPG_REQUIRE(np < ((int64_t)1 << 31), "too many SNAP pairs for 32-bit pair indices");
  const int64_t npa = np > 0 ? np : 1;
  L.nbr = scratch_t<int32_t>(ctx, (t + "_nbr").c_str(), npa);
  out.center = scratch_t<int32_t>(ctx, (t + "_center").c_str(), npa);
  out.rvec = scratch_t<double>(ctx, (t + "_rvec").c_str(), 3 * npa);
  out.R = (need & SNAPL_R) ? scratch_t<double>(ctx, (t + "_R").c_str(), npa) : nullptr;
  const int32_t* ell = scratch_t<int32_t>(ctx, (t + "_ell").c_str(), (size_t)std::max<int64_t>(n, 1) * K);
  out.rev = (need & SNAPL_SYM) ? scratch_t<int32_t>(ctx, (t + "_rev").c_str(), npa) : nullptr;
  if (n > 0 && np > 0) {
    const int FL = K >= 128 ? 64 : 32;
    const size_t sm = sizeof(int32_t) * (size_t)(256 / FL) * K;
    IPG_REQUIRE(sm <= 64 * 1024, "internal: neighbor row buffer too wide");
    IpgBox none{};
    launch(ctx, "batch_list_finish", [&] {
      if (FL == 64)
        k_snapc_finish<64><<<grid_for(n * 64, 256), 256, sm, ctx->stream>>>(
            n, K, ell, cnt, L.off, b.pos, none, b.cfg, b.boxes, L.nbr, out.center, out.rvec, out.R);
      else
        k_snapc_finish<32><<<grid_for(n * 32, 256), 256, sm, ctx->stream>>>(
            n, K, ell, cnt, L.off, b.pos, none, b.cfg, b.boxes, L.nbr, out.center, out.rvec, out.R);
    });
    if (out.rev)
      launch(ctx, "batch_list_rev", [&] {
        k_rev_search<<<grid_for(np, 256), 256, 0, ctx->stream>>>(np, n, L.off, L.nbr, out.center, out.rev, flag);
      });
  }
  if (need & SNAPL_REVERSE) {
    out.roff = scratch_t<int64_t>(ctx, (t + "_roff").c_str(), n + 1);
    out.rpair = scratch_t<int32_t>(ctx, (t + "_rpair").c_str(), npa);
    build_center_and_reverse(ctx, n, np, L.off, L.nbr, nullptr, out.roff, out.rpair, t.c_str());
  }
  return out;
}

// Largest squared PeriodicEuclidean displacement of any atom since the
// reference positions, as the bit pattern of a non-negative double (whose
// unsigned order is the numeric order), so an integer max gives it exactly.
__global__ void k_max_disp2(int64_t n, const double* __restrict__ pos, const double* __restrict__ ref, IpgBox box,
                            uint64_t* __restrict__ out) {
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, ipg_pe_sumsq(box, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], ref[3 * i], ref[3 * i + 1],
                             ref[3 * i + 2]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __shared__ double wmax[4];  // blockDim 256: one atomic per block
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    atomicMax((unsigned long long*)out, (unsigned long long)__double_as_longlong(m));
  }
}

// (Re)builds a skin list at cut + skin from the system's current positions
// and records them as the reference.  The new buffers replace the old ones
// only once the build succeeded.
static void skin_build(ipg_nlist* h, const DevSystem& s) {
  ipg_ctx* ctx = h->ctx;
  DevList L = build_list_owned(ctx, s, h->rcut, h->kind == IPG_LIST_HALF ? KIND_HALF : KIND_FULL, false, false);
  if (!h->ref) {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(double) * 3 * std::max<int64_t>(s.n, 1) + sizeof(uint64_t)) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(L.off);
      (void)hipFree(L.nbr);
      fail(IPG_ERR_NOMEM, "hipMalloc failed for a skin list");
    }
    h->ref = (double*)p;
    h->dmax = (uint64_t*)(h->ref + 3 * std::max<int64_t>(s.n, 1));
  }
  if (s.n > 0)
    IPG_HIP(hipMemcpyAsync(h->ref, s.pos, sizeof(double) * 3 * s.n, hipMemcpyDeviceToDevice, ctx->stream));
  IPG_HIP(hipStreamSynchronize(ctx->stream));  // the old list may still be read by enqueued work
  if (h->list.off) (void)hipFree(h->list.off);
  if (h->list.nbr) (void)hipFree(h->list.nbr);
  if (h->list.disp) (void)hipFree(h->list.disp);
  h->list = L;
  h->box = s.box;
  h->builds++;
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
    h->cut = rcut;
    h->kind = kind;
    h->box = s.box;
    h->builds = 1;
    *out = h;
  });
}

int ipg_nlist_build_skin_dev(ipg_ctx* ctx, const ipg_system* sys, double rcut, double skin, int kind,
                             ipg_nlist** out) {
  return guard([&] {
    IPG_REQUIRE(ctx && sys && out, "NULL argument");
    IPG_REQUIRE(kind == IPG_LIST_HALF || kind == IPG_LIST_FULL, "kind must be IPG_LIST_HALF or IPG_LIST_FULL");
    IPG_REQUIRE(std::isfinite(rcut) && rcut > 0, "rcutoff must be finite and > 0");
    IPG_REQUIRE(std::isfinite(skin) && skin > 0, "skin must be finite and > 0");
    IPG_HIP(hipSetDevice(ctx->device));
    DevSystem s = make_dev_system(sys);
    auto* h = new ipg_nlist();
    h->ctx = ctx;
    h->cut = rcut;
    h->skin = skin;
    h->rcut = rcut + skin;
    h->kind = kind;
    try {
      skin_build(h, s);
    } catch (...) {
      if (h->ref) (void)hipFree(h->ref);
      delete h;
      throw;
    }
    *out = h;
  });
}

int ipg_nlist_update_dev(ipg_nlist* list, const ipg_system* sys, int32_t* rebuilt) {
  return guard([&] {
    IPG_REQUIRE(list && sys, "NULL argument");
    IPG_REQUIRE(list->skin > 0, "not a skin list (build it with ipg_nlist_build_skin_dev)");
    ipg_ctx* ctx = list->ctx;
    IPG_HIP(hipSetDevice(ctx->device));
    DevSystem s = make_dev_system(sys);
    IPG_REQUIRE(s.n == list->list.n, "natoms differs from the list's");
    bool same_box = true;
    for (int d = 0; d < 3; d++)
      same_box = same_box && s.box.per[d] == list->box.per[d] && (!s.box.per[d] || s.box.L[d] == list->box.L[d]);
    bool rebuild = !same_box;
    if (!rebuild && s.n > 0) {
      // every pair within cut now was within cut + 2 max|dx| at the last
      // build; the margin covers the rounding of the distance sums
      const double half = 0.5 * list->skin - 1e-9 * list->rcut;
      zero(ctx, list->dmax, sizeof(uint64_t));
      launch(ctx, "nlist_max_disp", [&] {
        k_max_disp2<<<std::min<unsigned>(grid_for(s.n, 256), 512), 256, 0, ctx->stream>>>(s.n, s.pos, list->ref,
                                                                                          s.box, list->dmax);
      });
      uint64_t bits = 0;
      d2h(ctx, &bits, list->dmax, sizeof(uint64_t));
      sync(ctx);
      double m2;
      std::memcpy(&m2, &bits, sizeof(double));
      rebuild = !(m2 <= half * half) || half <= 0.0;
    }
    if (rebuild) skin_build(list, s);
    list->updates++;
    if (rebuilt) *rebuilt = rebuild ? 1 : 0;
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
    if (list->ref) (void)hipFree(list->ref);
    delete list;
  });
}

}  // extern "C"
