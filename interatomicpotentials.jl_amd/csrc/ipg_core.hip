// ipg_core.hip -- context, errors, device memory, profiling, scans and
// deterministic reductions shared by every entry point.
#include "ipg_internal.h"

#include <algorithm>

#include <cfloat>
#include <cstddef>

// The Julia binding (julia/InteratomicPotentialsHIP.jl: IPGSystem,
// IPGPairTerm, IPGSnapParams) and the Python mirror (ipgpu/_lib.py) lay
// these structs out by hand with natural alignment; a change here must be
// made there too.
static_assert(sizeof(ipg_system) == 64, "ipg_system size");
static_assert(offsetof(ipg_system, natoms) == 0 && offsetof(ipg_system, positions) == 8 &&
                  offsetof(ipg_system, species) == 16 && offsetof(ipg_system, box) == 24 &&
                  offsetof(ipg_system, periodic) == 48,
              "ipg_system layout");
static_assert(sizeof(ipg_pair_term) == 80, "ipg_pair_term size");
static_assert(offsetof(ipg_pair_term, kind) == 0 && offsetof(ipg_pair_term, reserved) == 4 &&
                  offsetof(ipg_pair_term, coeff) == 8 && offsetof(ipg_pair_term, rcut) == 16 &&
                  offsetof(ipg_pair_term, p) == 24 && offsetof(ipg_pair_term, species_mask) == 72,
              "ipg_pair_term layout");
static_assert(sizeof(ipg_snap_params) == 72, "ipg_snap_params size");
static_assert(offsetof(ipg_snap_params, twojmax) == 0 && offsetof(ipg_snap_params, nspecies) == 4 &&
                  offsetof(ipg_snap_params, rcutfac) == 8 && offsetof(ipg_snap_params, rmin0) == 16 &&
                  offsetof(ipg_snap_params, rfac0) == 24 && offsetof(ipg_snap_params, radii) == 32 &&
                  offsetof(ipg_snap_params, weight) == 40 && offsetof(ipg_snap_params, chem_flag) == 48 &&
                  offsetof(ipg_snap_params, bzero_flag) == 52 && offsetof(ipg_snap_params, bnorm_flag) == 56 &&
                  offsetof(ipg_snap_params, switch_flag) == 60 && offsetof(ipg_snap_params, wselfall_flag) == 64,
              "ipg_snap_params layout");

namespace ipg {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

void* scratch(ipg_ctx* ctx, const char* name, size_t bytes) {
  auto& b = ctx->scratch[name];
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) {
    // the old buffer may still be read by queued kernels
    IPG_HIP(hipStreamSynchronize(ctx->stream));
    IPG_HIP(hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  want = want + want / 4;  // headroom for slowly growing inputs
  if (hipMalloc(&b.ptr, want) != hipSuccess) {
    (void)hipGetLastError();
    b.ptr = nullptr;
    fail(IPG_ERR_NOMEM, std::string("device allocation of ") + std::to_string(want) + " bytes failed (" + name + ")");
  }
  b.bytes = want;
  ctx->scratch_gen++;
  return b.ptr;
}

hipEvent_t get_event(ipg_ctx* ctx) {
  if (ctx->free_events.empty()) {
    // create in batches: event creation is far slower than a launch
    for (int k = 0; k < 256; k++) {
      hipEvent_t e;
      IPG_HIP(hipEventCreate(&e));
      ctx->free_events.push_back(e);
    }
  }
  hipEvent_t e = ctx->free_events.back();
  ctx->free_events.pop_back();
  return e;
}

void flush_prof(ipg_ctx* ctx) {
  if (ctx->pending.empty()) return;
  IPG_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& p : ctx->pending) {
    float ms = 0.f;
    IPG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& s = ctx->stats[p.name];
    s.first += 1;
    s.second += (double)ms;
    ctx->free_events.push_back(p.a);
    ctx->free_events.push_back(p.b);
  }
  ctx->pending.clear();
}

void add_graph_prof(ipg_ctx* ctx, const std::vector<ipg_ctx::Pending>& ev) {
  if (ev.empty()) return;
  IPG_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& p : ev) {
    float ms = 0.f;
    IPG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& s = ctx->stats[p.name];
    s.first += 1;
    s.second += (double)ms;
  }
}

void h2d(ipg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  IPG_REQUIRE(!ctx->dry, "internal: host copy in an allocation-only pass");
  IPG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
}
void d2h(ipg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  IPG_REQUIRE(!ctx->dry, "internal: readback in an allocation-only pass");
  IPG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
}
void sync(ipg_ctx* ctx) { IPG_HIP(hipStreamSynchronize(ctx->stream)); }

__global__ void k_fill_zero(uint32_t* __restrict__ p, int64_t nwords) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nwords; i += stride) p[i] = 0u;
}

// Two ranges in one launch: words [0, n) of p, then [0, m) of q.
__global__ void k_fill_zero2(uint32_t* __restrict__ p, int64_t n, uint32_t* __restrict__ q, int64_t m) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n + m; i += stride) {
    if (i < n) p[i] = 0u;
    else q[i - n] = 0u;
  }
}

// Zero-fill with a kernel (no hipMemsetAsync, so the call stays capturable
// in a HIP graph).  bytes must be a multiple of 4.
void zero(ipg_ctx* ctx, void* dst, size_t bytes) {
  if (bytes == 0) return;
  int64_t nw = (int64_t)(bytes / 4);
  unsigned g = grid_for(nw, 256);
  if (g > 4096) g = 4096;
  launch(ctx, "fill_zero", [&] { k_fill_zero<<<g, 256, 0, ctx->stream>>>((uint32_t*)dst, nw); });
}

// Two zero-fills in one launch (one kernel instead of two in a captured
// step).  Byte counts must be multiples of 4; a zero count skips that range.
void zero2(ipg_ctx* ctx, void* dst, size_t bytes, void* dst2, size_t bytes2) {
  if (bytes2 == 0) return zero(ctx, dst, bytes);
  if (bytes == 0) return zero(ctx, dst2, bytes2);
  const int64_t n = (int64_t)(bytes / 4), m = (int64_t)(bytes2 / 4);
  unsigned g = grid_for(n + m, 256);
  if (g > 4096) g = 4096;
  launch(ctx, "fill_zero",
         [&] { k_fill_zero2<<<g, 256, 0, ctx->stream>>>((uint32_t*)dst, n, (uint32_t*)dst2, m); });
}

// ---------------------------------------------------------------------------
// exclusive scan (int32 counts -> int64 offsets), three-phase
// ---------------------------------------------------------------------------
constexpr int SCAN_BLOCK = 1024;

__global__ void k_scan_block_sums(const int32_t* __restrict__ c, int64_t n, int64_t* __restrict__ bsum) {
  __shared__ int64_t s[SCAN_BLOCK / 64];
  int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  int64_t v = (i < n) ? c[i] : 0;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < SCAN_BLOCK / 64; w++) t += s[w];
    bsum[blockIdx.x] = t;
  }
}

// single-block exclusive scan of the block sums (nb <= 1024*64 handled by a loop)
__global__ void k_scan_sums(int64_t* __restrict__ bsum, int64_t nb, int64_t* __restrict__ total) {
  __shared__ int64_t carry;
  __shared__ int64_t ws[SCAN_BLOCK / 64];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nb; base += SCAN_BLOCK) {
    int64_t i = base + threadIdx.x;
    int64_t v = (i < nb) ? bsum[i] : 0;
    // inclusive wave scan
    int lane = threadIdx.x & 63;
    int64_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      int64_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) ws[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t t = 0;
      for (int w = 0; w < SCAN_BLOCK / 64; w++) {
        int64_t q = ws[w];
        ws[w] = t;
        t += q;
      }
    }
    __syncthreads();
    int64_t excl = carry + ws[threadIdx.x >> 6] + x - v;
    if (i < nb) bsum[i] = excl;
    __syncthreads();
    if (threadIdx.x == SCAN_BLOCK - 1) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_final(const int32_t* __restrict__ c, int64_t n, const int64_t* __restrict__ bsum,
                             const int64_t* __restrict__ total, int64_t* __restrict__ off, int64_t cap,
                             int32_t* __restrict__ flag, int32_t bit) {
  __shared__ int64_t ws[SCAN_BLOCK / 64];
  int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  int64_t v = (i < n) ? c[i] : 0;
  int lane = threadIdx.x & 63;
  int64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < SCAN_BLOCK / 64; w++) {
      int64_t q = ws[w];
      ws[w] = t;
      t += q;
    }
  }
  __syncthreads();
  if (i < n) {
    const int64_t e = bsum[blockIdx.x] + ws[threadIdx.x >> 6] + x - v;
    off[i] = cap > 0 && e > cap ? cap : e;
  }
  if (i == 0) {
    const int64_t t = *total;
    off[n] = cap > 0 && t > cap ? cap : t;
    if (cap > 0 && t > cap) atomicOr(flag, bit);
  }
}

// Whole scan in one block (short arrays: one launch instead of three).  Each
// thread takes SCAN_PER consecutive counts per pass; passes carry the total.
constexpr int SCAN_PER = 16;
constexpr int64_t SCAN_SINGLE_MAX = 4 * SCAN_BLOCK * SCAN_PER;

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_single(const int32_t* __restrict__ c, int64_t n,
                                                            int64_t* __restrict__ off, int64_t* __restrict__ total,
                                                            int64_t cap, int32_t* __restrict__ flag, int32_t bit) {
  __shared__ int64_t ws[SCAN_BLOCK / 64];
  __shared__ int64_t pass_total;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int64_t carry = 0;
  // 16-byte loads and stores where the buffers allow (they are context
  // scratch, so in practice always)
  const bool vec = ((reinterpret_cast<uintptr_t>(c) | reinterpret_cast<uintptr_t>(off)) & 15) == 0;
  for (int64_t base = 0; base < n; base += (int64_t)SCAN_BLOCK * SCAN_PER) {
    const int64_t i0 = base + (int64_t)threadIdx.x * SCAN_PER;
    int32_t v[SCAN_PER];
    int64_t sum = 0;
    if (vec && i0 + SCAN_PER <= n) {
      const int4* c4 = reinterpret_cast<const int4*>(c + i0);
#pragma unroll
      for (int k = 0; k < SCAN_PER / 4; k++) {
        const int4 q = c4[k];
        v[4 * k] = q.x;
        v[4 * k + 1] = q.y;
        v[4 * k + 2] = q.z;
        v[4 * k + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < SCAN_PER; k++) v[k] = i0 + k < n ? c[i0 + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) sum += v[k];
    int64_t x = sum;  // inclusive wave scan of the thread sums
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) ws[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t t = 0;
      for (int w = 0; w < SCAN_BLOCK / 64; w++) {
        const int64_t q = ws[w];
        ws[w] = t;
        t += q;
      }
      pass_total = t;
    }
    __syncthreads();
    int64_t e = carry + ws[wv] + x - sum;
    if (vec && i0 + SCAN_PER <= n) {
      longlong2* o2 = reinterpret_cast<longlong2*>(off + i0);
#pragma unroll
      for (int k = 0; k < SCAN_PER / 2; k++) {
        const int64_t e0 = e, e1 = e + v[2 * k];
        e = e1 + v[2 * k + 1];
        o2[k] = make_longlong2(cap > 0 && e0 > cap ? cap : e0, cap > 0 && e1 > cap ? cap : e1);
      }
    } else {
#pragma unroll
      for (int k = 0; k < SCAN_PER; k++) {
        if (i0 + k < n) off[i0 + k] = cap > 0 && e > cap ? cap : e;
        e += v[k];
      }
    }
    carry += pass_total;
    __syncthreads();  // ws and pass_total are rewritten by the next pass
  }
  if (threadIdx.x == 0) {
    off[n] = cap > 0 && carry > cap ? cap : carry;
    *total = carry;
    if (cap > 0 && carry > cap) atomicOr(flag, bit);
  }
}

int64_t exclusive_scan(ipg_ctx* ctx, const int32_t* d_counts, int64_t n, int64_t* d_offsets, bool want_total,
                       int64_t cap, int32_t* flag, int32_t bit) {
  IPG_REQUIRE(cap <= 0 || flag != nullptr, "internal: capacity scan without an overflow flag");
  int64_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
  if (nb < 1) nb = 1;
  int64_t* bsum = scratch_t<int64_t>(ctx, "scan_bsum", nb + 2);
  int64_t* total = bsum + nb;
  if (n <= SCAN_SINGLE_MAX) {
    launch(ctx, "scan", [&] {
      k_scan_single<<<1, SCAN_BLOCK, 0, ctx->stream>>>(d_counts, n, d_offsets, total, cap, flag, bit);
    });
  } else {
    launch(ctx, "scan", [&] {
      k_scan_block_sums<<<(unsigned)nb, SCAN_BLOCK, 0, ctx->stream>>>(d_counts, n, bsum);
      k_scan_sums<<<1, SCAN_BLOCK, 0, ctx->stream>>>(bsum, nb, total);
      k_scan_final<<<(unsigned)nb, SCAN_BLOCK, 0, ctx->stream>>>(d_counts, n, bsum, total, d_offsets, cap, flag,
                                                                  bit);
    });
  }
  int64_t t = 0;
  if (want_total) {
    d2h(ctx, &t, total, sizeof(int64_t));
    sync(ctx);
  }
  return t;
}

// ---------------------------------------------------------------------------
// deterministic row reduction: out[w] (+)= sum_i rows[i*width + w]
// Two phases with fixed partition and fixed tree order.
// ---------------------------------------------------------------------------
constexpr int RED_BLOCK = 256;
constexpr int RED_MAXW = 64;

__global__ void k_reduce_rows_p1(const double* __restrict__ rows, int64_t n, int width, int64_t chunk,
                                 double* __restrict__ part) {
  __shared__ double sh[RED_BLOCK];
  int64_t lo = (int64_t)blockIdx.x * chunk;
  int64_t hi = lo + chunk < n ? lo + chunk : n;
  for (int w = 0; w < width; w++) {
    double acc = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += RED_BLOCK) acc += rows[i * width + w];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = RED_BLOCK / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * width + w] = sh[0];
    __syncthreads();
  }
}

__global__ void k_reduce_rows_p2(const double* __restrict__ part, int64_t nb, int width, double* __restrict__ out,
                                 int accumulate, double offset) {
  __shared__ double sh[RED_BLOCK];
  for (int w = 0; w < width; w++) {
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < nb; i += RED_BLOCK) acc += part[i * width + w];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = RED_BLOCK / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[w] = offset + (accumulate ? out[w] + sh[0] : sh[0]);
    __syncthreads();
  }
}

// Short inputs: the whole reduction in one block (fixed strided partition,
// fixed tree).
constexpr int RED1_BLOCK = 1024;
__global__ __launch_bounds__(RED1_BLOCK) void k_reduce_rows_single(const double* __restrict__ rows, int64_t n,
                                                                   int width, double* __restrict__ out,
                                                                   int accumulate, double offset) {
  __shared__ double sh[RED1_BLOCK];
  for (int w = 0; w < width; w++) {
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += RED1_BLOCK) acc += rows[i * width + w];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = RED1_BLOCK / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[w] = offset + (accumulate ? out[w] + sh[0] : sh[0]);
    __syncthreads();
  }
}

// Narrow rows (width <= 8) in one pass: every thread sums its strided rows
// for all columns at once, then fixed butterflies inside each wave and the
// wave partials in wave order.
template <int W>
__global__ __launch_bounds__(RED1_BLOCK) void k_reduce_rows_narrow(const double* __restrict__ rows, int64_t n,
                                                                   double* __restrict__ out, int accumulate,
                                                                   double offset) {
  __shared__ double sh[RED1_BLOCK / 64][W];
  double acc[W];
#pragma unroll
  for (int w = 0; w < W; w++) acc[w] = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += RED1_BLOCK)
#pragma unroll
    for (int w = 0; w < W; w++) acc[w] += rows[i * W + w];
#pragma unroll
  for (int w = 0; w < W; w++) {
    for (int o = 32; o > 0; o >>= 1) acc[w] += __shfl_xor(acc[w], o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][w] = acc[w];
  }
  __syncthreads();
  if ((int)threadIdx.x < W) {
    double t = 0.0;
    for (int k = 0; k < RED1_BLOCK / 64; k++) t += sh[k][threadIdx.x];
    out[threadIdx.x] = offset + (accumulate ? out[threadIdx.x] + t : t);
  }
}

// Wide rows (width > 8): a lane per column, so each wave loads 64
// consecutive doubles of one row.  Blocks take fixed chunks of rows (the 4
// waves of a block every 4th row of the chunk, combined in wave order); the
// second kernel adds the chunk partials in chunk order.
constexpr int REDW_CHUNK = 256;
__global__ __launch_bounds__(256) void k_reduce_wide_p1(const double* __restrict__ rows, int64_t n, int width,
                                                        int64_t chunk, double* __restrict__ part) {
  __shared__ double sh[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = blockIdx.x * 64 + lane;
  const int64_t lo = (int64_t)blockIdx.y * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  double acc = 0.0;
  if (col < width)
    for (int64_t i = lo + wv; i < hi; i += 4) acc += rows[i * width + col];
  sh[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && col < width)
    part[(int64_t)blockIdx.y * width + col] = ((sh[0][lane] + sh[1][lane]) + sh[2][lane]) + sh[3][lane];
}

__global__ void k_reduce_wide_p2(const double* __restrict__ part, int64_t nchunk, int width, double* __restrict__ out,
                                 int accumulate, double offset) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= width) return;
  double t = 0.0;
  for (int64_t c = 0; c < nchunk; c++) t += part[c * width + col];
  out[col] = offset + (accumulate ? out[col] + t : t);
}

void reduce_rows(ipg_ctx* ctx, const double* d_rows, int64_t n, int width, double* d_out, bool accumulate,
                 double offset) {
  if (width > 8 && n > 0) {
    int64_t chunk = REDW_CHUNK;
    if ((n + chunk - 1) / chunk > 65535) chunk = (n + 65534) / 65535;
    const int64_t nchunk = (n + chunk - 1) / chunk;
    double* part = scratch_t<double>(ctx, "reduce_wide_part", nchunk * (int64_t)width);
    launch(ctx, "reduce_rows", [&] {
      k_reduce_wide_p1<<<dim3((unsigned)((width + 63) / 64), (unsigned)nchunk), 256, 0, ctx->stream>>>(
          d_rows, n, width, chunk, part);
      k_reduce_wide_p2<<<grid_for(width, 256), 256, 0, ctx->stream>>>(part, nchunk, width, d_out,
                                                                       accumulate ? 1 : 0, offset);
    });
    return;
  }
  if (n <= 64 * RED1_BLOCK && width >= 1 && width <= 8) {
    launch(ctx, "reduce_rows", [&] {
      const int acc = accumulate ? 1 : 0;
      switch (width) {
        case 1: k_reduce_rows_narrow<1><<<1, RED1_BLOCK, 0, ctx->stream>>>(d_rows, n, d_out, acc, offset); break;
        case 2: k_reduce_rows_narrow<2><<<1, RED1_BLOCK, 0, ctx->stream>>>(d_rows, n, d_out, acc, offset); break;
        case 3: k_reduce_rows_narrow<3><<<1, RED1_BLOCK, 0, ctx->stream>>>(d_rows, n, d_out, acc, offset); break;
        case 4: k_reduce_rows_narrow<4><<<1, RED1_BLOCK, 0, ctx->stream>>>(d_rows, n, d_out, acc, offset); break;
        case 5: k_reduce_rows_narrow<5><<<1, RED1_BLOCK, 0, ctx->stream>>>(d_rows, n, d_out, acc, offset); break;
        case 6: k_reduce_rows_narrow<6><<<1, RED1_BLOCK, 0, ctx->stream>>>(d_rows, n, d_out, acc, offset); break;
        case 7: k_reduce_rows_narrow<7><<<1, RED1_BLOCK, 0, ctx->stream>>>(d_rows, n, d_out, acc, offset); break;
        default: k_reduce_rows_narrow<8><<<1, RED1_BLOCK, 0, ctx->stream>>>(d_rows, n, d_out, acc, offset); break;
      }
    });
    return;
  }
  if (n <= 64 * RED1_BLOCK) {
    launch(ctx, "reduce_rows", [&] {
      k_reduce_rows_single<<<1, RED1_BLOCK, 0, ctx->stream>>>(d_rows, n, width, d_out, accumulate ? 1 : 0, offset);
    });
    return;
  }
  int64_t nb = (n + 4095) / 4096;
  if (nb < 1) nb = 1;
  if (nb > 1024) nb = 1024;
  int64_t chunk = (n + nb - 1) / nb;
  if (chunk < 1) chunk = 1;
  double* part = scratch_t<double>(ctx, "reduce_part", nb * (int64_t)width);
  launch(ctx, "reduce_rows", [&] {
    k_reduce_rows_p1<<<(unsigned)nb, RED_BLOCK, 0, ctx->stream>>>(d_rows, n, width, chunk, part);
    k_reduce_rows_p2<<<1, RED_BLOCK, 0, ctx->stream>>>(part, nb, width, d_out, accumulate ? 1 : 0, offset);
  });
}

void check_system_shape(const ipg_system* s) {
  IPG_REQUIRE(s != nullptr, "system is NULL");
  IPG_REQUIRE(s->natoms >= 0, "natoms must be >= 0");
  IPG_REQUIRE(s->natoms < (int64_t)INT32_MAX, "natoms must be < 2^31");
  IPG_REQUIRE(s->natoms == 0 || s->positions != nullptr, "positions is NULL");
  for (int d = 0; d < 3; d++)
    if (s->periodic[d])
      IPG_REQUIRE(std::isfinite(s->box[d]) && s->box[d] > 0.0,
                  "box length of a periodic dimension must be finite and > 0");
}

void check_system_host(const ipg_system* s) {
  check_system_shape(s);
  for (int64_t i = 0; i < 3 * s->natoms; i++)
    IPG_REQUIRE(std::isfinite(s->positions[i]), "positions must be finite (atom " + std::to_string(i / 3) + ")");
  if (s->species)
    for (int64_t i = 0; i < s->natoms; i++)
      IPG_REQUIRE(s->species[i] >= 0 && s->species[i] < 64, "species index out of range [0, 64)");
}

double sqrt_le_threshold(double r) {
  double s = r * r;
  while (s > 0 && std::sqrt(s) > r) s = std::nextafter(s, 0.0);
  for (;;) {
    double t = std::nextafter(s, DBL_MAX);
    if (std::sqrt(t) <= r) s = t;
    else break;
  }
  return s;
}

}  // namespace ipg

using namespace ipg;

extern "C" {

int ipg_get_version(void) { return IPG_VERSION; }

const char* ipg_last_error(void) { return g_last_error.c_str(); }

int ipg_ctx_create(int device, ipg_ctx** out) {
  return guard([&] {
    IPG_REQUIRE(out != nullptr, "out is NULL");
    int ndev = 0;
    IPG_HIP(hipGetDeviceCount(&ndev));
    IPG_REQUIRE(device >= 0 && device < ndev, "device index out of range");
    IPG_HIP(hipSetDevice(device));
    auto* c = new ipg_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      fail(IPG_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    *out = c;
  });
}

int ipg_ctx_destroy(ipg_ctx* ctx) {
  return guard([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->scratch)
      if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    for (auto& p : ctx->pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->free_events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int ipg_ctx_synchronize(ipg_ctx* ctx) {
  return guard([&] {
    IPG_REQUIRE(ctx, "ctx is NULL");
    sync(ctx);
  });
}

int ipg_device_malloc(ipg_ctx* ctx, size_t bytes, void** dptr) {
  return guard([&] {
    IPG_REQUIRE(ctx && dptr, "NULL argument");
    IPG_HIP(hipSetDevice(ctx->device));
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
      (void)hipGetLastError();
      fail(IPG_ERR_NOMEM, "hipMalloc failed");
    }
  });
}

int ipg_device_free(ipg_ctx* ctx, void* dptr) {
  return guard([&] {
    IPG_REQUIRE(ctx, "ctx is NULL");
    sync(ctx);
    if (dptr) IPG_HIP(hipFree(dptr));
  });
}

int ipg_memcpy_h2d(ipg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    IPG_REQUIRE(ctx, "ctx is NULL");
    h2d(ctx, dst, src, bytes);
    sync(ctx);
  });
}

int ipg_memcpy_d2h(ipg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    IPG_REQUIRE(ctx, "ctx is NULL");
    d2h(ctx, dst, src, bytes);
    sync(ctx);
  });
}

int ipg_prof_enable(ipg_ctx* ctx, int on) {
  return guard([&] {
    IPG_REQUIRE(ctx, "ctx is NULL");
    flush_prof(ctx);
    ctx->prof = on != 0;
    ctx->prof_gen++;
  });
}

int ipg_prof_select(ipg_ctx* ctx, const char* names) {
  return guard([&] {
    IPG_REQUIRE(ctx, "ctx is NULL");
    flush_prof(ctx);
    ctx->prof_only.clear();
    if (names) {
      std::string all(names), cur;
      for (char ch : all + ",") {
        if (ch == ',') {
          if (!cur.empty()) ctx->prof_only.push_back(cur);
          cur.clear();
        } else if (ch != ' ') {
          cur += ch;
        }
      }
    }
    ctx->prof_gen++;
  });
}

// Plain 16-byte-vector copy (grid-stride): a kernel of known traffic, used
// to calibrate hardware byte counters against.
__global__ __launch_bounds__(256) void k_device_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int ipg_device_copy(ipg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    IPG_REQUIRE(ctx && dst && src, "NULL argument");
    IPG_REQUIRE(bytes % 16 == 0 && ((uintptr_t)dst | (uintptr_t)src) % 16 == 0,
                "bytes and both pointers must be multiples of 16");
    IPG_HIP(hipSetDevice(ctx->device));
    const int64_t n16 = (int64_t)(bytes / 16);
    if (n16 == 0) return;
    const unsigned g = (unsigned)std::min<int64_t>((n16 + 255) / 256, (int64_t)ctx->num_cus * 16);
    launch(ctx, "device_copy", [&] {
      k_device_copy<<<g, 256, 0, ctx->stream>>>((const uint4*)src, (uint4*)dst, n16);
    });
  });
}

int ipg_ctx_set_graphs(ipg_ctx* ctx, int on) {
  return guard([&] {
    IPG_REQUIRE(ctx, "ctx is NULL");
    ctx->graphs = on != 0;
  });
}

int ipg_prof_reset(ipg_ctx* ctx) {
  return guard([&] {
    IPG_REQUIRE(ctx, "ctx is NULL");
    flush_prof(ctx);
    ctx->stats.clear();
  });
}

int ipg_prof_query(ipg_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms) {
  return guard([&] {
    IPG_REQUIRE(ctx && kernel, "NULL argument");
    flush_prof(ctx);
    auto it = ctx->stats.find(kernel);
    if (launches) *launches = it == ctx->stats.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.second;
  });
}

}  // extern "C"
