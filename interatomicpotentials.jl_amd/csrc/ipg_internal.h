// ipg_internal.h -- shared internals of libipgpu.so (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "ipgpu.h"

namespace ipg {

// ---------------------------------------------------------------------------
// errors: internal code throws ipg::Error; every C entry point catches it at
// the boundary and turns it into a status code + thread-local message.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string& msg);

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

#define IPG_HIP(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      ::ipg::fail(IPG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

#define IPG_REQUIRE(cond, msg)                              \
  do {                                                      \
    if (!(cond)) ::ipg::fail(IPG_ERR_INVALID, (msg));       \
  } while (0)

template <typename F>
int guard(F&& f) {
  try {
    f();
    return IPG_OK;
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return IPG_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return IPG_ERR_INVALID;
  } catch (...) {
    set_error("unknown error");
    return IPG_ERR_INVALID;
  }
}

}  // namespace ipg

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct ipg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
  };
  std::map<std::string, Buf> scratch;  // named device buffers, grown on demand
  bool prof = false;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
  std::map<std::string, std::pair<int64_t, double>> stats;
  int num_cus = 256;
  uint64_t scratch_gen = 0;  // bumped whenever a scratch buffer is (re)allocated
  bool graphs = true;        // replay captured HIP graphs where a path supports it
  bool dry = false;          // allocation-only pass: launches are skipped
  std::vector<Pending>* capture_events = nullptr;  // event pairs of a graph being captured
  std::vector<std::string> prof_only;  // kernels timed when profiling (empty: all)
  uint64_t prof_gen = 0;               // bumped when the profiling setup changes
};

namespace ipg {

// Device buffer owned by the context, keyed by name, grown (never shrunk).
void* scratch(ipg_ctx* ctx, const char* name, size_t bytes);
template <typename T>
T* scratch_t(ipg_ctx* ctx, const char* name, size_t count) {
  return static_cast<T*>(scratch(ctx, name, count * sizeof(T) + 16));
}

void flush_prof(ipg_ctx* ctx);
hipEvent_t get_event(ipg_ctx* ctx);

// Launch wrapper: records HIP events around the launch on the context
// stream when profiling is on, and checks the launch status.
// Under stream capture the events become event-record nodes of the graph
// and are kept by the graph's owner (ctx->capture_events).  In a dry pass
// nothing is launched (only the scratch allocations of the caller happen).
inline void record_event(ipg_ctx* ctx, hipEvent_t e) {
  if (ctx->capture_events) IPG_HIP(hipEventRecordWithFlags(e, ctx->stream, hipEventRecordExternal));
  else IPG_HIP(hipEventRecord(e, ctx->stream));
}

inline bool prof_selected(const ipg_ctx* ctx, const char* name) {
  if (ctx->prof_only.empty()) return true;
  for (const std::string& k : ctx->prof_only)
    if (k == name) return true;
  return false;
}

template <typename L>
void launch(ipg_ctx* ctx, const char* name, L&& l) {
  if (ctx->dry) return;
  hipEvent_t a = nullptr, b = nullptr;
  const bool timed = ctx->prof && prof_selected(ctx, name);
  if (timed) {
    a = get_event(ctx);
    b = get_event(ctx);
    record_event(ctx, a);
  }
  l();
  IPG_HIP(hipGetLastError());
  if (timed) {
    record_event(ctx, b);
    (ctx->capture_events ? *ctx->capture_events : ctx->pending).push_back({name, a, b});
  }
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

void h2d(ipg_ctx* ctx, void* dst, const void* src, size_t bytes);
void d2h(ipg_ctx* ctx, void* dst, const void* src, size_t bytes);
void sync(ipg_ctx* ctx);
void zero(ipg_ctx* ctx, void* dst, size_t bytes);
void zero2(ipg_ctx* ctx, void* dst, size_t bytes, void* dst2, size_t bytes2);

// Add the elapsed times of event pairs recorded by a graph replay to the
// statistics (the events stay owned by the graph).
void add_graph_prof(ipg_ctx* ctx, const std::vector<ipg_ctx::Pending>& ev);

// exclusive scan of n int32 counts into n+1 int64 offsets (device), returns
// total (synchronizes when want_total).
// cap > 0: offsets are clamped to cap and bit is or-ed into *flag when the
// total exceeds it (capacity mode, no readback).
int64_t exclusive_scan(ipg_ctx* ctx, const int32_t* d_counts, int64_t n, int64_t* d_offsets, bool want_total,
                       int64_t cap = 0, int32_t* flag = nullptr, int32_t bit = 0);

// Deterministic sum of n rows of `width` doubles (row-major) into out[width].
// out[w] = offset + sum (+ out[w] when accumulate).
void reduce_rows(ipg_ctx* ctx, const double* d_rows, int64_t n, int width, double* d_out, bool accumulate,
                 double offset = 0.0);

// Host-side validation of a system description (host pointers).
void check_system_host(const ipg_system* s);
void check_system_shape(const ipg_system* s);

// Largest double s with sqrt(s) <= r (correctly rounded host sqrt), so that
// "s <= smax" on a squared distance equals "sqrt(s) <= r".
double sqrt_le_threshold(double r);

}  // namespace ipg

// ---------------------------------------------------------------------------
// device helpers: exact restatements of the reference distance semantics.
// FP contraction is disabled inside them so that every product and sum
// rounds on its own, as Julia evaluates them.
// ---------------------------------------------------------------------------
struct IpgBox {
  double L[3];
  int per[3];
};

// Julia Base.mod(x, L) for L > 0: r = rem(x, L); r == 0 -> +0; r < 0 -> r + L.
__device__ __forceinline__ double ipg_jl_mod(double x, double L) {
#pragma clang fp contract(off)
  double r = (fabs(x) < L) ? x : fmod(x, L);  // fmod(x, L) == x exactly when |x| < L
  if (r == 0.0) return 0.0;
  if (r < 0.0) r = r + L;
  return r;
}

// Distances.PeriodicEuclidean squared sum, ((0 + sx^2) + sy^2) + sz^2.
__device__ __forceinline__ double ipg_pe_sumsq(const IpgBox& b, double ax, double ay, double az, double bx,
                                               double by, double bz) {
#pragma clang fp contract(off)
  double a[3] = {ax, ay, az}, c[3] = {bx, by, bz};
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double s1 = fabs(a[d] - c[d]);
    double s3 = s1;
    if (b.per[d]) {
      double s2 = (s1 < b.L[d]) ? s1 : fmod(s1, b.L[d]);
      double s2c = b.L[d] - s2;
      s3 = (s2 < s2c) ? s2 : s2c;
    }
    s = s + s3 * s3;
  }
  return s;
}

// get_displacement(L, x, y) (src/nnlist.jl:18-21), one component.
__device__ __forceinline__ double ipg_disp1(double x, double y, double L, int per) {
#pragma clang fp contract(off)
  double diff = x - y;
  if (!per) return diff;
  double d = ipg_jl_mod(diff, L);
  return (2.0 * d < L) ? d : d - L;
}
