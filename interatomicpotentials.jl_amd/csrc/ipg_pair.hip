// ipg_pair.hip -- fused pair energy / force / virial kernel.
//
// Semantics of src/types/empirical_potential.jl:20-67 and
// src/mixed/linear_combination_potential.jl:27-39, fused: one full list at
// the largest rcut; for every pair, each term whose species filter holds
// for both atoms (empirical_potential.jl:36-38) and whose own rcut admits
// the pair (PeriodicEuclidean distance <= rcut_k, the membership of that
// term's own list) contributes coeff_k * V_k and coeff_k * F_k.
//
// Layout: LANES lanes cooperate on one atom and stride over its neighbor
// segment (coalesced index reads); forces are gathered per atom (no
// atomics, deterministic).  For the pair {lo, hi} (lo < hi) both atoms
// evaluate r = get_displacement(x_lo, x_hi) and F(r) bit-identically; lo
// adds +F and hi adds -F, exactly the reference's f[i] += fo, f[j] -= fo.
// Energy and virial are taken from the lo side only (half-pair sums) and
// reduced in a fixed tree.
#include <mutex>

#include "ipg_nlist.h"

namespace ipg {

constexpr int MAX_TERMS = 8;

struct PairTermDev {
  int kind;
  int check;  // 1: the term's rcut is below the list's, test membership
  uint64_t mask;
  double coeff;
  double smax;  // sqrt-free membership threshold of the term's rcut
  double q[6];  // precomputed parameters, see prepare_terms()
};
struct PairTable {
  int nterms;
  int any_check;
  PairTermDev t[MAX_TERMS];
};

// Pair functions written from the closed forms of src/empirical/*.jl.
// Returns V and sets fs so that F = fs * r; ir = 1/R, formed once per pair
// and shared by the terms (the only division on the pair's path).
template <typename T>
__device__ __forceinline__ T pair_eval(const PairTermDev& t, T R, T ir, T& fs) {
  switch (t.kind) {
    case IPG_PAIR_LJ: {  // lj.jl:35-36; q = {4 eps, 24 eps, sigma}
      T s = (T)t.q[2] * ir;
      T s2 = s * s;
      T s6 = s2 * s2 * s2;
      T s12 = s6 * s6;
      fs = (T)t.q[1] * ((T)2 * s12 - s6) * (ir * ir);
      return (T)t.q[0] * (s12 - s6);
    }
    case IPG_PAIR_MORSE: {  // morse.jl:36-39 (force sign as in the reference); q = {D, alpha, sigma}
      T E = exp(-(T)t.q[1] * (R - (T)t.q[2]));
      T om = (T)1 - E;
      fs = (T)2 * (T)t.q[0] * (T)t.q[1] * om * E * ir;
      return (T)t.q[0] * (om * om);
    }
    case IPG_PAIR_ZBL: {  // zbl.jl:27-36; q = {kZZe^2, 1/a_screen, ln(e)}
      const T c[4] = {(T)0.18175, (T)0.50986, (T)0.28022, (T)0.02817};
      const T x[4] = {(T)3.19980, (T)0.94229, (T)0.40290, (T)0.20162};
      T d = R * (T)t.q[1];
      T le = (T)t.q[2];
      T phi = 0, dphi = 0;
      if (t.q[2] == 0.0) {
        // e = 1: every e^(-x_k d) is exactly 1 (the same sums as below)
#pragma unroll
        for (int k = 0; k < 4; k++) {
          phi += c[k];
          dphi += x[k] * c[k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < 4; k++) {
          T ek = exp(-x[k] * d * le);  // e^(-x_k d)
          phi += c[k] * ek;
          dphi += x[k] * c[k] * ek;
        }
      }
      dphi = -dphi;
      T cf = (T)t.q[0];
      fs = cf * (phi + d * dphi) * (ir * ir * ir);
      return cf * phi * ir;
    }
    case IPG_PAIR_BORN_MAYER: {  // bm.jl:38-39; q = {A, rho, sigma, C, D}
      T ex = (T)t.q[0] * exp(((T)t.q[2] - R) / (T)t.q[1]);
      T c6 = (T)t.q[3] * ir, d8 = (T)t.q[4] * ir;
      c6 = c6 * c6 * c6;
      c6 = c6 * c6;
      T d2 = d8 * d8;
      T d4 = d2 * d2;
      d8 = d4 * d4;
      fs = ex * ir / (T)t.q[1] + (c6 - d8) * (ir * ir);
      return ex + c6 - d8;
    }
    case IPG_PAIR_COULOMB: {  // coulomb.jl:36-37; q = {k q1 q2}
      fs = (T)t.q[0] * (ir * ir * ir);
      return (T)t.q[0] * ir;
    }
    case IPG_PAIR_MOCK: {  // test/mocks.jl:28-33; q = {seed}
      T v = R * R * (T)t.q[0];
      fs = v;
      return v;
    }
  }
  fs = 0;
  return 0;
}

template <typename T>
__device__ __forceinline__ T wave_sum_lanes(T v, int lanes) {
  for (int o = 1; o < lanes; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One block: 256 threads; LANES lanes per atom.  ROWS: the neighbors of
// atom a are rows[a*K .. a*K + cnt[a]) (the list's rows before compaction,
// in the CSR order), else nbr[off[a] .. off[a+1]).
// Occupancy: the FP32 build at 6 waves per SIMD (at most 80 VGPRs) measured
// 11% faster than its natural 5; the FP64 build spills at any bound above
// its natural 4 waves.
template <typename T, int LANES, bool ROWS>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 6 : 1) void k_pair(int64_t n, const int64_t* __restrict__ off,
                                              const int32_t* __restrict__ nbr, const int32_t* __restrict__ cnt,
                                              int K, const double* __restrict__ pos,
                                              const int32_t* __restrict__ species, IpgBox box, PairTable tab,
                                              double* __restrict__ forces, double* __restrict__ part,
                                              int32_t* __restrict__ flag) {
  const int tid = threadIdx.x;
  const int64_t a = ((int64_t)blockIdx.x * 256 + tid) / LANES;
  const int lane = tid % LANES;
  double e = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0, v4 = 0.0, v5 = 0.0;
  T fx = 0, fy = 0, fz = 0;
  if (a < n) {
    const double xa = pos[3 * a], ya = pos[3 * a + 1], za = pos[3 * a + 2];
    const int sa = species ? species[a] : 0;
    // a species index outside [0, 64) belongs to no term's species set
    // (it contributes nothing) and raises bit 16 for the caller
    if ((unsigned)sa >= 64u) atomicOr(flag, 16);
    const int64_t q0 = ROWS ? a * K : off[a], q1 = ROWS ? a * K + cnt[a] : off[a + 1];
    for (int64_t q = q0 + lane; q < q1; q += LANES) {
      const int32_t j = nbr[q];
      const double xj = pos[3 * j], yj = pos[3 * j + 1], zj = pos[3 * j + 2];
      const int sj = species ? species[j] : 0;
      const bool lo = a < j;
      // r = get_displacement(x_lo, x_hi): operands picked by selects, so
      // lanes of both orders share one evaluation
      const double rx = ipg_disp1(lo ? xa : xj, lo ? xj : xa, box.L[0], box.per[0]);
      const double ry = ipg_disp1(lo ? ya : yj, lo ? yj : ya, box.L[1], box.per[1]);
      const double rz = ipg_disp1(lo ? za : zj, lo ? zj : za, box.L[2], box.per[2]);
      double spe = 0.0;
      if (tab.any_check) spe = ipg_pe_sumsq(box, xa, ya, za, xj, yj, zj);
      const T trx = (T)rx, try_ = (T)ry, trz = (T)rz;
      const T R = sqrt(trx * trx + try_ * try_ + trz * trz);
      const T ir = (T)1 / R;
      T vsum = 0, fsum = 0;
      for (int k = 0; k < tab.nterms; k++) {
        const PairTermDev& t = tab.t[k];
        if ((unsigned)sa >= 64u || (unsigned)sj >= 64u || !((t.mask >> sa) & 1ull) || !((t.mask >> sj) & 1ull))
          continue;
        if (t.check && !(spe <= t.smax)) continue;
        T fs;
        T vv = pair_eval<T>(t, R, ir, fs);
        vsum += (T)t.coeff * vv;
        fsum += (T)t.coeff * fs;
      }
      const T sgn = lo ? (T)1 : (T)-1;
      fx += sgn * fsum * trx;
      fy += sgn * fsum * try_;
      fz += sgn * fsum * trz;
      if (lo) {
        const double fd = (double)fsum;
        e += (double)vsum;
        const double Fx = fd * rx, Fy = fd * ry, Fz = fd * rz;
        v0 += rx * Fx;
        v1 += ry * Fy;
        v2 += rz * Fz;
        v3 += rz * Fy;
        v4 += rz * Fx;
        v5 += ry * Fx;
      }
    }
  }
  // per-atom force: sum over the LANES lanes (in double)
  double dfx = wave_sum_lanes<double>((double)fx, LANES);
  double dfy = wave_sum_lanes<double>((double)fy, LANES);
  double dfz = wave_sum_lanes<double>((double)fz, LANES);
  if (a < n && lane == 0 && forces) {
    forces[3 * a] = dfx;
    forces[3 * a + 1] = dfy;
    forces[3 * a + 2] = dfz;
  }
  // block partial of energy and virial (fixed order)
  __shared__ double sh[7][4];
  double vals[7] = {e, v0, v1, v2, v3, v4, v5};
#pragma unroll
  for (int w = 0; w < 7; w++) {
    double x = vals[w];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((tid & 63) == 0) sh[w][tid >> 6] = x;
  }
  __syncthreads();
  if (tid < 7) part[(int64_t)blockIdx.x * 7 + tid] = ((sh[tid][0] + sh[tid][1]) + sh[tid][2]) + sh[tid][3];
}

__global__ void k_split_ev(const double* __restrict__ ev, double* __restrict__ energy, double* __restrict__ virial) {
  int t = threadIdx.x;
  if (t == 0 && energy) energy[0] = ev[0];
  if (t >= 1 && t < 7 && virial) virial[t - 1] = ev[t];
}

static PairTable prepare_terms(const ipg_pair_term* terms, int nterms, double list_rcut) {
  IPG_REQUIRE(nterms >= 1 && nterms <= MAX_TERMS, "nterms must be in [1, 8]");
  PairTable tab{};
  tab.nterms = nterms;
  double list_smax = sqrt_le_threshold(list_rcut);
  for (int k = 0; k < nterms; k++) {
    const ipg_pair_term& s = terms[k];
    PairTermDev& t = tab.t[k];
    IPG_REQUIRE(std::isfinite(s.rcut) && s.rcut > 0, "term rcutoff must be finite and > 0");
    IPG_REQUIRE(std::isfinite(s.coeff), "term coefficient must be finite");
    t.kind = s.kind;
    t.coeff = s.coeff;
    t.mask = s.species_mask;
    t.smax = sqrt_le_threshold(s.rcut);
    t.check = t.smax < list_smax ? 1 : 0;
    tab.any_check |= t.check;
    const double* p = s.p;
    switch (s.kind) {
      case IPG_PAIR_LJ:
        t.q[0] = 4.0 * p[0];
        t.q[1] = 24.0 * p[0];
        t.q[2] = p[1];
        break;
      case IPG_PAIR_MORSE:
        t.q[0] = p[0];
        t.q[1] = p[1];
        t.q[2] = p[2];
        break;
      case IPG_PAIR_ZBL: {
        IPG_REQUIRE(p[2] > 0, "ZBL charge parameter e must be > 0");
        t.q[0] = 1.0 * p[0] * p[1] * (p[2] * p[2]);
        t.q[1] = (std::pow(p[0], 0.23) + std::pow(p[1], 0.23)) / (0.8854 * 0.529);
        t.q[2] = std::log(p[2]);
        break;
      }
      case IPG_PAIR_BORN_MAYER:
        for (int i = 0; i < 5; i++) t.q[i] = p[i];
        break;
      case IPG_PAIR_COULOMB:
        t.q[0] = 1.0 * p[0] * p[1];
        break;
      case IPG_PAIR_MOCK:
        t.q[0] = p[0];
        break;
      default:
        fail(IPG_ERR_INVALID, "unknown pair term kind " + std::to_string(s.kind));
    }
  }
  return tab;
}

static double max_rcut(const ipg_pair_term* terms, int nterms) {
  IPG_REQUIRE(terms != nullptr, "terms is NULL");
  IPG_REQUIRE(nterms >= 1 && nterms <= MAX_TERMS, "nterms must be in [1, 8]");
  double r = 0;
  for (int k = 0; k < nterms; k++) {
    IPG_REQUIRE(std::isfinite(terms[k].rcut) && terms[k].rcut > 0, "term rcutoff must be finite and > 0");
    r = std::max(r, terms[k].rcut);
  }
  return r;
}

static int pick_lanes(int64_t n, int64_t npairs) {
  double avg = n > 0 ? (double)npairs / (double)n : 0.0;
  if (avg >= 48) return 16;
  if (avg >= 12) return 8;
  return 4;
}

// Blocks of the pair kernel over a list of npairs (7 partial sums each).
static int64_t pair_blocks(int64_t n, int64_t npairs) {
  const int64_t nblk = (n * pick_lanes(n, npairs) + 255) / 256;
  return nblk < 1 ? 1 : nblk;
}

// Launch the pair kernel over a full list; e/v to d_ev[7] (device).  part
// (pair_blocks() rows of 7 doubles) holds the per-block partial sums; null
// takes it from the context's scratch.
static void run_pair(ipg_ctx* ctx, const DevSystem& s, const DevList& L, const PairTable& tab, int precision,
                     double* d_forces, double* d_ev, int32_t* flag, double* part = nullptr) {
  int lanes = pick_lanes(s.n, L.npairs);
  const int64_t nblk = pair_blocks(s.n, L.npairs);
  if (!part) part = scratch_t<double>(ctx, "pair_part", nblk * 7);
  auto go = [&](auto tag, auto lanes_c) {
    using T = decltype(tag);
    constexpr int LN = decltype(lanes_c)::value;
    const char* nm = std::is_same<T, double>::value ? "pair_fp64" : "pair_fp32";
    launch(ctx, nm, [&] {
      if (L.ellK > 0)
        k_pair<T, LN, true><<<(unsigned)nblk, 256, 0, ctx->stream>>>(s.n, nullptr, L.ell, L.cnt, L.ellK, s.pos,
                                                                     s.species, s.box, tab, d_forces, part, flag);
      else
        k_pair<T, LN, false><<<(unsigned)nblk, 256, 0, ctx->stream>>>(s.n, L.off, L.nbr, nullptr, 0, s.pos,
                                                                      s.species, s.box, tab, d_forces, part, flag);
    });
  };
  if (precision == IPG_FP64) {
    if (lanes == 16) go(0.0, std::integral_constant<int, 16>());
    else if (lanes == 8) go(0.0, std::integral_constant<int, 8>());
    else go(0.0, std::integral_constant<int, 4>());
  } else {
    if (lanes == 16) go(0.0f, std::integral_constant<int, 16>());
    else if (lanes == 8) go(0.0f, std::integral_constant<int, 8>());
    else go(0.0f, std::integral_constant<int, 4>());
  }
  reduce_rows(ctx, part, nblk, 7, d_ev, false);
}

// Without a prebuilt list the call scans into neighbor rows of the width
// estimated from the density and uses them at once; bit 8 of the returned
// flag word says a row outgrew that width, and the call must then be
// repeated with counted = true.  Bit 16 says a species index was outside
// [0, 64) (such an atom took part in no term).  The flag is read by the
// caller, together with its results (one synchronization).
static int32_t* pair_dev(ipg_ctx* ctx, const DevSystem& s, const ipg_pair_term* terms, int nterms, int precision,
                         double* d_energy, double* d_forces, double* d_virial, const DevList* prebuilt,
                         double prebuilt_rcut, bool counted = false) {
  IPG_REQUIRE(precision == IPG_FP64 || precision == IPG_FP32, "precision must be IPG_FP64 or IPG_FP32");
  double rc = max_rcut(terms, nterms);
  DevList L;
  double list_rcut = prebuilt ? prebuilt_rcut : rc;
  int32_t* flag = scratch_t<int32_t>(ctx, "pair_list_flag", 1);
  zero(ctx, flag, sizeof(int32_t));
  if (prebuilt) {
    L = *prebuilt;
    IPG_REQUIRE(L.n == s.n, "list natoms does not match the system");
  } else {
    L = build_list(ctx, s, rc, KIND_FULL, false, false, "pair_list", 0, flag, counted ? 0 : -1, !counted);
    if (L.deferred) {
      PairTable tab = prepare_terms(terms, nterms, list_rcut);
      double* ev = scratch_t<double>(ctx, "pair_ev", 8);
      double* f = d_forces ? d_forces : scratch_t<double>(ctx, "pair_fscratch", 3 * (s.n > 0 ? s.n : 1));
      run_pair(ctx, s, L, tab, precision, f, ev, flag);
      launch(ctx, "split_ev", [&] { k_split_ev<<<1, 64, 0, ctx->stream>>>(ev, d_energy, d_virial); });
      return flag;
    }
  }
  // terms whose rcut is below the list's test their own membership
  PairTable tab = prepare_terms(terms, nterms, list_rcut);
  double* ev = scratch_t<double>(ctx, "pair_ev", 8);
  double* f = d_forces ? d_forces : scratch_t<double>(ctx, "pair_fscratch", 3 * (s.n > 0 ? s.n : 1));
  if (s.n == 0) {
    zero(ctx, ev, sizeof(double) * 8);
  } else {
    run_pair(ctx, s, L, tab, precision, f, ev, flag);
  }
  launch(ctx, "split_ev", [&] { k_split_ev<<<1, 64, 0, ctx->stream>>>(ev, d_energy, d_virial); });
  return flag;
}

static void check_pair_flag(int32_t over) {
  if (over & 16) fail(IPG_ERR_INVALID, "species index out of range [0, 64)");
}

// ---------------------------------------------------------------------------
// allocation-free evaluation over a prebuilt list
// ---------------------------------------------------------------------------
// The device memory of ipg_pair_eval_noalloc_dev beyond its inputs and
// outputs: the status word (bytes 0..3), the energy/virial total (7
// doubles at byte 256) and the per-block partial sums (7 doubles per block
// of the pair kernel, from byte 512).  Every GPU holds its own copy for the
// life of the process; all calls on one GPU share it.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)1 << 20;
constexpr size_t NOALLOC_PART_OFFSET = 512;

}  // namespace ipg

// at global scope: inside namespace ipg the name scratch is the context's
// scratch-buffer function
static __device__ __align__(256) unsigned char scratch[ipg::NOALLOC_SCRATCH_BYTES];

namespace ipg {

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES];

// Device address of this GPU's copy of the static scratch, looked up once
// per device on its first call (the device must be current).  Host code
// uses this address, never the array's name.
static char* noalloc_scratch(int device) {
  IPG_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, "device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    IPG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    IPG_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the static scratch");
    g_noalloc_ptr[device] = p;
  }
  return static_cast<char*>(g_noalloc_ptr[device]);
}

// Largest pair-kernel block count the static scratch holds.
constexpr int64_t NOALLOC_MAX_BLOCKS = (int64_t)((NOALLOC_SCRATCH_BYTES - NOALLOC_PART_OFFSET) / (7 * sizeof(double)));
// reduce_rows sums up to 65536 rows of width 7 in one block, with no
// buffer of its own
static_assert(NOALLOC_MAX_BLOCKS <= 65536, "partial-sum rows beyond the single-block reduction");

// Profiling events are created on demand, so an allocation-free call
// records none; the context's setting is restored on every exit.
struct ProfOff {
  ipg_ctx* ctx;
  bool was;
  explicit ProfOff(ipg_ctx* c) : ctx(c), was(c->prof) { c->prof = false; }
  ~ProfOff() { ctx->prof = was; }
};

// Device-buffer form: the flag is read here (one synchronization).
static void pair_dev_checked(ipg_ctx* ctx, const DevSystem& s, const ipg_pair_term* terms, int nterms, int precision,
                             double* d_energy, double* d_forces, double* d_virial) {
  int32_t* flag = pair_dev(ctx, s, terms, nterms, precision, d_energy, d_forces, d_virial, nullptr, 0.0);
  int32_t over = 0;
  d2h(ctx, &over, flag, sizeof(int32_t));
  sync(ctx);
  check_pair_flag(over);
  if (over & 8) pair_dev(ctx, s, terms, nterms, precision, d_energy, d_forces, d_virial, nullptr, 0.0, true);
}

}  // namespace ipg

using namespace ipg;

extern "C" {

int ipg_pair_energy_forces(ipg_ctx* ctx, const ipg_system* sys, const ipg_pair_term* terms, int nterms,
                           int precision, double* energy, double* forces, double* virial) {
  return guard([&] {
    IPG_REQUIRE(ctx && sys, "NULL argument");
    IPG_HIP(hipSetDevice(ctx->device));
    DevSystem s = upload_system(ctx, sys, "pair_api");
    double* dout = scratch_t<double>(ctx, "pair_api_out", 8 + 3 * (s.n > 0 ? s.n : 1));
    double* dev_e = dout;
    double* dev_v = dout + 1;
    double* dev_f = dout + 8;
    // results and the row-overflow flag come back in one synchronization;
    // an overflow repeats the call with a counted list
    double ev[8];
    int32_t over = 0;
    for (int pass = 0; pass < 2; pass++) {
      int32_t* flag = pair_dev(ctx, s, terms, nterms, precision, dev_e, forces ? dev_f : nullptr, dev_v, nullptr, 0.0,
                               pass == 1);
      d2h(ctx, ev, dout, sizeof(double) * 8);
      if (forces) d2h(ctx, forces, dev_f, sizeof(double) * 3 * s.n);
      over = 0;
      d2h(ctx, &over, flag, sizeof(int32_t));
      sync(ctx);
      check_pair_flag(over);
      if ((over & 8) == 0) break;
    }
    if (energy) *energy = ev[0];
    if (virial)
      for (int k = 0; k < 6; k++) virial[k] = ev[1 + k];
  });
}

int ipg_pair_energy_forces_dev(ipg_ctx* ctx, const ipg_system* sys, const ipg_pair_term* terms, int nterms,
                               int precision, double* energy, double* forces, double* virial) {
  return guard([&] {
    IPG_REQUIRE(ctx && sys, "NULL argument");
    IPG_HIP(hipSetDevice(ctx->device));
    DevSystem s = make_dev_system(sys);
    pair_dev_checked(ctx, s, terms, nterms, precision, energy, forces, virial);
  });
}

int ipg_pair_eval_noalloc_dev(ipg_ctx* ctx, const ipg_nlist* full_list, const ipg_system* sys,
                              const ipg_pair_term* terms, int nterms, int precision, double* energy, double* forces,
                              double* virial) {
  return guard([&] {
    IPG_REQUIRE(ctx && sys && full_list, "NULL argument");
    IPG_REQUIRE(full_list->kind == IPG_LIST_FULL, "ipg_pair_eval_noalloc_dev needs an IPG_LIST_FULL list");
    IPG_REQUIRE(precision == IPG_FP64 || precision == IPG_FP32, "precision must be IPG_FP64 or IPG_FP32");
    IPG_REQUIRE(max_rcut(terms, nterms) <= full_list->cut, "list cutoff is below a term's rcut");
    IPG_HIP(hipSetDevice(ctx->device));
    DevSystem s = make_dev_system(sys);
    const DevList& L = full_list->list;
    IPG_REQUIRE(L.n == s.n, "list natoms does not match the system");
    const int64_t nblk = pair_blocks(s.n, L.npairs);
    if (nblk > NOALLOC_MAX_BLOCKS)
      fail(IPG_ERR_INVALID, "system too large for the allocation-free pair evaluation: " + std::to_string(nblk) +
                                " pair-kernel blocks need " +
                                std::to_string(NOALLOC_PART_OFFSET + (size_t)nblk * 7 * sizeof(double)) +
                                " bytes of its static scratch of " + std::to_string(NOALLOC_SCRATCH_BYTES));
    PairTable tab = prepare_terms(terms, nterms, full_list->rcut);
    char* base = noalloc_scratch(ctx->device);
    int32_t* flag = reinterpret_cast<int32_t*>(base);
    double* ev = reinterpret_cast<double*>(base + 256);
    double* part = reinterpret_cast<double*>(base + NOALLOC_PART_OFFSET);
    ProfOff quiet(ctx);
    if (s.n == 0) {
      zero(ctx, ev, sizeof(double) * 8);
    } else {
      zero(ctx, flag, sizeof(int32_t));  // a fill kernel: the previous call's status stays in the array
      run_pair(ctx, s, L, tab, precision, forces, ev, flag, part);
    }
    launch(ctx, "split_ev", [&] { k_split_ev<<<1, 64, 0, ctx->stream>>>(ev, energy, virial); });
  });
}

int ipg_pair_eval_dev(ipg_ctx* ctx, const ipg_nlist* full_list, const ipg_system* sys, const ipg_pair_term* terms,
                      int nterms, int precision, double* energy, double* forces, double* virial) {
  return guard([&] {
    IPG_REQUIRE(ctx && sys && full_list, "NULL argument");
    IPG_REQUIRE(full_list->kind == IPG_LIST_FULL, "ipg_pair_eval_dev needs an IPG_LIST_FULL list");
    IPG_REQUIRE(max_rcut(terms, nterms) <= full_list->cut, "list cutoff is below a term's rcut");
    IPG_HIP(hipSetDevice(ctx->device));
    DevSystem s = make_dev_system(sys);
    int32_t* flag =
        pair_dev(ctx, s, terms, nterms, precision, energy, forces, virial, &full_list->list, full_list->rcut);
    int32_t over = 0;
    d2h(ctx, &over, flag, sizeof(int32_t));
    sync(ctx);
    check_pair_flag(over);
  });
}

}  // extern "C"
