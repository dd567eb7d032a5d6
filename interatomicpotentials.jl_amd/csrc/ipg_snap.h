// ipg_snap.h -- SNAP tables, device views and the forward-mode Dual scalar.
#pragma once
#include <array>
#include <functional>
#include <type_traits>

#include "ipg_nlist.h"

namespace ipg {

constexpr int SNAP_JMAX = 8;       // largest twojmax with a compiled kernel set
constexpr int SNAP_JMAX_ALL = 10;  // largest twojmax of the runtime-table kernels (ipg_snap_gen.hip): from 2J = 11
                                   // on the reference's Int64 factorial products in init_clebsch_gordan overflow

// One Z entry (types/prebuilt_arrays.jl:6-17 SNA_ZINDICES) plus the
// Clebsch-Gordan block offset and the Z-block id.
struct ZEnt {
  int j1, j2, j, ma1min, ma2max, mb1min, mb2max, na, nb, jju, cgoff, blk;
};
// Contributions of one Z block (j1,j2,j) to Y_j: coef = sum_k beta[jjb[k]] * fac[k].
struct YBlk {
  int n;
  int jjb[3];
  double fac[3];
};
// The three half-sums of dB_jjb/dr (compute_b.jl:145-337): level, Z block
// start, factor.
struct BTerm {
  int l[3];
  int jjz0[3];
  double fac[3];
  int j;
};

struct SnapHost {
  int J = 0, nspecies = 1;
  int idxu_max = 0, idxz_max = 0, idxb_max = 0, idxcg_max = 0, nblk = 0, nhalf = 0;
  int ncoeff = 0;
  bool bnorm = false, bzero = false, switch_flag = false, wselfall = false;
  double rcutfac = 0, rmin0 = 0, rfac0 = 0, wself = 1.0;
  std::vector<double> radii, weight;
  std::vector<int> idxu_block;                 // J+1
  std::vector<int> idxz_block;                 // (J+1)^3
  std::vector<int> idxcg_block;                // (J+1)^3
  std::vector<int> idxb_block;                 // (J+1)^3
  std::vector<std::array<int, 3>> idxb;        // idxb_max
  std::vector<ZEnt> zent;                      // idxz_max
  std::vector<YBlk> yblk;                      // nblk
  std::vector<BTerm> bterm;                    // idxb_max
  std::vector<double> cg;                      // idxcg_max
  std::vector<double> bzero_j;                 // J+1
  std::vector<int> half_jju;                   // nhalf left-half positions (mb <= j/2)
  std::vector<int> yoff;                       // nhalf+1
  std::vector<int> yjjz;                       // jjz list per left-half position
  std::vector<int> yjblk;                      // Z block of each yjjz entry
  std::vector<uint32_t> zplan;                 // compact Z loop plan, 3 words per entry
  // runtime-table path (chem_flag = true, or twojmax > SNAP_JMAX)
  bool generic = false, chem = false;
  int nel = 1;                                 // prebuilt num_elements: nspecies with chem_flag, else 1
  std::vector<double> rootpq;                  // (J+1)^2, [p][q] = sqrt(p/q) (prebuilt_arrays.jl:314-321)
  std::vector<int> jmm;                        // idxu_max: j | mb << 8 | ma << 16 of each U entry
};

void snap_build_host(SnapHost& h, const ipg_snap_params* p);

// Device copy of the tables (one allocation).
struct SnapDev {
  int J, nspecies, idxu_max, idxz_max, idxb_max, ncoeff, nblk, nhalf;
  int bnorm, bzero, switch_flag;
  double rcutfac, rmin0, rfac0, wself;
  double radii[8], weight[8];
  double bzero_j[SNAP_JMAX_ALL + 1];
  int idxu_block[SNAP_JMAX_ALL + 2];
  int hoff[SNAP_JMAX_ALL + 2];  // offsets of each level's left-half rows (mb <= j/2)
  const ZEnt* zent;
  const YBlk* yblk;
  const BTerm* bterm;
  const double* cg_d;
  const float* cg_f;
  const int* half_jju;
  const int* yoff;
  const int* yjjz;
  const int* yjblk;        // Z block of each yjjz entry
  const uint32_t* zplan;   // 3 words per Z entry, see snap_build_host
  int idxcg_max;
  // runtime-table path
  int generic, chem, nel, wselfall;
  const double* rootpq;    // (J+1)^2
  const int* jmm;          // idxu_max
  const int* idxb3;        // idxb_max x 3 (j1, j2, j)
  const int* idxz_block;   // (J+1)^3
};

// Per-call device data of one SNAP evaluation.
struct SnapRun {
  int64_t n = 0, npairs = 0;
  const int64_t* off = nullptr;   // SNAP list offsets (n+1)
  const int32_t* nbr = nullptr;   // neighbor of each pair
  const int32_t* center = nullptr;
  const double* rvec = nullptr;   // x_j - x_i per pair (3)
  const int32_t* species = nullptr;
  const int64_t* roff = nullptr;  // reverse list
  const int32_t* rpair = nullptr;
  const double* tang = nullptr;   // per-atom tangent direction (3n) for Dual runs
  const int32_t* rev = nullptr;   // reverse pair of each pair (symmetric CUTOFF list) or null
  const int64_t* d_np = nullptr;  // device pair count when npairs is a capacity (else null)
  double avg_nbr = -1.0;          // mean neighbors per atom for launch heuristics (<0: npairs/n)
  int apw = 0;                    // U kernel atoms per wave (1 or 2; 0: from avg_nbr)
  int dbdr_threads = 0;           // dB kernel block size (64 or 128; 0: 128)
  int32_t* flag = nullptr;        // call's flag word: the U kernel raises bit 16 for a species index
                                  // outside [0, nspecies)
};

// Species index of atom i clamped into [0, ns).  Device entry points take
// species arrays unchecked, so no kernel lets a bad index address a table;
// the U kernel reports it (bit 16 of the call's flag word) and the call
// fails.
__device__ __forceinline__ int snap_species(const int32_t* sp, int64_t i, int ns) {
  const int s = sp ? sp[i] : 0;
  return (unsigned)s < (unsigned)ns ? s : 0;
}

// ---------------------------------------------------------------------------
// Dual numbers: value + tangent, for exact forward-mode directional
// derivatives through the SNAP kernels.
// ---------------------------------------------------------------------------
// keep the float/double overloads visible next to the Dual ones below
using ::cos;
using ::sin;
using ::sqrt;

template <typename T>
struct Dual {
  T v, t;
  Dual() = default;  // trivial, so Dual can live in __shared__ arrays
  __host__ __device__ Dual(T a) : v(a), t(0) {}
  __host__ __device__ Dual(T a, T b) : v(a), t(b) {}
  template <typename U, typename = typename std::enable_if<std::is_arithmetic<U>::value && !std::is_same<U, T>::value>::type>
  __host__ __device__ explicit Dual(U a) : v((T)a), t(0) {}
  template <typename U, typename = typename std::enable_if<!std::is_same<U, T>::value>::type>
  __host__ __device__ explicit Dual(Dual<U> a) : v((T)a.v), t((T)a.t) {}
};
template <typename T> __device__ __forceinline__ Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.t + b.t}; }
template <typename T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.t - b.t}; }
template <typename T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a) { return {-a.v, -a.t}; }
template <typename T> __device__ __forceinline__ Dual<T> operator*(Dual<T> a, Dual<T> b) { return {a.v * b.v, a.v * b.t + a.t * b.v}; }
template <typename T> __device__ __forceinline__ Dual<T> operator*(Dual<T> a, T b) { return {a.v * b, a.t * b}; }
template <typename T> __device__ __forceinline__ Dual<T> operator*(T b, Dual<T> a) { return {a.v * b, a.t * b}; }
template <typename T> __device__ __forceinline__ Dual<T> operator/(Dual<T> a, Dual<T> b) {
  T iv = (T)1 / b.v;
  T q = a.v * iv;
  return {q, (a.t - q * b.t) * iv};
}
template <typename T> __device__ __forceinline__ Dual<T> operator/(Dual<T> a, T b) { return {a.v / b, a.t / b}; }
template <typename T> __device__ __forceinline__ Dual<T> operator/(T a, Dual<T> b) {
  T q = a / b.v;
  return {q, -q * b.t / b.v};
}
template <typename T> __device__ __forceinline__ Dual<T> operator+(Dual<T> a, T b) { return {a.v + b, a.t}; }
template <typename T> __device__ __forceinline__ Dual<T> operator+(T b, Dual<T> a) { return {a.v + b, a.t}; }
template <typename T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a, T b) { return {a.v - b, a.t}; }
template <typename T> __device__ __forceinline__ Dual<T> operator-(T b, Dual<T> a) { return {b - a.v, -a.t}; }
template <typename T> __device__ __forceinline__ bool operator<=(Dual<T> a, T b) { return a.v <= b; }
template <typename T> __device__ __forceinline__ bool operator>(Dual<T> a, T b) { return a.v > b; }
template <typename T> __device__ __forceinline__ Dual<T>& operator+=(Dual<T>& a, Dual<T> b) { a.v += b.v; a.t += b.t; return a; }
template <typename T> __device__ __forceinline__ Dual<T>& operator-=(Dual<T>& a, Dual<T> b) { a.v -= b.v; a.t -= b.t; return a; }
template <typename T> __device__ __forceinline__ Dual<T>& operator*=(Dual<T>& a, Dual<T> b) { a = a * b; return a; }
template <typename T> __device__ __forceinline__ bool operator<=(Dual<T> a, Dual<T> b) { return a.v <= b.v; }
template <typename T> __device__ __forceinline__ bool operator>(Dual<T> a, Dual<T> b) { return a.v > b.v; }
template <typename T> __device__ __forceinline__ Dual<T> sqrt(Dual<T> a) {
  T s = sqrt(a.v);
  return {s, a.t / ((T)2 * s)};
}
template <typename T> __device__ __forceinline__ Dual<T> sin(Dual<T> a) { return {sin(a.v), cos(a.v) * a.t}; }
template <typename T> __device__ __forceinline__ Dual<T> cos(Dual<T> a) { return {cos(a.v), -sin(a.v) * a.t}; }

template <typename S> struct BaseOf { using type = S; };
template <typename T> struct BaseOf<Dual<T>> { using type = T; };
template <typename S> __device__ __forceinline__ double value_of(S x) { return (double)x; }
template <typename T> __device__ __forceinline__ double value_of(Dual<T> x) { return (double)x.v; }
template <typename S> __device__ __forceinline__ double tangent_of(S) { return 0.0; }
template <typename T> __device__ __forceinline__ double tangent_of(Dual<T> x) { return (double)x.t; }
template <typename S> struct IsDual { static constexpr bool value = false; };
template <typename T> struct IsDual<Dual<T>> { static constexpr bool value = true; };

// Kernel launchers, instantiated per scalar type S in ipg_snap_k_*.hip.
template <typename S>
void snap_launch_ui(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, S* ulisttot);
template <typename S>
void snap_launch_zby(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const S* ulisttot, const double* beta,
                     const typename BaseOf<S>::type* ycoef, double* Bout, double* eatom, S* Yout, S* Zout,
                     double* tanB);
template <typename T>
void snap_launch_ycoef(ipg_ctx* ctx, const SnapDev& P, const double* beta, T* ycoef);
// Per-pair dB_i/dr_ij [P][ncoeff][3] from the center atoms' Z [N][idxz_max][2].
template <typename S>
void snap_launch_dbdr(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const S* Z, S* dblist);
// FP32 adjoint Z -> (energy per atom, Y) on the atom-blocked U/Y layout
// [ceil(n/32)][idxu_max][32][2]; beta expanded [nspecies][ncoeff].
void snap_launch_zy32(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const float* U, const double* beta, float* Y,
                      double* eatom);
// FP32 descriptors from the atom-blocked U: Z [N][idxz_max][2] (FP32) and
// B [N][ncoeff] (FP64 sums).
void snap_launch_zb32(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const float* U, float* Z, double* B);

// Runtime-table FP64 kernels (ipg_snap_gen.hip): U per element
// [N][nel][idxu_max][2], Z of the first element pair [N][idxz_max][2],
// B [N][ncoeff] and the per-pair dB_i/dr_ij [P][ncoeff][3] in the layout of
// snap_launch_dbdr; eatom (optional) = beta_s(i) . B_i.
void snap_gen_descriptors(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, double* U, double* Z, double* B,
                          const double* beta, double* eatom);
void snap_gen_dbdr(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const double* Z, double* dblist);

}  // namespace ipg
