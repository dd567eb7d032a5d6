// ipg_snap_kernels.h -- SNAP device kernels, templated on the scalar S
// (float, double, Dual<float>, Dual<double>) and, for the per-neighbor
// Wigner recursions, on twojmax J.  Included by the ipg_snap_k_*.hip files,
// each of which instantiates one scalar type.
//
// Wigner U recursion (compute_ui.jl:50-129 compute_uarray, :186-326
// compute_duarray): level j, row mb is built from level j-1, row mb, except
// the first level of a row (j = 2mb), whose source row mb of level 2mb-1 is
// the inversion-symmetry mirror of row mb-1 of that level.  So every row mb
// is a chain over levels 2mb..J that needs only one row of registers plus
// the mirrored start row saved from chain mb-1.  The middle row of an even
// level gets the reference's treatment: its right part is the mirror of its
// left part and the centre element's imaginary part is negated (the
// reference's copy loop at :103-125 does exactly that).
#pragma once
#include "ipg_snap.h"

#include <algorithm>

namespace ipg {

__host__ __device__ constexpr int ub_of(int j) { return j * (j + 1) * (2 * j + 1) / 6; }  // idxu_block[j]

// ---------------------------------------------------------------------------
// Deterministic sums over groups of G = 16, 32 or 64 lanes: DPP within each
// 16-lane row (xor 1, xor 2, half-row mirror, row mirror: every lane ends with
// its row total), then readlane of the row totals.  All lanes of a group
// receive bit-identical totals.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_v(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_v(double v) {
  const int lo = dpp_i<CTRL>(__double2loint(v)), hi = dpp_i<CTRL>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_v(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane_v(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
template <int G, typename T>
__device__ __forceinline__ T group_sum_t(T v) {
  v = v + dpp_v<0xB1>(v);   // quad_perm [1,0,3,2]
  v = v + dpp_v<0x4E>(v);   // quad_perm [2,3,0,1]
  v = v + dpp_v<0x141>(v);  // row_half_mirror
  v = v + dpp_v<0x140>(v);  // row_mirror
  if constexpr (G == 16) {
    return v;
  } else {
    const T r0 = readlane_v(v, 0), r1 = readlane_v(v, 16), r2 = readlane_v(v, 32), r3 = readlane_v(v, 48);
    if constexpr (G == 32) return (threadIdx.x & 32) ? r2 + r3 : r0 + r1;
    else return (r0 + r1) + (r2 + r3);
  }
}
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
  return group_sum_t<G, T>(v);
}
template <int G, typename T>
__device__ __forceinline__ Dual<T> group_sum(Dual<T> v) {
  return Dual<T>(group_sum_t<G, T>(v.v), group_sum_t<G, T>(v.t));
}

// Wider accumulator type: float -> double (so FP32 kernels keep FP64 sums).
template <typename S> struct AccOf { using type = double; };
template <typename T> struct AccOf<Dual<T>> { using type = Dual<double>; };
template <typename S> __device__ __forceinline__ typename AccOf<S>::type to_acc(S x) { return (double)x; }
template <typename T> __device__ __forceinline__ Dual<double> to_acc(Dual<T> x) {
  return Dual<double>((double)x.v, (double)x.t);
}
template <typename S> __device__ __forceinline__ S from_acc(double x) { return (S)x; }
template <typename S> __device__ __forceinline__ S from_acc(Dual<double> x) {
  using T = typename BaseOf<S>::type;
  return S((T)x.v, (T)x.t);
}

template <typename S>
struct CK {
  S a_r, a_i, b_r, b_i;
  S da_r[3], da_i[3], db_r[3], db_i[3];
  S u[3];
  S sfac, dsfac;  // switching function and derivative, times w_j
};

// Cayley-Klein parameters of one neighbor and their gradients
// (compute_ui.jl:328-347 compute_duidrj, :191-216; sfac/dsfac :350-380).
template <typename S, bool DERIV>
__device__ __forceinline__ void cayley_klein_impl(S x, S y, S z, double rcut, double wj, const SnapDev& P, CK<S>& c) {
  using T = typename BaseOf<S>::type;
  S rsq = x * x + y * y + z * z;
  S r = sqrt(rsq);
  const T rscale0 = (T)(P.rfac0 * M_PI / (rcut - P.rmin0));
  S theta0 = (r - (T)P.rmin0) * rscale0;
  S cs = cos(theta0), sn = sin(theta0);
  S z0 = r * cs / sn;
  S r0inv = (T)1 / sqrt(r * r + z0 * z0);
  c.a_r = z0 * r0inv;
  c.a_i = -z * r0inv;
  c.b_r = y * r0inv;
  c.b_i = -x * r0inv;
  S sfac, dsfac;
  if (!P.switch_flag || r <= (T)P.rmin0) {
    sfac = S((T)1);
    dsfac = S((T)0);
  } else {
    const T rcutfac = (T)(M_PI / (rcut - P.rmin0));
    S ang = (r - (T)P.rmin0) * rcutfac;
    sfac = (r > (T)rcut) ? S((T)0) : (T)0.5 * (cos(ang) + (T)1);
    dsfac = (T)-0.5 * (sin(ang) * rcutfac);
  }
  c.sfac = sfac * (T)wj;
  c.dsfac = dsfac * (T)wj;
  if (DERIV) {
    S dz0dr = z0 / r - (r * rscale0) * (rsq + z0 * z0) / rsq;
    S rinv = (T)1 / r;
    c.u[0] = x * rinv;
    c.u[1] = y * rinv;
    c.u[2] = z * rinv;
    S dr0invdr = -r0inv * r0inv * r0inv * (r + z0 * dz0dr);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      S dr0inv = dr0invdr * c.u[k];
      S dz0 = dz0dr * c.u[k];
      c.da_r[k] = dz0 * r0inv + z0 * dr0inv;
      c.da_i[k] = -z * dr0inv;
      c.db_r[k] = y * dr0inv;
      c.db_i[k] = -x * dr0inv;
    }
    c.da_i[2] = c.da_i[2] - r0inv;
    c.db_i[0] = c.db_i[0] - r0inv;
    c.db_r[1] = c.db_r[1] + r0inv;
  }
}

// The Cayley-Klein parameters are evaluated at FP64 width (Dual<double> for
// tangent runs) and rounded once to the kernel's scalar type: cot(theta0)
// near the cutoff amplifies FP32 rounding by ~1/sin^2(theta0).
template <typename S> struct WideOf { using type = double; };
template <typename T> struct WideOf<Dual<T>> { using type = Dual<double>; };
template <typename S> __device__ __forceinline__ S narrow(double x) { return (S)x; }
template <typename S> __device__ __forceinline__ S narrow(Dual<double> x) {
  using T = typename BaseOf<S>::type;
  return S((T)x.v, (T)x.t);
}
template <typename S, bool DERIV>
__device__ __forceinline__ void cayley_klein(typename WideOf<S>::type x, typename WideOf<S>::type y,
                                             typename WideOf<S>::type z, double rcut, double wj, const SnapDev& P,
                                             CK<S>& c) {
  using W = typename WideOf<S>::type;
  CK<W> w;
  cayley_klein_impl<W, DERIV>(x, y, z, rcut, wj, P, w);
  c.a_r = narrow<S>(w.a_r);
  c.a_i = narrow<S>(w.a_i);
  c.b_r = narrow<S>(w.b_r);
  c.b_i = narrow<S>(w.b_i);
  c.sfac = narrow<S>(w.sfac);
  c.dsfac = narrow<S>(w.dsfac);
  if (DERIV) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      c.da_r[k] = narrow<S>(w.da_r[k]);
      c.da_i[k] = narrow<S>(w.da_i[k]);
      c.db_r[k] = narrow<S>(w.db_r[k]);
      c.db_i[k] = narrow<S>(w.db_i[k]);
      c.u[k] = narrow<S>(w.u[k]);
    }
  }
}

template <int V>
using IC = std::integral_constant<int, V>;

// Compile-time loop: f(IC<B>), f(IC<B+1>), ..., f(IC<E-1>).  Every index the
// Wigner recursion touches is a template constant, so the register rows are
// indexed statically and sqrt(p/q) folds to an immediate.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(IC<B>{});
    static_for<B + 1, E>(static_cast<F&&>(f));
  }
}

// sqrt(p/q) (prebuilt_arrays.jl:314-321) with constant p, q.
template <typename T, int P, int Q>
__device__ __forceinline__ T rootpq_c() {
  return (T)__builtin_sqrt((double)P / (double)Q);
}

// One recursion step, in place: level JJ row MB from level JJ-1 row MB
// (compute_ui.jl:71-100, :227-267).  Descending ma keeps the old values of
// ma and ma-1 available.
template <int J, typename S, int NK, int JJ, int MB>
__device__ __forceinline__ void wigner_step(const CK<S>& c, S (&ur)[J + 1], S (&ui)[J + 1],
                                            S (&dur)[NK > 0 ? NK : 1][J + 1], S (&dui)[NK > 0 ? NK : 1][J + 1]) {
  using T = typename BaseOf<S>::type;
  static_for<0, JJ + 1>([&](auto ii) {
    constexpr int ma = JJ - decltype(ii)::value;
    S nr = S((T)0), ni = S((T)0);
    S ndr[NK > 0 ? NK : 1], ndi[NK > 0 ? NK : 1];
#pragma unroll
    for (int k = 0; k < NK; k++) {
      ndr[k] = S((T)0);
      ndi[k] = S((T)0);
    }
    if constexpr (ma >= 1) {
      const T rb = rootpq_c<T, ma, JJ - MB>();
      const S pr = ur[ma - 1], pi = ui[ma - 1];
      nr = -rb * (c.b_r * pr + c.b_i * pi);
      ni = -rb * (c.b_r * pi - c.b_i * pr);
#pragma unroll
      for (int k = 0; k < NK; k++) {
        ndr[k] = -rb * (c.db_r[k] * pr + c.db_i[k] * pi + c.b_r * dur[k][ma - 1] + c.b_i * dui[k][ma - 1]);
        ndi[k] = -rb * (c.db_r[k] * pi - c.db_i[k] * pr + c.b_r * dui[k][ma - 1] - c.b_i * dur[k][ma - 1]);
      }
    }
    if constexpr (ma <= JJ - 1) {
      const T ra = rootpq_c<T, JJ - ma, JJ - MB>();
      const S pr = ur[ma], pi = ui[ma];
      nr += ra * (c.a_r * pr + c.a_i * pi);
      ni += ra * (c.a_r * pi - c.a_i * pr);
#pragma unroll
      for (int k = 0; k < NK; k++) {
        ndr[k] += ra * (c.da_r[k] * pr + c.da_i[k] * pi + c.a_r * dur[k][ma] + c.a_i * dui[k][ma]);
        ndi[k] += ra * (c.da_r[k] * pi - c.da_i[k] * pr + c.a_r * dui[k][ma] - c.a_i * dur[k][ma]);
      }
    }
    ur[ma] = nr;
    ui[ma] = ni;
#pragma unroll
    for (int k = 0; k < NK; k++) {
      dur[k][ma] = ndr[k];
      dui[k][ma] = ndi[k];
    }
  });
}

// Visits every unique left-half entry (level j, row mb, column ma with
// 2mb < j, or 2mb == j and ma <= mb) once, calling
//   emit(IC<j>, IC<mb>, IC<ma>, IC<half>, IC<last>, u_r, u_i, du_r, du_i)
// where half = 1 marks the centre of an even level (weight 1/2 in the
// half-column sums of compute_b.jl), last = 1 marks the last entry of an
// emitted row, and du_r/du_i are the raw recursion derivative rows (before
// the switching-function scaling of compute_ui.jl:300-325).
template <int J, typename S, int NK, typename Emit>
__device__ __forceinline__ void wigner_chain(const CK<S>& c, Emit&& emit) {
  using T = typename BaseOf<S>::type;
  constexpr int NKA = NK > 0 ? NK : 1;
  S ur[J + 1], ui[J + 1], dur[NKA][J + 1], dui[NKA][J + 1];
  S sr[J + 1], si[J + 1], sdr[NKA][J + 1], sdi[NKA][J + 1];
  static_for<0, J / 2 + 1>([&](auto mbc) {
    constexpr int MB = decltype(mbc)::value;
    if constexpr (MB == 0) {
      ur[0] = S((T)1);
      ui[0] = S((T)0);
#pragma unroll
      for (int k = 0; k < NK; k++) {
        dur[k][0] = S((T)0);
        dui[k][0] = S((T)0);
      }
      emit(IC<0>{}, IC<0>{}, IC<0>{}, IC<1>{}, IC<1>{}, ur[0], ui[0], dur, dui);
    } else {
      // row MB of level 2MB-1 = (-1)^(MB+x) conj(row MB-1 of level 2MB-1)[2MB-1-x]
      static_for<0, 2 * MB>([&](auto xc) {
        constexpr int x = decltype(xc)::value;
        constexpr bool neg = (MB + x) & 1;
        constexpr int y = 2 * MB - 1 - x;
        ur[x] = neg ? -sr[y] : sr[y];
        ui[x] = neg ? si[y] : -si[y];
#pragma unroll
        for (int k = 0; k < NK; k++) {
          dur[k][x] = neg ? -sdr[k][y] : sdr[k][y];
          dui[k][x] = neg ? sdi[k][y] : -sdi[k][y];
        }
      });
      wigner_step<J, S, NK, 2 * MB, MB>(c, ur, ui, dur, dui);
      // middle row of the even level 2MB
      static_for<MB + 1, 2 * MB + 1>([&](auto mac) {
        constexpr int ma = decltype(mac)::value;
        constexpr bool neg = (MB + ma) & 1;
        constexpr int y = 2 * MB - ma;
        ur[ma] = neg ? -ur[y] : ur[y];
        ui[ma] = neg ? ui[y] : -ui[y];
#pragma unroll
        for (int k = 0; k < NK; k++) {
          dur[k][ma] = neg ? -dur[k][y] : dur[k][y];
          dui[k][ma] = neg ? dui[k][y] : -dui[k][y];
        }
      });
      ui[MB] = -ui[MB];
#pragma unroll
      for (int k = 0; k < NK; k++) dui[k][MB] = -dui[k][MB];
      static_for<0, MB + 1>([&](auto mac) {
        constexpr int ma = decltype(mac)::value;
        emit(IC<2 * MB>{}, IC<MB>{}, IC<ma>{}, IC<(ma == MB) ? 1 : 0>{}, IC<(ma == MB) ? 1 : 0>{}, ur[ma], ui[ma],
             dur, dui);
      });
    }
    static_for<2 * MB + 1, J + 1>([&](auto jc) {
      constexpr int JJ = decltype(jc)::value;
      wigner_step<J, S, NK, JJ, MB>(c, ur, ui, dur, dui);
      static_for<0, JJ + 1>([&](auto mac) {
        constexpr int ma = decltype(mac)::value;
        emit(IC<JJ>{}, IC<MB>{}, IC<ma>{}, IC<0>{}, IC<(ma == JJ) ? 1 : 0>{}, ur[ma], ui[ma], dur, dui);
      });
      if constexpr (JJ == 2 * MB + 1 && 2 * (MB + 1) <= J) {
#pragma unroll
        for (int x = 0; x <= JJ; x++) {
          sr[x] = ur[x];
          si[x] = ui[x];
#pragma unroll
          for (int k = 0; k < NK; k++) {
            sdr[k][x] = dur[k][x];
            sdi[k][x] = dui[k][x];
          }
        }
      }
    });
  });
}

// Pair geometry (+ tangent of r_ij = x_j - x_i along R.tang for Dual runs).
template <typename S>
__device__ __forceinline__ void load_pair(const SnapRun& R, int64_t q, int32_t i, int32_t j, S& x, S& y, S& z) {
  using T = typename BaseOf<S>::type;
  const double* rv = R.rvec + 3 * q;
  if constexpr (IsDual<S>::value) {
    T tx = 0, ty = 0, tz = 0;
    if (R.tang) {
      tx = (T)(R.tang[3 * j] - R.tang[3 * i]);
      ty = (T)(R.tang[3 * j + 1] - R.tang[3 * i + 1]);
      tz = (T)(R.tang[3 * j + 2] - R.tang[3 * i + 2]);
    }
    x = S((T)rv[0], tx);
    y = S((T)rv[1], ty);
    z = S((T)rv[2], tz);
  } else {
    x = (S)rv[0];
    y = (S)rv[1];
    z = (S)rv[2];
  }
}

// Position of entry (j, mb, ma) in wigner_chain's emission order, and the
// number of emitted entries.
constexpr int emit_index(int J, int j, int mb, int ma) {
  int k = 0;
  for (int MB = 0; 2 * MB <= J; MB++) {
    for (int a = 0; a <= MB; a++) {  // level 2MB, columns 0..MB
      if (j == 2 * MB && mb == MB && ma == a) return k;
      k++;
    }
    for (int JJ = 2 * MB + 1; JJ <= J; JJ++)
      for (int a = 0; a <= JJ; a++) {
        if (j == JJ && mb == MB && ma == a) return k;
        k++;
      }
  }
  return k;  // past the end: the count
}
template <int J>
struct EmitJju {  // jju of the k-th emitted entry
  int v[emit_index(J, -1, -1, -1)] = {};
  constexpr EmitJju() {
    int k = 0;
    for (int MB = 0; 2 * MB <= J; MB++) {
      for (int a = 0; a <= MB; a++) v[k++] = ub_of(2 * MB) + (2 * MB + 1) * MB + a;
      for (int JJ = 2 * MB + 1; JJ <= J; JJ++)
        for (int a = 0; a <= JJ; a++) v[k++] = ub_of(JJ) + (JJ + 1) * MB + a;
    }
  }
};

// Transposed sum over the G lanes of a group of the NV values every lane
// holds (NV a power of two, NV <= G): exchanges at xor distances G/2, G/4,
// ... each keep half of the values (the upper half where the lane has that
// bit), then plain butterfly adds cover the remaining distances.  Lane l
// ends with the group total of value l >> log2(G/NV) in v[0]; lanes that
// differ only in the low log2(G/NV) bits hold the same total.
template <int M, int D, typename T>
__device__ __forceinline__ void transpose_sum(T* v, int l) {
  if constexpr (D >= 1) {
    if constexpr (M > 1) {
      const bool hi = (l & D) != 0;
      static_for<0, M / 2>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const T send = hi ? v[k] : v[k + M / 2];
        const T keep = hi ? v[k + M / 2] : v[k];
        v[k] = keep + __shfl_xor(send, D, 64);
      });
      transpose_sum<M / 2, D / 2>(v, l);
    } else {
      v[0] = v[0] + __shfl_xor(v[0], D, 64);
      transpose_sum<1, D / 2>(v, l);
    }
  }
}

// The same transposed sum for float values without the LDS crossbar:
// exchanges at distance 32 and 16 are v_permlane32_swap / v_permlane16_swap
// of the two halves (after the swap, keep + partner's send is simply the
// sum of the two registers in every lane), shorter distances read the
// partner through DPP.  The DPP partner is a mirror within 16 or 8 lanes
// or a quad permutation: it differs from the lane in bit D and agrees on
// the higher bits, so each stage still adds two disjoint lane sets holding
// the same values, and lane l ends with the same value as in transpose_sum.
template <int D>
__device__ __forceinline__ float dpp_partner(float x) {
  static_assert(D == 8 || D == 4 || D == 2 || D == 1, "DPP partner distance");
  // row_mirror (15 - l), row_half_mirror (7 - l), quad_perm [2,3,0,1], [1,0,3,2]
  constexpr int ctrl = D == 8 ? 0x140 : D == 4 ? 0x141 : D == 2 ? 0x4E : 0xB1;
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), ctrl, 0xF, 0xF, false));
}
template <int D>
__device__ __forceinline__ float swap_sum(float a, float b) {
  if constexpr (D == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  } else {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
}
template <int M, int D>
__device__ __forceinline__ void transpose_sum_f32(float* v, int l) {
  if constexpr (D >= 1) {
    if constexpr (M > 1) {
      static_for<0, M / 2>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (D >= 16) {
          v[k] = swap_sum<D>(v[k], v[k + M / 2]);
        } else {
          const bool hi = (l & D) != 0;
          const float send = hi ? v[k] : v[k + M / 2];
          const float keep = hi ? v[k + M / 2] : v[k];
          v[k] = keep + dpp_partner<D>(send);
        }
      });
      transpose_sum_f32<M / 2, D / 2>(v, l);
    } else {
      if constexpr (D >= 16) v[0] = swap_sum<D>(v[0], v[0]);
      else v[0] = v[0] + dpp_partner<D>(v[0]);
      transpose_sum_f32<1, D / 2>(v, l);
    }
  }
}

// Components of a U value for the lane sums: (re, im), and for Dual runs
// the tangents after them.
template <typename S> struct UComp { static constexpr int NC = 2; };
template <typename T> struct UComp<Dual<T>> { static constexpr int NC = 4; };
template <typename T>
__device__ __forceinline__ void put_comp(T* buf, int B, int e, T r, T i) {
  buf[e] = r;
  buf[B + e] = i;
}
template <typename T>
__device__ __forceinline__ void put_comp(T* buf, int B, int e, Dual<T> r, Dual<T> i) {
  buf[e] = r.v;
  buf[B + e] = i.v;
  buf[2 * B + e] = r.t;
  buf[3 * B + e] = i.t;
}

// ---------------------------------------------------------------------------
// K1: ulisttot (compute_ui.jl:5-44, zero_uarraytot :131-152, add_uarraytot
// :155-175).  One wave holds APW atoms; the G = 64/APW lanes of an atom take
// one neighbor each (looping over chunks of G neighbors) and run the Wigner
// row chain in registers.  The emitted entries w_j u_j are buffered in
// batches of B entries; each full batch is summed over the group with one
// transposed reduction, after which every lane owns one (entry, component)
// of the batch and keeps its running total in a register across chunks.
// The totals are written to LDS once per atom; the right half is filled
// from the left half by the inversion symmetry and wself is added on the
// diagonal.  All sums run in a fixed order, so the result is deterministic.
// ---------------------------------------------------------------------------
template <int J, typename S, int APW>
__global__ __launch_bounds__(64) void k_snap_ui(SnapDev P, SnapRun R, S* __restrict__ ulisttot) {
  using T = typename BaseOf<S>::type;
  constexpr int G = 64 / APW;
  constexpr int IDXU = ub_of(J + 1);
  constexpr int NC = UComp<S>::NC;
  constexpr int B = 32 * 4 / (NC * (int)sizeof(T));  // entries per batch: NC * B values of 32 VGPRs
  constexpr int NV = NC * B;
  constexpr int NE = emit_index(J, -1, -1, -1);
  constexpr int NB = (NE + B - 1) / B;
  constexpr int SHIFT = (G == 64 ? 6 : 5) - (NV == 64 ? 6 : NV == 32 ? 5 : NV == 16 ? 4 : 3);
  static_assert(NV <= G, "batch larger than the lane group");
  __shared__ S acc[APW][IDXU][2];
  const int lane = threadIdx.x;
  const int g = lane / G, l = lane % G;
  const int64_t a = (int64_t)blockIdx.x * APW + g;
  const bool valid = a < R.n;
  T tot[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) tot[b] = (T)0;
  const int64_t q0 = valid ? R.off[a] : 0, q1 = valid ? R.off[a + 1] : 0;
  int nch = (int)((q1 - q0 + G - 1) / G);
  nch = max(max(__builtin_amdgcn_readlane(nch, 0), __builtin_amdgcn_readlane(nch, 16)),
            max(__builtin_amdgcn_readlane(nch, 32), __builtin_amdgcn_readlane(nch, 48)));
  if (valid && l == 0 && R.species && R.flag && (unsigned)R.species[a] >= (unsigned)P.nspecies) atomicOr(R.flag, 16);
  const int si = valid ? snap_species(R.species, a, P.nspecies) : 0;
  for (int ch = 0; ch < nch; ch++) {
    const int64_t q = q0 + (int64_t)ch * G + l;
    const bool act = q < q1;
    CK<S> c;
    if (act) {
      const int32_t j = R.nbr[q];
      const int sj = snap_species(R.species, j, P.nspecies);
      typename WideOf<S>::type x, y, z;
      load_pair<typename WideOf<S>::type>(R, q, (int32_t)a, j, x, y, z);
      const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
      cayley_klein<S, false>(x, y, z, rcut, P.weight[sj], P, c);
    } else {
      c.a_r = S((T)1);
      c.a_i = c.b_r = c.b_i = S((T)0);
      c.sfac = S((T)0);
    }
    const S w = c.sfac;
    T buf[NV];
    wigner_chain<J, S, 0>(c, [&](auto jc, auto mbc, auto mac, auto, auto, S u_r, S u_i, auto&, auto&) {
      constexpr int k = emit_index(J, decltype(jc)::value, decltype(mbc)::value, decltype(mac)::value);
      constexpr int e = k % B, b = k / B;
      put_comp(buf, B, e, w * u_r, w * u_i);
      if constexpr (e == B - 1 || k == NE - 1) {
        if constexpr (e < B - 1)  // last, partial batch: clear the unused slots
          static_for<e + 1, B>([&](auto ec) {
            static_for<0, NC>([&](auto cc) { buf[decltype(cc)::value * B + decltype(ec)::value] = (T)0; });
          });
        if constexpr (std::is_same<T, float>::value) transpose_sum_f32<NV, G / 2>(buf, l);
        else transpose_sum<NV, G / 2>(buf, l);
        tot[b] += buf[0];
      }
    });
  }
  // each lane's totals to LDS: lane l owns value l >> SHIFT of every batch
  if ((l & ((1 << SHIFT) - 1)) == 0) {
    constexpr EmitJju<J> ej{};
    const int idx = l >> SHIFT, comp = idx / B, e = idx % B;
    T* accT = reinterpret_cast<T*>(&acc[g][0][0]);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int k = b * B + e;
      if (k < NE) {
        const int jju = ej.v[k];
        // S is (value) or (value, tangent); comp: 0 re, 1 im, 2 re tangent, 3 im tangent
        accT[(jju * 2 + (comp & 1)) * (NC / 2) + (comp >> 1)] = tot[b];
      }
    }
  }
  __syncthreads();
  if (!valid) return;
  // FP32 runs write the atom-blocked layout [a/32][jju][a%32] of the Z/Y kernel
  constexpr bool BLK = std::is_same<S, float>::value;
  constexpr int ES = BLK ? 64 : 2;  // element stride
  S* out = BLK ? ulisttot + ((a >> 5) * (int64_t)IDXU * 32 + (a & 31)) * 2 : ulisttot + (int64_t)a * IDXU * 2;
  for (int e = l; e < IDXU; e += G) {
    int j = 0;
    while (e >= ub_of(j + 1)) j++;
    const int rem = e - ub_of(j);
    const int mb = rem / (j + 1), ma = rem % (j + 1);
    S vr, vi;
    if (2 * mb < j || (2 * mb == j && ma <= mb)) {
      vr = acc[g][e][0];
      vi = acc[g][e][1];
    } else {
      const int src = ub_of(j) + (j - mb) * (j + 1) + (j - ma);
      const bool neg = (mb + ma) & 1;
      vr = neg ? -acc[g][src][0] : acc[g][src][0];
      vi = neg ? acc[g][src][1] : -acc[g][src][1];
    }
    if (ma == mb) vr = vr + (T)P.wself;
    out[ES * e] = vr;
    out[ES * e + 1] = vi;
  }
}

// ---------------------------------------------------------------------------
// K2: Z (compute_z.jl:1-62), B (compute_b.jl:5-95) and the adjoint
// Y = dE/dU.  Persistent: each 256-thread block stages the Clebsch-Gordan
// table in LDS once, keeps its Z entries' compact plans in registers and
// loops over atoms; U and Z of the current atom live in LDS.
// Y_l(e) = sum over Z blocks (a,c,l) of coef(a,c,l) Z_(a,c,l)(e), with coef
// collecting the beta entries whose dB/dr half-sums use that block.
// ---------------------------------------------------------------------------
// Block shape: 256 threads, each with up to 10 Z entries (2,386 at
// 2J = 8); FP64 dual numbers take 512 threads with 5 entries each, since
// their U and Z in LDS (85 KB) leave room for one block per CU (measured
// 1.14 ms against 1.47 ms with 256; FP32 duals, three blocks per CU, were
// slower at 512).
template <typename S>
struct ZbyShape {
  static constexpr int NT = 256, KMAX = 10;
};
template <>
struct ZbyShape<Dual<double>> {
  static constexpr int NT = 512, KMAX = 5;
};

template <typename S>
__global__ __launch_bounds__(ZbyShape<S>::NT) void k_snap_zby(SnapDev P, SnapRun R, const S* __restrict__ ulisttot,
                                                          const double* __restrict__ beta,
                                                          const typename BaseOf<S>::type* __restrict__ ycoef,
                                                          double* __restrict__ Bout, double* __restrict__ eatom,
                                                          S* __restrict__ Yout, S* __restrict__ Zout,
                                                          double* __restrict__ tanB) {
  using T = typename BaseOf<S>::type;
  using A = typename AccOf<S>::type;
  constexpr int ZBY_THREADS = ZbyShape<S>::NT, ZBY_KMAX = ZbyShape<S>::KMAX;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* cgL = reinterpret_cast<T*>(smem_raw);
  S* U = reinterpret_cast<S*>(smem_raw + (((size_t)P.idxcg_max * sizeof(T) + 15) & ~(size_t)15));
  S* Z = U + 2 * P.idxu_max;
  __shared__ double red[ZBY_THREADS / 64];
  const int tid = threadIdx.x;
  const T* cgt;
  if constexpr (std::is_same<T, float>::value) cgt = P.cg_f;
  else cgt = P.cg_d;
  for (int e = tid; e < P.idxcg_max; e += ZBY_THREADS) cgL[e] = cgt[e];
  uint32_t zp0[ZBY_KMAX], zp1[ZBY_KMAX], zp2[ZBY_KMAX];
#pragma unroll
  for (int k = 0; k < ZBY_KMAX; k++) {
    const int jjz = tid + ZBY_THREADS * k;
    zp0[k] = zp1[k] = zp2[k] = 0;
    if (jjz < P.idxz_max) {
      zp0[k] = P.zplan[3 * jjz];
      zp1[k] = P.zplan[3 * jjz + 1];
      zp2[k] = P.zplan[3 * jjz + 2];
    }
  }
  const int M = P.ncoeff * P.nspecies;
  for (int64_t a = blockIdx.x; a < R.n; a += gridDim.x) {
    __syncthreads();  // the previous atom's readers of U and Z are done
    const int s = snap_species(R.species, a, P.nspecies);
    const S* ug = ulisttot + a * (int64_t)P.idxu_max * 2;
    for (int e = tid; e < 2 * P.idxu_max; e += ZBY_THREADS) U[e] = ug[e];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ZBY_KMAX; k++) {
      const int jjz = tid + ZBY_THREADS * k;
      if (jjz < P.idxz_max) {
        // plan: u1 base | u2 base << 16; cga0 | cgb0 << 16; na | nb<<4 | (j1+1)<<8 | j2<<12 | (j+1)<<16
        const int u1b = (int)(zp0[k] & 0xFFFF), u2b = (int)(zp0[k] >> 16);
        const int cga0 = (int)(zp1[k] & 0xFFFF), cgb0 = (int)(zp1[k] >> 16);
        const int na = (int)(zp2[k] & 0xF), nb = (int)((zp2[k] >> 4) & 0xF);
        const int s1 = (int)((zp2[k] >> 8) & 0xF), j2 = (int)((zp2[k] >> 12) & 0xF);
        A zr = A(0.0), zi = A(0.0);
        int u1 = u1b, u2 = u2b, cgb = cgb0;
        for (int ib = 0; ib < nb; ib++) {
          S sr = S((T)0), si_ = S((T)0);
          int cga = cga0;
          for (int ia = 0; ia < na; ia++) {
            const S u1r = U[2 * (u1 + ia)], u1i = U[2 * (u1 + ia) + 1];
            const S u2r = U[2 * (u2 - ia)], u2i = U[2 * (u2 - ia) + 1];
            const T c = cgL[cga];
            sr += c * (u1r * u2r - u1i * u2i);
            si_ += c * (u1r * u2i + u1i * u2r);
            cga += j2;
          }
          const T cb = cgL[cgb];
          zr += to_acc(cb * sr);
          zi += to_acc(cb * si_);
          u1 += s1;
          u2 -= j2 + 1;
          cgb += j2;
        }
        if (P.bnorm) {
          const double jp1 = (double)(zp2[k] >> 16);
          zr = zr / jp1;
          zi = zi / jp1;
        }
        Z[2 * jjz] = from_acc<S>(zr);
        Z[2 * jjz + 1] = from_acc<S>(zi);
      }
    }
    __syncthreads();
    if (Zout) {
      S* zg = Zout + a * (int64_t)P.idxz_max * 2;
      for (int e = tid; e < 2 * P.idxz_max; e += ZBY_THREADS) zg[e] = Z[e];
    }
    // B and the per-atom energy beta_s . B
    double epart = 0.0;
    if (tanB)
      for (int c = tid; c < M; c += ZBY_THREADS)
        if (c / P.ncoeff != s) tanB[a * M + c] = 0.0;
    for (int jjb = tid; jjb < P.idxb_max; jjb += ZBY_THREADS) {
      const BTerm bt = P.bterm[jjb];
      const int j = bt.j;
      int jju = P.idxu_block[j], jjz = bt.jjz0[0];
      A sum = A(0.0);
      for (int mb = 0; 2 * mb < j; mb++)
        for (int ma = 0; ma <= j; ma++, jju++, jjz++)
          sum += to_acc(U[2 * jju] * Z[2 * jjz] + U[2 * jju + 1] * Z[2 * jjz + 1]);
      if (j % 2 == 0) {
        const int mb = j / 2;
        for (int ma = 0; ma < mb; ma++, jju++, jjz++)
          sum += to_acc(U[2 * jju] * Z[2 * jjz] + U[2 * jju + 1] * Z[2 * jjz + 1]);
        sum += 0.5 * to_acc(U[2 * jju] * Z[2 * jjz] + U[2 * jju + 1] * Z[2 * jjz + 1]);
      }
      A b = 2.0 * sum;
      if (P.bzero) b = b - P.bzero_j[j];
      const double bv = value_of(b);
      if (Bout) Bout[a * P.ncoeff + jjb] = bv;
      if (beta) epart += beta[s * P.ncoeff + jjb] * bv;
      if (tanB) tanB[a * M + s * P.ncoeff + jjb] = tangent_of(b);
    }
    if (eatom) {
      for (int o = 32; o > 0; o >>= 1) epart += __shfl_xor(epart, o, 64);
      if ((tid & 63) == 0) red[tid >> 6] = epart;
      __syncthreads();
      if (tid == 0) {
        double t = red[0];
        for (int w = 1; w < ZBY_THREADS / 64; w++) t += red[w];
        eatom[a] = t;
      }
    }
    if (Yout) {
      S* yg = Yout + a * (int64_t)P.idxu_max * 2;
      const T* yc = ycoef + (int64_t)s * P.nblk;
      for (int h = tid; h < P.nhalf; h += ZBY_THREADS) {
        A yr = A(0.0), yi = A(0.0);
        for (int k = P.yoff[h]; k < P.yoff[h + 1]; k++) {
          const int jjz = P.yjjz[k];
          const T cf = yc[P.yjblk[k]];
          yr += to_acc(cf * Z[2 * jjz]);
          yi += to_acc(cf * Z[2 * jjz + 1]);
        }
        const int jju = P.half_jju[h];
        yg[2 * jju] = from_acc<S>(yr);
        yg[2 * jju + 1] = from_acc<S>(yi);
      }
    }
  }
}

template <typename T>
__global__ void k_snap_ycoef(SnapDev P, const double* __restrict__ beta, T* __restrict__ ycoef) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.nspecies * P.nblk) return;
  const int s = t / P.nblk, b = t % P.nblk;
  const YBlk yb = P.yblk[b];
  double c = 0.0;
  for (int k = 0; k < yb.n; k++) c += beta[s * P.ncoeff + yb.jjb[k]] * yb.fac[k];
  ycoef[t] = (T)c;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename S, int J>
static void launch_ui_j(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, S* ulisttot, int apw) {
  if (apw == 2) {
    unsigned g = (unsigned)((R.n + 1) / 2);
    k_snap_ui<J, S, 2><<<g, 64, 0, ctx->stream>>>(P, R, ulisttot);
  } else {
    k_snap_ui<J, S, 1><<<(unsigned)R.n, 64, 0, ctx->stream>>>(P, R, ulisttot);
  }
}

template <typename S>
void snap_launch_ui(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, S* ulisttot) {
  if (R.n == 0) return;
  const double avg = R.avg_nbr >= 0.0 ? R.avg_nbr : (double)R.npairs / (double)R.n;
  // 32 lanes per atom hold up to 32 neighbors in one chunk
  const int apw = R.apw == 1 || R.apw == 2 ? R.apw : (avg <= 30.0 ? 2 : 1);
  const char* nm = IsDual<S>::value ? "snap_ui_dual" : (sizeof(S) == 4 ? "snap_ui_fp32" : "snap_ui_fp64");
  launch(ctx, nm, [&] {
    switch (P.J) {
      case 0: launch_ui_j<S, 0>(ctx, P, R, ulisttot, apw); break;
      case 1: launch_ui_j<S, 1>(ctx, P, R, ulisttot, apw); break;
      case 2: launch_ui_j<S, 2>(ctx, P, R, ulisttot, apw); break;
      case 3: launch_ui_j<S, 3>(ctx, P, R, ulisttot, apw); break;
      case 4: launch_ui_j<S, 4>(ctx, P, R, ulisttot, apw); break;
      case 5: launch_ui_j<S, 5>(ctx, P, R, ulisttot, apw); break;
      case 6: launch_ui_j<S, 6>(ctx, P, R, ulisttot, apw); break;
      case 7: launch_ui_j<S, 7>(ctx, P, R, ulisttot, apw); break;
      default: launch_ui_j<S, 8>(ctx, P, R, ulisttot, apw); break;
    }
  });
}

template <typename S>
void snap_launch_zby(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const S* ulisttot, const double* beta,
                     const typename BaseOf<S>::type* ycoef, double* Bout, double* eatom, S* Yout, S* Zout,
                     double* tanB) {
  using T = typename BaseOf<S>::type;
  if (R.n == 0) return;
  constexpr int NT = ZbyShape<S>::NT;
  IPG_REQUIRE(P.idxz_max <= NT * ZbyShape<S>::KMAX, "internal: Z table larger than the kernel plan");
  const size_t sm = (((size_t)P.idxcg_max * sizeof(T) + 15) & ~(size_t)15) +
                    sizeof(S) * 2 * (size_t)(P.idxu_max + P.idxz_max);
  const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / (sm + 1024))));
  const int64_t grid = std::min<int64_t>(R.n, (int64_t)ctx->num_cus * per_cu);
  const char* nm = IsDual<S>::value ? "snap_zby_dual" : (sizeof(S) == 4 ? "snap_zby_fp32" : "snap_zby_fp64");
  launch(ctx, nm, [&] {
    if (sm > 64 * 1024)
      IPG_HIP(hipFuncSetAttribute((const void*)k_snap_zby<S>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    k_snap_zby<S><<<(unsigned)grid, NT, sm, ctx->stream>>>(P, R, ulisttot, beta, ycoef, Bout, eatom, Yout, Zout,
                                                           tanB);
  });
}

#define IPG_SNAP_INSTANTIATE_U(S) template void snap_launch_ui<S>(ipg_ctx*, const SnapDev&, const SnapRun&, S*);
#define IPG_SNAP_INSTANTIATE(S)                                                                                  \
  IPG_SNAP_INSTANTIATE_U(S)                                                                                      \
  template void snap_launch_zby<S>(ipg_ctx*, const SnapDev&, const SnapRun&, const S*, const double*,            \
                                   const typename BaseOf<S>::type*, double*, double*, S*, S*, double*);

// Each launcher is instantiated once, in ipg_snap_k_*.hip; every other
// source that includes this header only calls them.
#define IPG_SNAP_EXTERN(S)                                                                                       \
  extern template void snap_launch_ui<S>(ipg_ctx*, const SnapDev&, const SnapRun&, S*);                          \
  extern template void snap_launch_zby<S>(ipg_ctx*, const SnapDev&, const SnapRun&, const S*, const double*,     \
                                          const typename BaseOf<S>::type*, double*, double*, S*, S*, double*);
extern template void snap_launch_ui<float>(ipg_ctx*, const SnapDev&, const SnapRun&, float*);
IPG_SNAP_EXTERN(double)
IPG_SNAP_EXTERN(Dual<float>)
IPG_SNAP_EXTERN(Dual<double>)

}  // namespace ipg
