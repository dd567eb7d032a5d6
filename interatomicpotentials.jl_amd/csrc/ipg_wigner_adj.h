// ipg_wigner_adj.h -- reverse-mode (adjoint) differentiation of the Wigner U
// recursion of one neighbor, for the FP32 adjoint force kernels.
//
// For a neighbor with Cayley-Klein parameters (a, b) the force kernels need
//   Phi(a, b)   = sum over the emitted left-half entries e of
//                 Re(conj(u_e(a, b)) s_e)
// and its gradient with respect to (a, b), where s_e is a per-entry seed
// (2 w(e) Y_e: Y the adjoint of the energy with respect to U, w(e) = 1/2 on
// the centre of an even level, the half-sum weights of compute_b.jl).  The
// gradient with respect to r_ij follows from the Jacobian of (a, b), so the
// derivative rows du/dr of compute_duarray (compute_ui.jl:186-326) are never
// formed: one backward sweep through the recursion replaces the three
// forward-mode derivative sweeps.
//
// Structure of the recursion (compute_ui.jl:50-129).  Level j, row mb is
// built from level j-1, row mb:
//   u_j[ma] = ra(ma) conj(a) u_(j-1)[ma] - rb(ma) conj(b) u_(j-1)[ma-1],
//   ra = sqrt((j-ma)/(j-mb)), rb = sqrt(ma/(j-mb)),
// except the first level of a row (j = 2mb), whose source (row mb of level
// 2mb-1) is the inversion-symmetry mirror of row mb-1 of that level, and the
// middle row of an even level, whose right part is the mirror of its left
// part with the centre's imaginary part negated (the reference's copy loop
// at compute_ui.jl:103-125).  Row mb is therefore a chain over levels
// 2mb..J hanging off row mb-1 at level 2mb-1.
//
// The backward sweep runs the rows in descending mb.  For row mb it
// recomputes the row's start (the staircase rows 0..mb-1 up to level 2mb-1),
// runs the row forward keeping every level in registers (the recursion below
// holds level j in the frame of level j), then walks back down: the adjoint
// of level j-1 is conj-weighted from level j's adjoint plus the seeds of
// level j-1, and every step adds its share to the adjoints of a and b.  The
// adjoint that reaches the row's start is mirrored back onto row mb-1 at
// level 2mb-1, where that row's own sweep picks it up.
//
// The chain runs on rescaled entries v_j[ma] = u_j[ma] / N(j, mb, ma) with
// N = sqrt(ma! (j-ma)! / (mb! (j-mb)!)), for which the step is simply
//   v_j[ma] = conj(a) v_(j-1)[ma] - conj(b) v_(j-1)[ma-1]
// (N(j,mb,ma) / N(j-1,mb,ma) = ra and N(j,mb,ma) / N(j-1,mb,ma-1) = rb).  N is
// symmetric under ma -> j-ma and equal on both sides of a row start (rows mb
// and mb-1 at level 2mb-1), so both mirrors carry over unchanged, and the
// factor N of each entry is folded into its compile-time seed weight.
//
// Complex adjoints use z_bar = dL/dRe(z) + i dL/dIm(z): for w = c z,
// z_bar += conj(c) w_bar and c_bar += conj(z) w_bar; with c = conj(a) the
// adjoint of a itself is conj(c_bar) = z conj(w_bar).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace ipg {
namespace wadj {

template <int V>
using K = std::integral_constant<int, V>;

template <int B, int E, typename F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(K<B>{});
    sfor<B + 1, E>(static_cast<F&&>(f));
  }
}

// sqrt(p/q) at compile time (Newton's iteration, within one ulp in FP64).
constexpr double csqrt(double x) {
  if (!(x > 0.0)) return 0.0;
  double g = x < 1.0 ? 1.0 : x;
  for (int it = 0; it < 400; it++) {
    const double ng = 0.5 * (g + x / g);
    if (ng == g) break;
    g = ng;
  }
  return g;
}
constexpr double cfact(int n) {
  double r = 1.0;
  for (int k = 2; k <= n; k++) r *= (double)k;
  return r;
}
// N(j, mb, ma) of the rescaled chain
constexpr double norm(int j, int mb, int ma) {
  return csqrt(cfact(ma) * cfact(j - ma) / (cfact(mb) * cfact(j - mb)));
}
// Weight 2 w(e) N(e) of the seed of left-half entry e = (j, mb, ma) in the
// rescaled sweep (w = 1/2 on the centre of an even level).
constexpr double seed_weight(int j, int mb, int ma) {
  return ((j == 2 * mb && ma == mb) ? 1.0 : 2.0) * norm(j, mb, ma);
}

// Cayley-Klein parameters of one neighbor.
template <typename T>
struct AB {
  T ar, ai, br, bi;
};

// Phi and the adjoints of a and b.  Each level's partial sums (in the
// chain's scalar type) are added into the wider type W once: FP64, or a
// dual number over FP64 for a sweep run on dual numbers.
template <typename W>
struct AccW {
  W phi = W(0.0), ar = W(0.0), ai = W(0.0), br = W(0.0), bi = W(0.0);
};
using Acc = AccW<double>;

// Scalar type under T: T itself, specialised for dual numbers where they
// are defined.
template <typename T>
struct Base {
  using type = T;
};

// Keeps an accumulator update where it is written: otherwise the compiler
// sinks the FP64 additions to the end of the sweep and every partial sum
// stays live until then.  (Accumulators of other types supply their own
// pin, found by argument-dependent lookup.)
__host__ __device__ __forceinline__ void pin(double& x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#else
  (void)x;
#endif
}

// dst += x in the accumulator's type, pinned where it is written.
template <typename W, typename T>
__host__ __device__ __forceinline__ void acc_add(W& dst, const T& x) {
  dst += W(x);
  pin(dst);
}

// Level JJ of row MB from level JJ-1 (JJ entries in, JJ+1 out), rescaled.
template <typename T, int JJ, int MB>
__host__ __device__ __forceinline__ void fstep(const AB<T>& c, const T (&pr)[JJ], const T (&pi)[JJ],
                                               T (&nr)[JJ + 1], T (&ni)[JJ + 1]) {
  sfor<0, JJ + 1>([&](auto mc) {
    constexpr int ma = decltype(mc)::value;
    // written as chains of single products so that each term is one FMA
    T xr = (T)0, xi = (T)0;
    if constexpr (ma <= JJ - 1) {
      xr = c.ar * pr[ma];
      xr = xr + c.ai * pi[ma];
      xi = c.ar * pi[ma];
      xi = xi - c.ai * pr[ma];
    }
    if constexpr (ma >= 1) {
      xr = xr - c.br * pr[ma - 1];
      xr = xr - c.bi * pi[ma - 1];
      xi = xi - c.br * pi[ma - 1];
      xi = xi + c.bi * pr[ma - 1];
    }
    nr[ma] = xr;
    ni[ma] = xi;
  });
}

// Backward of fstep<T, JJ, MB>: lr/li is the adjoint of level JJ (entries past
// LMAX are zero and skipped), qr/qi receives the adjoint of level JJ-1 and
// acc the adjoints of a and b.  Returns the step's partial sum of Re(a_bar),
// which depends on every entry of the step (an ordering handle).
template <typename T, int JJ, int MB, int LMAX, typename A>
__host__ __device__ __forceinline__ T rstep(const AB<T>& c, const T (&pr)[JJ], const T (&pi)[JJ],
                                               const T (&lr)[JJ + 1], const T (&li)[JJ + 1],
                                               T (&qr)[JJ], T (&qi)[JJ], A& acc) {
  T abr = (T)0, abi = (T)0, bbr = (T)0, bbi = (T)0;
  sfor<0, LMAX + 1>([&](auto mc) {
    constexpr int ma = decltype(mc)::value;
    const T vr = lr[ma], vi = li[ma];
    if constexpr (ma <= JJ - 1) {
      qr[ma] = qr[ma] + c.ar * vr;
      qr[ma] = qr[ma] - c.ai * vi;
      qi[ma] = qi[ma] + c.ar * vi;
      qi[ma] = qi[ma] + c.ai * vr;
      abr = abr + pr[ma] * vr;
      abr = abr + pi[ma] * vi;
      abi = abi + pi[ma] * vr;
      abi = abi - pr[ma] * vi;
    }
    if constexpr (ma >= 1) {
      qr[ma - 1] = qr[ma - 1] - c.br * vr;
      qr[ma - 1] = qr[ma - 1] + c.bi * vi;
      qi[ma - 1] = qi[ma - 1] - c.br * vi;
      qi[ma - 1] = qi[ma - 1] - c.bi * vr;
      bbr = bbr - pr[ma - 1] * vr;
      bbr = bbr - pi[ma - 1] * vi;
      bbi = bbi - pi[ma - 1] * vr;
      bbi = bbi + pr[ma - 1] * vi;
    }
  });
  acc_add(acc.ar, abr);
  acc_add(acc.ai, abi);
  acc_add(acc.br, bbr);
  acc_add(acc.bi, bbi);
  return abr;
}

// Row MB of level 2MB-1 from row MB-1 of that level (inversion symmetry):
// t[x] = (-1)^(MB+x) conj(s[2MB-1-x]).
template <typename T, int MB>
__host__ __device__ __forceinline__ void mirror(const T (&sr)[2 * MB], const T (&si)[2 * MB],
                                                T (&tr)[2 * MB], T (&ti)[2 * MB]) {
  sfor<0, 2 * MB>([&](auto xc) {
    constexpr int x = decltype(xc)::value, y = 2 * MB - 1 - x;
    constexpr bool neg = (MB + x) & 1;
    tr[x] = neg ? -sr[y] : sr[y];
    ti[x] = neg ? si[y] : -si[y];
  });
}

// The middle row of level 2MB: right part mirrored from the left part, the
// centre's imaginary part negated.
template <typename T, int MB>
__host__ __device__ __forceinline__ void middle(T (&wr)[2 * MB + 1], T (&wi)[2 * MB + 1]) {
  sfor<MB + 1, 2 * MB + 1>([&](auto mc) {
    constexpr int ma = decltype(mc)::value, y = 2 * MB - ma;
    constexpr bool neg = (MB + ma) & 1;
    wr[ma] = neg ? -wr[y] : wr[y];
    wi[ma] = neg ? wi[y] : -wi[y];
  });
  wi[MB] = -wi[MB];
}

// Row MB-1 at level 2MB-1 (2MB entries): the staircase of rows 0..MB-1,
// each up to the level where the next row starts.
template <typename T, int MB>
__host__ __device__ __forceinline__ void stair(const AB<T>& c, T (&sr)[2 * MB], T (&si)[2 * MB]) {
  if constexpr (MB == 1) {
    const T v0r[1] = {(T)1}, v0i[1] = {(T)0};
    fstep<T, 1, 0>(c, v0r, v0i, sr, si);
  } else {
    T pr[2 * MB - 2], pi[2 * MB - 2], tr[2 * MB - 2], ti[2 * MB - 2];
    stair<T, MB - 1>(c, pr, pi);
    mirror<T, MB - 1>(pr, pi, tr, ti);
    T wr[2 * MB - 1], wi[2 * MB - 1];
    fstep<T, 2 * MB - 2, MB - 1>(c, tr, ti, wr, wi);
    middle<T, MB - 1>(wr, wi);
    fstep<T, 2 * MB - 1, MB - 1>(c, wr, wi, sr, si);
  }
}

// Level JJ of row MB with values (vr, vi): recurses to level JJ+1, then
// forms this level's adjoint (lr, li): the adjoint handed over by row MB+1
// (at level 2MB+1), plus level JJ+1's adjoint stepped back, plus the seeds
// of this level's emitted entries, which also add to Phi.  Emitted: every
// column above level 2MB, columns 0..MB at level 2MB.
//
// Seed reads are ordered by tokens (seed.order(x) returns an int that
// depends on x).  A level reads its seeds under the token of its child's
// completed backward step and uses them after its own backward step, so the
// reads overlap one step of arithmetic and no level's seeds are held longer.
// done returns a value that depends on this level's completed step.
template <typename T, int J, int MB, int JJ, typename Seed, typename A>
__host__ __device__ __forceinline__ void level(const AB<T>& c, const T (&vr)[JJ + 1], const T (&vi)[JJ + 1],
                                               T (&lr)[JJ + 1], T (&li)[JJ + 1],
                                               const T (&incr)[J + 1], const T (&inci)[J + 1], A& acc,
                                               Seed& seed, T& done) {
  // deeper levels first: on the way up only the values of the levels below
  // stay live
  T mr[JJ < J ? JJ + 2 : 1], mi[JJ < J ? JJ + 2 : 1];
  T cdone = vr[0];
  if constexpr (JJ < J) {
    T nr[JJ + 2], ni[JJ + 2];
    fstep<T, JJ + 1, MB>(c, vr, vi, nr, ni);
    level<T, J, MB, JJ + 1>(c, nr, ni, mr, mi, incr, inci, acc, seed, cdone);
  }
  const int tk = seed.order(cdone);
  T sr[JJ + 1], si[JJ + 1];
  sfor<0, JJ + 1>([&](auto mc) {
    constexpr int ma = decltype(mc)::value;
    if constexpr (JJ > 2 * MB || ma <= MB) {
      const auto s = seed(K<JJ>{}, K<MB>{}, K<ma>{}, tk);
      sr[ma] = s.x;
      si[ma] = s.y;
    }
  });
  sfor<0, JJ + 1>([&](auto mc) {
    constexpr int ma = decltype(mc)::value;
    if constexpr (JJ == 2 * MB + 1 && 2 * (MB + 1) <= J) {
      lr[ma] = incr[ma];
      li[ma] = inci[ma];
    } else {
      lr[ma] = (T)0;
      li[ma] = (T)0;
    }
  });
  done = cdone;
  if constexpr (JJ < J) done = rstep<T, JJ + 1, MB, JJ + 1>(c, vr, vi, mr, mi, lr, li, acc);
  T ph = (T)0;
  sfor<0, JJ + 1>([&](auto mc) {
    constexpr int ma = decltype(mc)::value;
    if constexpr (JJ > 2 * MB || ma <= MB) {
      // 2 w(e) N(e), unless the seed comes with it (Seed::weighted)
      using B = typename Base<T>::type;
      constexpr B wn = Seed::weighted ? (B)1 : (B)seed_weight(JJ, MB, ma);
      const T gr = sr[ma] * wn, gi = si[ma] * wn;
      lr[ma] = lr[ma] + gr;
      li[ma] = li[ma] + gi;
      ph = ph + vr[ma] * gr;
      ph = ph + vi[ma] * gi;
    }
  });
  acc_add(acc.phi, ph);
  if constexpr (JJ == J) done = ph;
}

// Row MB: forward and backward sweep.  inc holds, on entry, the adjoint that
// row MB+1 handed to level 2MB+1 of this row, and on return the adjoint this
// row hands to row MB-1 at level 2MB-1.
template <typename T, int J, int MB, typename Seed, typename A>
__host__ __device__ __forceinline__ void row(const AB<T>& c, T (&incr)[J + 1], T (&inci)[J + 1], A& acc,
                                             Seed& seed) {
  if constexpr (MB == 0) {
    const T vr[1] = {(T)1}, vi[1] = {(T)0};
    T lr[1], li[1], done;
    level<T, J, 0, 0>(c, vr, vi, lr, li, incr, inci, acc, seed, done);
  } else {
    T sr[2 * MB], si[2 * MB], tr[2 * MB], ti[2 * MB];
    stair<T, MB>(c, sr, si);
    mirror<T, MB>(sr, si, tr, ti);
    T wr[2 * MB + 1], wi[2 * MB + 1];
    fstep<T, 2 * MB, MB>(c, tr, ti, wr, wi);
    middle<T, MB>(wr, wi);
    T lr[2 * MB + 1], li[2 * MB + 1], done;
    level<T, J, MB, 2 * MB>(c, wr, wi, lr, li, incr, inci, acc, seed, done);
    // adjoint of the middle-row treatment: columns past MB were overwritten
    T pr[2 * MB + 1], pi[2 * MB + 1];
    sfor<0, MB>([&](auto mc) {
      constexpr int ma = decltype(mc)::value, mm = 2 * MB - ma;
      constexpr bool neg = (MB + mm) & 1;
      pr[ma] = lr[ma] + (neg ? -lr[mm] : lr[mm]);
      pi[ma] = li[ma] + (neg ? li[mm] : -li[mm]);
    });
    pr[MB] = lr[MB];
    pi[MB] = -li[MB];
    sfor<MB + 1, 2 * MB + 1>([&](auto mc) {
      pr[decltype(mc)::value] = (T)0;
      pi[decltype(mc)::value] = (T)0;
    });
    T qr[2 * MB], qi[2 * MB];
    sfor<0, 2 * MB>([&](auto xc) {
      qr[decltype(xc)::value] = (T)0;
      qi[decltype(xc)::value] = (T)0;
    });
    rstep<T, 2 * MB, MB, MB>(c, tr, ti, pr, pi, qr, qi, acc);
    // t[x] = sigma conj(s[y])  =>  s_bar[y] = sigma conj(t_bar[x])
    sfor<0, 2 * MB>([&](auto xc) {
      constexpr int x = decltype(xc)::value, y = 2 * MB - 1 - x;
      constexpr bool neg = (MB + x) & 1;
      incr[y] = neg ? -qr[x] : qr[x];
      inci[y] = neg ? qi[x] : -qi[x];
    });
  }
}

// Phi and its adjoints with respect to a and b for one neighbor.
// seed(K<j>, K<mb>, K<ma>, tok) returns the seed s_e of left-half entry
// (j, mb, ma) as an object with members x (real part) and y (imaginary
// part), read under the ordering token tok; seed.order(x) makes a token that
// depends on x (0 where the order does not matter) and seed.tie(v, x)
// returns v made to depend on x.  Seed::weighted (constexpr bool) says that
// the seeds already carry their weights seed_weight(j, mb, ma).
template <typename T, int J, typename Seed, typename A>
__host__ __device__ __forceinline__ void vjp(const AB<T>& c, Seed& seed, A& acc) {
  T incr[J + 1], inci[J + 1];
  sfor<0, J + 1>([&](auto xc) {
    incr[decltype(xc)::value] = (T)0;
    inci[decltype(xc)::value] = (T)0;
  });
  sfor<0, J / 2 + 1>([&](auto kc) {
    constexpr int MB = J / 2 - decltype(kc)::value;
    if constexpr (MB == J / 2) {
      row<T, J, MB>(c, incr, inci, acc, seed);
    } else {
      // row MB starts once row MB+1 has handed over its adjoint
      const AB<T> ct{seed.tie(c.ar, incr[0]), seed.tie(c.ai, incr[0]), seed.tie(c.br, incr[0]),
                     seed.tie(c.bi, incr[0])};
      row<T, J, MB>(ct, incr, inci, acc, seed);
    }
  });
}

// Position of entry (level j, column ma) among the emitted entries of row MB
// (columns 0..MB of level 2MB, then every column of levels 2MB+1..J), and
// the row's entry count.
constexpr int row_index(int MB, int j, int ma) {
  int k = MB + 1;
  if (j == 2 * MB) return ma;
  for (int l = 2 * MB + 1; l < j; l++) k += l + 1;
  return k + ma;
}
constexpr int row_count(int J, int MB) { return row_index(MB, J + 1, 0); }

// Levels JJ+1..J of row MB after level JJ, each emitted in full.
template <typename T, int J, int MB, int JJ, typename Emit>
__host__ __device__ __forceinline__ void fwd_rest(const AB<T>& c, const T (&vr)[JJ + 1], const T (&vi)[JJ + 1],
                                                  Emit& emit) {
  if constexpr (JJ < J) {
    T nr[JJ + 2], ni[JJ + 2];
    fstep<T, JJ + 1, MB>(c, vr, vi, nr, ni);
    sfor<0, JJ + 2>([&](auto mc) {
      constexpr int ma = decltype(mc)::value;
      emit(K<JJ + 1>{}, mc, nr[ma], ni[ma]);
    });
    const AB<T> cn = emit.next(c);
    fwd_rest<T, J, MB, JJ + 1>(cn, nr, ni, emit);
  }
}

// The forward recursion of row MB alone, from level 0 (the staircase of the
// rows below, then the row itself), emitting its entries in row order with
// their rescaled values v = u / N: emit(K<j>, K<ma>, v_r, v_i).  After each
// level the chain continues with emit.next(c), through which a caller can
// make the next level wait for its use of this one.  The
// staircase up to level 2MB-1 forms MB (2MB+1) entries, the same count as
// levels 0..2MB-1 of a row, so every row forms (J+1)(J+2)/2 entries (45 at
// 2J = 8).
template <typename T, int J, int MB, typename Emit>
__host__ __device__ __forceinline__ void row_forward(const AB<T>& c, Emit& emit) {
  if constexpr (MB == 0) {
    const T vr[1] = {(T)1}, vi[1] = {(T)0};
    emit(K<0>{}, K<0>{}, vr[0], vi[0]);
    const AB<T> cn = emit.next(c);
    fwd_rest<T, J, 0, 0>(cn, vr, vi, emit);
  } else {
    T sr[2 * MB], si[2 * MB], tr[2 * MB], ti[2 * MB];
    stair<T, MB>(c, sr, si);
    mirror<T, MB>(sr, si, tr, ti);
    T wr[2 * MB + 1], wi[2 * MB + 1];
    fstep<T, 2 * MB, MB>(c, tr, ti, wr, wi);
    middle<T, MB>(wr, wi);
    sfor<0, MB + 1>([&](auto mc) { emit(K<2 * MB>{}, mc, wr[decltype(mc)::value], wi[decltype(mc)::value]); });
    const AB<T> cn = emit.next(c);
    fwd_rest<T, J, MB, 2 * MB>(cn, wr, wi, emit);
  }
}

// Adapter for forward(): scales row MB's entries back by N.
template <typename T, int MB, typename Emit>
struct ScaledEmit {
  Emit& emit;
  template <int j, int ma>
  __host__ __device__ __forceinline__ void operator()(K<j> jc, K<ma> mc, T vr, T vi) {
    constexpr T nm = (T)norm(j, MB, ma);
    emit(jc, K<MB>{}, mc, nm * vr, nm * vi);
  }
  __host__ __device__ __forceinline__ AB<T> next(const AB<T>& c) const { return c; }
};

// The plain forward recursion, emitting every left-half entry with its
// values (scaled back by N): emit(K<j>, K<mb>, K<ma>, u_r, u_i).  Used by
// the host-side checks of vjp.
template <typename T, int J, typename Emit>
__host__ __device__ __forceinline__ void forward(const AB<T>& c, Emit& emit) {
  sfor<0, J / 2 + 1>([&](auto mbc) {
    constexpr int MB = decltype(mbc)::value;
    ScaledEmit<T, MB, Emit> e{emit};
    row_forward<T, J, MB>(c, e);
  });
}

}  // namespace wadj
}  // namespace ipg
