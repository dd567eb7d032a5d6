// ipg_snap_zy.h -- FP32 adjoint kernel: Z (compute_z.jl:1-62), the energy
// beta . B (compute_b.jl:5-95, linear_potential.jl:20-26) and the adjoint
// Y = dE/dU for batches of 32 atoms, with every lane owning one atom.
//
// Work layout.  The output rows (j, mb), mb <= j/2, of Y are the units of
// work; a unit sums coef(j1,j2,j) * Z_(j1,j2,j)[mb][:] over every block
// (j1, j2) that produces level j.  Units are assigned to the waves of a
// block at compile time (longest-processing-time first), so all control
// flow is wave-uniform.  Inside a block (j1, j2, j) the loop over the
// mb1 rows is the only run-time loop: the two 32-lane halves of a wave take
// alternate mb1 rows, load row mb1 of U_j1 and row mb2 of U_j2 once into
// registers and form every column ma of the output row from them, with the
// Clebsch-Gordan factors of the column sums as compile-time literals.  U of
// the batch sits in LDS as [jju][32 atoms] (conflict-free 8-byte reads).
// Each Z entry is summed in FP32 over its rows (the column sums and at most
// five row terms per half), then widened once: Y and the energy accumulate in
// FP64 registers.  Z is never stored in the adjoint mode.
//
// The U and Y buffers use the atom-blocked layout [batch][jju][32][re, im]
// written by the FP32 U kernel and read by the FP32 force kernel.  Y holds
// the entries the force kernel's adjoint sweep seeds (mb < j/2, and
// ma <= mb on the middle row of an even level), each already multiplied by
// its seed weight 2 w(e) N(e) (ipg_wigner_adj.h).
#pragma once
#include "ipg_snap_kernels.h"
#include "ipg_wigner_adj.h"

namespace ipg {
namespace zy {

constexpr int AB = 32;  // atoms per batch (one per lane of a half wave)

// x + (x of the lane 32 apart): the sum of the two halves' partial values,
// by v_permlane32_swap (one lane sees its own value in one result and the
// partner's in the other) instead of a shuffle through the LDS crossbar.
__device__ __forceinline__ float half_sum(float x) { return swap_sum<32>(x, x); }
__device__ __forceinline__ double half_sum(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)b, (uint32_t)b, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(b >> 32), (uint32_t)(b >> 32), false, false);
  const double d0 = __longlong_as_double((long long)(((uint64_t)hi[0] << 32) | lo[0]));
  const double d1 = __longlong_as_double((long long)(((uint64_t)hi[1] << 32) | lo[1]));
  return d0 + d1;
}

constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr double fact(int n) {
  double r = 1.0;
  for (int k = 2; k <= n; k++) r *= (double)k;
  return r;
}
// Newton's iteration for sqrt at compile time (converges to within one ulp).
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
// <j1/2 a/2, j2/2 b/2 | j/2 (a+b)/2>, Racah formula with doubled arguments
// (the same closed form as the host table).
constexpr double cg2(int j1, int j2, int j, int a, int b) {
  const int c = a + b;
  if (c < -j || c > j) return 0.0;
  const double tri =
      fact((j1 + j2 - j) / 2) * fact((j1 - j2 + j) / 2) * fact((-j1 + j2 + j) / 2) / fact((j1 + j2 + j) / 2 + 1);
  const double nrm = (double)(j + 1) * fact((j1 + a) / 2) * fact((j1 - a) / 2) * fact((j2 + b) / 2) *
                     fact((j2 - b) / 2) * fact((j + c) / 2) * fact((j - c) / 2);
  const int zmin = cmax(0, cmax(-(j - j2 + a) / 2, -(j - j1 - b) / 2));
  const int zmax = cmin((j1 + j2 - j) / 2, cmin((j1 - a) / 2, (j2 + b) / 2));
  double s = 0.0;
  for (int z = zmin; z <= zmax; z++) {
    const double den = fact(z) * fact((j1 + j2 - j) / 2 - z) * fact((j1 - a) / 2 - z) * fact((j2 + b) / 2 - z) *
                       fact((j - j2 + a) / 2 + z) * fact((j - j1 - b) / 2 + z);
    s += ((z & 1) ? -1.0 : 1.0) / den;
  }
  return csqrt(tri) * csqrt(nrm) * s;
}
// table entry [m1][m2] of block (j1, j2, j)
constexpr double cgm(int j1, int j2, int j, int m1, int m2) { return cg2(j1, j2, j, 2 * m1 - j1, 2 * m2 - j2); }

constexpr bool valid(int J, int j1, int j2, int j) {
  return j2 <= j1 && j1 <= J && j >= j1 - j2 && j <= cmin(J, j1 + j2) && ((j1 + j2 + j) % 2 == 0);
}
// Position of block (j1, j2, j) in the j1 >= j2, j = j1-j2 .. min(J, j1+j2)
// step 2 order (types/prebuilt_arrays.jl:48-321): Z-block index, B index
// (blocks with j >= j1) and Clebsch-Gordan table offset.
// zoff: start of the block's Z rows (jjz); bslot: first of the j/2 + 1 row
// slots of its B partial sums (blocks with j >= j1 only).
struct BlockPos {
  int blk, jjb, cgoff, zoff, bslot;
};
constexpr BlockPos block_pos(int J, int j1, int j2, int j) {
  BlockPos p{0, 0, 0, 0, 0};
  for (int a = 0; a <= J; a++)
    for (int b = 0; b <= a; b++)
      for (int c = a - b; c <= cmin(J, a + b); c += 2) {
        if (a == j1 && b == j2 && c == j) return p;
        p.blk++;
        if (c >= a) {
          p.jjb++;
          p.bslot += c / 2 + 1;
        }
        p.cgoff += (a + 1) * (b + 1);
        p.zoff += (c / 2 + 1) * (c + 1);
      }
  return p;
}
// B-sum slot table: slot start and level of every coefficient, and the
// slot and Z counts.
template <int J>
struct BSlots {
  static constexpr int count_b() {
    int n = 0;
    for (int a = 0; a <= J; a++)
      for (int b = 0; b <= a; b++)
        for (int c = a - b; c <= cmin(J, a + b); c += 2)
          if (c >= a) n++;
    return n;
  }
  static constexpr int NBC = count_b();
  int slot0[NBC] = {}, lev[NBC] = {};
  int nslot = 0, nz = 0;
  constexpr BSlots() {
    int k = 0;
    for (int a = 0; a <= J; a++)
      for (int b = 0; b <= a; b++)
        for (int c = a - b; c <= cmin(J, a + b); c += 2) {
          nz += (c / 2 + 1) * (c + 1);
          if (c >= a) {
            slot0[k] = nslot;
            lev[k] = c;
            nslot += c / 2 + 1;
            k++;
          }
        }
  }
};
// Column-sum ranges of one output entry (the ma/mb rules of compute_z.jl).
constexpr int m1min(int j1, int j2, int j, int m) { return cmax(0, (2 * m - j - j2 + j1) / 2); }
constexpr int m2max(int j1, int j2, int j, int m) { return (2 * m - j - (2 * m1min(j1, j2, j, m) - j1) + j2) / 2; }
constexpr int mcount(int j1, int j2, int j, int m) {
  return cmin(j1, (2 * m - j + j2 + j1) / 2) - m1min(j1, j2, j, m) + 1;
}

// Cost model of a unit (j, mb): products plus row loads per mb1 step.
constexpr int unit_cost(int J, int j, int mb) {
  int cost = 0;
  for (int j1 = 0; j1 <= J; j1++)
    for (int j2 = 0; j2 <= j1; j2++) {
      if (!valid(J, j1, j2, j)) continue;
      int pr = 0;
      for (int ma = 0; ma <= j; ma++) pr += mcount(j1, j2, j, ma);
      const int nb = mcount(j1, j2, j, mb);
      cost += nb * (pr + (j1 + j2 + 2) / 2) + 4 * (j + 1);
    }
  return cost;
}
constexpr int num_units(int J) {
  int u = 0;
  for (int j = 0; j <= J; j++) u += j / 2 + 1;
  return u;
}

template <int J, int NW>
struct Plan {
  static constexpr int NU = num_units(J);
  int uj[NU] = {}, umb[NU] = {}, owner[NU] = {};
  constexpr Plan() {
    int cost[NU] = {}, order[NU] = {};
    int u = 0;
    for (int j = 0; j <= J; j++)
      for (int mb = 0; 2 * mb <= j; mb++) {
        uj[u] = j;
        umb[u] = mb;
        cost[u] = unit_cost(J, j, mb);
        order[u] = u;
        u++;
      }
    for (int a = 0; a < NU; a++)  // by decreasing cost
      for (int b = a + 1; b < NU; b++)
        if (cost[order[b]] > cost[order[a]]) {
          const int t = order[a];
          order[a] = order[b];
          order[b] = t;
        }
    int load[NW] = {};
    for (int k = 0; k < NU; k++) {
      int best = 0;
      for (int w = 1; w < NW; w++)
        if (load[w] < load[best]) best = w;
      owner[order[k]] = best;
      load[best] += cost[order[k]];
    }
  }
};

// One block (j1, j2, j) of output row mb: Z row (this half's mb1 rows),
// its energy share and its Y contribution.
// DESC: instead of the energy and Y, store the Z row (both halves hold it
// after the exchange; half h writes the columns of parity h) and put the
// row's share 2 sum w Re(conj(U_j) Z) of B_(j1,j2,j) in its LDS slot.
template <int J, int j1, int j2, int j, int mb, bool DESC>
__device__ __forceinline__ void triple(const float2* __restrict__ UL, const float* __restrict__ cgL, int h, int at,
                                       const double* __restrict__ ycs, const double* __restrict__ bts,
                                       const SnapDev& P, double* yr, double* yi, double& e,
                                       float2* __restrict__ Zat, double* __restrict__ Bp, bool store) {
  constexpr int mb1lo = m1min(j1, j2, j, mb);
  constexpr int mb2hi = m2max(j1, j2, j, mb);
  constexpr int nb = mcount(j1, j2, j, mb);
  constexpr BlockPos bp = block_pos(J, j1, j2, j);
  // columns formed: all of them for the stored Z (DESC); for Y and the
  // energy the middle row of an even level needs only its left half and
  // centre (the adjoint sweep and the B half-sums use no other entry)
  constexpr int NA = (!DESC && 2 * mb == j) ? mb + 1 : j + 1;
  // Z of this half's mb1 rows: the FP32 column sums times the row factor,
  // summed over at most five rows in FP32 and widened to FP64 once
  float zrf[j + 1], zif[j + 1];
  static_for<0, NA>([&](auto mc) {
    zrf[decltype(mc)::value] = 0.f;
    zif[decltype(mc)::value] = 0.f;
  });
  for (int ib = h; ib < nb; ib += 2) {
    const int mb1 = mb1lo + ib, mb2 = mb2hi - ib;
    const float2* r1 = UL + (ub_of(j1) + mb1 * (j1 + 1)) * AB + at;
    const float2* r2 = UL + (ub_of(j2) + mb2 * (j2 + 1)) * AB + at;
    float2 u1[j1 + 1], u2[j2 + 1];
    static_for<0, j1 + 1>([&](auto k) { u1[decltype(k)::value] = r1[decltype(k)::value * AB]; });
    static_for<0, j2 + 1>([&](auto k) { u2[decltype(k)::value] = r2[decltype(k)::value * AB]; });
    const float cb = cgL[bp.cgoff + mb1 * (j2 + 1) + mb2];
    static_for<0, NA>([&](auto mc) {
      constexpr int ma = decltype(mc)::value;
      constexpr int lo = m1min(j1, j2, j, ma), hi2 = m2max(j1, j2, j, ma), na = mcount(j1, j2, j, ma);
      float sr = 0.f, si = 0.f;
      static_for<0, na>([&](auto ic) {
        constexpr int ma1 = lo + decltype(ic)::value, ma2 = hi2 - decltype(ic)::value;
        constexpr float c = (float)cgm(j1, j2, j, ma1, ma2);
        const float2 p = u1[ma1], q = u2[ma2];
        sr += c * (p.x * q.x - p.y * q.y);
        si += c * (p.x * q.y + p.y * q.x);
      });
      zrf[ma] += cb * sr;
      zif[ma] += cb * si;
    });
  }
  double zr[j + 1], zi[j + 1];
  if constexpr (DESC) {
    static_for<0, j + 1>([&](auto mc) {
      constexpr int ma = decltype(mc)::value;
      zr[ma] = (double)half_sum(zrf[ma]);
      zi[ma] = (double)half_sum(zif[ma]);
    });
    if (store) {
      float2* zrow = Zat + bp.zoff + mb * (j + 1);
      static_for<0, (j + 2) / 2>([&](auto kc) {
        constexpr int m0 = 2 * decltype(kc)::value;
        if constexpr (m0 + 1 <= j) {
          const float2 v = h ? make_float2((float)zr[m0 + 1], (float)zi[m0 + 1])
                             : make_float2((float)zr[m0], (float)zi[m0]);
          zrow[m0 + h] = v;
        } else {
          if (h == 0) zrow[m0] = make_float2((float)zr[m0], (float)zi[m0]);
        }
      });
    }
    if constexpr (j >= j1) {
      const float2* uj = UL + (ub_of(j) + mb * (j + 1)) * AB + at;
      double bs = 0.0;
      static_for<0, j + 1>([&](auto mc) {
        constexpr int ma = decltype(mc)::value;
        constexpr double wgt = (2 * mb < j) ? 1.0 : (ma < mb ? 1.0 : (ma == mb ? 0.5 : 0.0));
        if constexpr (wgt != 0.0) {
          const float2 uu = uj[ma * AB];
          bs += wgt * ((double)uu.x * zr[ma] + (double)uu.y * zi[ma]);
        }
      });
      if (h == 0) Bp[(bp.bslot + mb) * AB] = 2.0 * bs;
    }
    return;
  }
  static_for<0, NA>([&](auto mc) {
    zr[decltype(mc)::value] = (double)zrf[decltype(mc)::value];
    zi[decltype(mc)::value] = (double)zif[decltype(mc)::value];
  });
  if constexpr (j >= j1) {
    // this half's share of B_(j1,j2,j) = 2 sum_half Re(conj(U_j) Z) (the
    // middle row of an even level counts its first half and half its centre)
    const float2* uj = UL + (ub_of(j) + mb * (j + 1)) * AB + at;
    double bs = 0.0;
    static_for<0, j + 1>([&](auto mc) {
      constexpr int ma = decltype(mc)::value;
      constexpr double wgt = (2 * mb < j) ? 1.0 : (ma < mb ? 1.0 : (ma == mb ? 0.5 : 0.0));
      if constexpr (wgt != 0.0) {
        const float2 uu = uj[ma * AB];
        bs += wgt * ((double)uu.x * zr[ma] + (double)uu.y * zi[ma]);
      }
    });
    double ev = 2.0 * bs;
    if constexpr (mb == 0)
      if (h == 0 && P.bzero) ev -= P.bzero_j[j];
    e += bts[bp.jjb] * ev;
  }
  const double cf = ycs[bp.blk];
  static_for<0, NA>([&](auto mc) {
    yr[decltype(mc)::value] += cf * zr[decltype(mc)::value];
    yi[decltype(mc)::value] += cf * zi[decltype(mc)::value];
    // pin the update here: otherwise it is sunk to the unit's end and the
    // Z rows of every block stay live until then
    asm volatile("" : "+v"(yr[decltype(mc)::value]), "+v"(yi[decltype(mc)::value]));
  });
  asm volatile("" : "+v"(e));
}

template <int J, int j, int mb, bool DESC>
__device__ __forceinline__ void unit(const float2* __restrict__ UL, const float* __restrict__ cgL, int h, int at,
                                     const double* __restrict__ ycs, const double* __restrict__ bts,
                                     const SnapDev& P, double& e, float2* __restrict__ Yrow, bool store,
                                     float2* __restrict__ Zat, double* __restrict__ Bp) {
  constexpr int NA = (!DESC && 2 * mb == j) ? mb + 1 : j + 1;  // columns used (see triple)
  double yr[j + 1], yi[j + 1];
  static_for<0, NA>([&](auto mc) {
    yr[decltype(mc)::value] = 0.0;
    yi[decltype(mc)::value] = 0.0;
  });
  static_for<0, J + 1>([&](auto c1) {
    constexpr int j1 = decltype(c1)::value;
    static_for<0, j1 + 1>([&](auto c2) {
      constexpr int j2 = decltype(c2)::value;
      if constexpr (valid(J, j1, j2, j))
        triple<J, j1, j2, j, mb, DESC>(UL, cgL, h, at, ycs, bts, P, yr, yi, e, Zat, Bp, store);
    });
  });
  if constexpr (DESC) return;
  // the two halves hold partial sums over alternate mb1 rows
  static_for<0, NA>([&](auto mc) {
    constexpr int ma = decltype(mc)::value;
    yr[ma] = half_sum(yr[ma]);
    yi[ma] = half_sum(yi[ma]);
  });
  if (!store) return;
  // half h stores the columns ma = 2k + h (one instruction per column pair),
  // each weighted by 2 w(e) N(e), the seed weight of the rescaled adjoint
  // sweep of the force kernel (its only reader)
  static_for<0, (NA + 1) / 2>([&](auto kc) {
    constexpr int m0 = 2 * decltype(kc)::value;
    constexpr double w0 = wadj::seed_weight(j, mb, m0);
    if constexpr (m0 + 1 < NA) {
      constexpr double w1 = wadj::seed_weight(j, mb, m0 + 1);
      const float2 v = h ? make_float2((float)(w1 * yr[m0 + 1]), (float)(w1 * yi[m0 + 1]))
                         : make_float2((float)(w0 * yr[m0]), (float)(w0 * yi[m0]));
      Yrow[(m0 + h) * AB] = v;
    } else {
      if (h == 0) Yrow[m0 * AB] = make_float2((float)(w0 * yr[m0]), (float)(w0 * yi[m0]));
    }
  });
}

// ---------------------------------------------------------------------------
// Adjoint schedule with shared products.  The rows (j, mb) with the same
// d = j - 2mb read the same pairs of U rows in every block (j1, j2):
// mb1 + mb2 = S = (j1 + j2 - d) / 2, and for a column sum c = ma1 + ma2 row
// j forms column ma = c - (j1 + j2 - j) / 2 from the same products
// U_j1(mb1, ma1) U_j2(mb2, ma2), the larger rows' columns covering the
// smaller rows'.  A piece is a run of rows of one d (a table for 2J = 8,
// otherwise runs of at most PIECE_CAP columns); the wave that owns it forms
// each product once and adds it,
// with each row's literal Clebsch-Gordan factor, into the rows that use it.
// Every Y row belongs to one piece, so no sums cross waves.
// ---------------------------------------------------------------------------
constexpr int PIECE_CAP = 9;
constexpr int PIECE_ROWS = 5;
constexpr int row_na(int j, int mb) { return 2 * mb == j ? mb + 1 : j + 1; }  // columns of Y used
constexpr int cnum(int j1, int j2, int c) { return cmin(j1, c) - cmax(0, c - j2) + 1; }  // products of sum c

// Pieces of 2J = 8: runs of rows (their levels j, largest first) of one
// d = j - 2mb, chosen offline to minimise the largest wave load after the
// LPT assignment below (cost as in piece_cost) with at most 16 columns per
// piece; -1 pads a run.
constexpr int PIECES_J8_N = 15;
constexpr int PIECES_J8_D[PIECES_J8_N] = {0, 0, 1, 1, 2, 2, 2, 3, 3, 4, 4, 5, 6, 7, 8};
constexpr int PIECES_J8[PIECES_J8_N][3] = {{8, 6, 4}, {2, 0, -1},           // d = 0
                                           {7, 5, -1}, {3, 1, -1},          // d = 1
                                           {8, -1, -1}, {6, -1, -1}, {4, 2, -1},  // d = 2
                                           {7, 5, -1}, {3, -1, -1},         // d = 3
                                           {8, -1, -1}, {6, 4, -1},         // d = 4
                                           {7, 5, -1},                      // d = 5
                                           {8, 6, -1},                      // d = 6
                                           {7, -1, -1},                     // d = 7
                                           {8, -1, -1}};                    // d = 8

template <int J>
struct PieceSet {
  static constexpr int NPMAX = 48;
  int np = 0;
  int pd[NPMAX] = {}, pn[NPMAX] = {}, pj[NPMAX][PIECE_ROWS] = {}, cost[NPMAX] = {};
  constexpr PieceSet() {
    if (J == 8) {
      np = PIECES_J8_N;
      for (int p = 0; p < np; p++) {
        pd[p] = PIECES_J8_D[p];
        for (int k = 0; k < 3 && PIECES_J8[p][k] >= 0; k++) pj[p][pn[p]++] = PIECES_J8[p][k];
        cost[p] = piece_cost(p);
      }
      return;
    }
    for (int d = 0; d <= J; d++) {
      int cur = -1, used = 0;
      for (int j = J - ((J - d) & 1); j >= d; j -= 2) {  // rows of d, largest first
        const int e = row_na(j, (j - d) / 2);
        if (cur < 0 || used + e > PIECE_CAP || pn[cur] == PIECE_ROWS) {
          cur = np++;
          pd[cur] = d;
          used = 0;
        }
        pj[cur][pn[cur]++] = j;
        used += e;
      }
    }
    for (int p = 0; p < np; p++) cost[p] = piece_cost(p);
  }
  // products, weighted adds and row updates over every block, plus the row
  // loads and the Y/energy work per row
  constexpr int piece_cost(int p) const {
    const int d = pd[p];
    int c = 0;
    for (int j1 = 0; j1 <= J; j1++)
      for (int j2 = 0; j2 <= j1; j2++) {
        if ((j1 + j2 + d) & 1) continue;
        const int S = (j1 + j2 - d) / 2;
        const int nb = cmin(j1, S) - cmax(0, S - j2) + 1;
        if (nb <= 0) continue;
        int prod = 0, uses = 0, rows = 0, top = -1;
        for (int k = 0; k < pn[p]; k++) {
          const int j = pj[p][k];
          if (!valid(J, j1, j2, j)) continue;
          const int off = (j1 + j2 - j) / 2, na = row_na(j, (j - d) / 2);
          for (int c2 = off; c2 < off + na; c2++) uses += cnum(j1, j2, c2);
          rows += na;
          if (top < 0) {
            top = j;
            for (int c2 = off; c2 < off + na; c2++) prod += cnum(j1, j2, c2);
          }
        }
        if (top < 0) continue;
        c += nb * (4 * prod + 2 * uses + 2 * rows + j1 + j2 + 2) + 8 * rows;
      }
    return c;
  }
};

template <int J, int NW>
struct PiecePlan {
  PieceSet<J> ps{};
  int owner[PieceSet<J>::NPMAX] = {};
  constexpr PiecePlan() {
    int order[PieceSet<J>::NPMAX] = {};
    for (int p = 0; p < ps.np; p++) order[p] = p;
    for (int a = 0; a < ps.np; a++)  // by decreasing cost
      for (int b = a + 1; b < ps.np; b++)
        if (ps.cost[order[b]] > ps.cost[order[a]]) {
          const int t = order[a];
          order[a] = order[b];
          order[b] = t;
        }
    int load[NW] = {};
    for (int k = 0; k < ps.np; k++) {
      int best = 0;
      for (int w = 1; w < NW; w++)
        if (load[w] < load[best]) best = w;
      owner[order[k]] = best;
      load[best] += ps.cost[order[k]];
    }
  }
};
template <int J, int NW>
struct PiecePlanOf {
  static constexpr PiecePlan<J, NW> v{};
};

// Block (j1, j2) of piece P: the Z rows of its rows valid in the block,
// over this half's row pairs, then each row's energy share and Y update.
template <int J, int NW, int P, int j1, int j2>
__device__ __forceinline__ void piece_block(const float2* __restrict__ UL, const float* __restrict__ cgL, int h,
                                            int at, const double* __restrict__ ycs, const double* __restrict__ bts,
                                            const SnapDev& Pd, double (&yr)[PIECE_ROWS][J + 1],
                                            double (&yi)[PIECE_ROWS][J + 1], double& e) {
  constexpr auto& ps = PiecePlanOf<J, NW>::v.ps;
  constexpr int d = ps.pd[P], NR = ps.pn[P];
  constexpr int S = (j1 + j2 - d) / 2;
  constexpr int lo = cmax(0, S - j2), nb = cmin(j1, S) - lo + 1;
  float zr[PIECE_ROWS][J + 1], zi[PIECE_ROWS][J + 1];
  static_for<0, NR>([&](auto kc) {
    constexpr int k = decltype(kc)::value, j = ps.pj[P][k];
    if constexpr (valid(J, j1, j2, j))
      static_for<0, row_na(j, (j - d) / 2)>([&](auto mc) {
        zr[k][decltype(mc)::value] = 0.f;
        zi[k][decltype(mc)::value] = 0.f;
      });
  });
  // the largest row valid in the block spans every column sum the others use
  constexpr int top = [] {
    for (int k = 0; k < NR; k++)
      if (valid(J, j1, j2, ps.pj[P][k])) return k;
    return -1;
  }();
  static_assert(top >= 0, "piece without a row in the block");
  constexpr int jt = ps.pj[P][top];
  constexpr int c0 = (j1 + j2 - jt) / 2, c1 = c0 + row_na(jt, (jt - d) / 2);
  for (int ib = h; ib < nb; ib += 2) {
    const int mb1 = lo + ib, mb2 = S - mb1;
    const float2* r1 = UL + (ub_of(j1) + mb1 * (j1 + 1)) * AB + at;
    const float2* r2 = UL + (ub_of(j2) + mb2 * (j2 + 1)) * AB + at;
    float2 u1[j1 + 1], u2[j2 + 1];
    static_for<0, j1 + 1>([&](auto k) { u1[decltype(k)::value] = r1[decltype(k)::value * AB]; });
    static_for<0, j2 + 1>([&](auto k) { u2[decltype(k)::value] = r2[decltype(k)::value * AB]; });
    float cb[PIECE_ROWS];
    static_for<0, NR>([&](auto kc) {
      constexpr int k = decltype(kc)::value, j = ps.pj[P][k];
      if constexpr (valid(J, j1, j2, j)) cb[k] = cgL[block_pos(J, j1, j2, j).cgoff + mb1 * (j2 + 1) + mb2];
    });
    static_for<c0, c1>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      float sr[PIECE_ROWS], si[PIECE_ROWS];
      static_for<0, NR>([&](auto kc) {
        sr[decltype(kc)::value] = 0.f;
        si[decltype(kc)::value] = 0.f;
      });
      static_for<cmax(0, c - j2), cmin(j1, c) + 1>([&](auto ac) {
        constexpr int ma1 = decltype(ac)::value, ma2 = c - ma1;
        const float2 p = u1[ma1], q = u2[ma2];
        const float pr = p.x * q.x - p.y * q.y, pi = p.x * q.y + p.y * q.x;
        static_for<0, NR>([&](auto kc) {
          constexpr int k = decltype(kc)::value, j = ps.pj[P][k];
          if constexpr (valid(J, j1, j2, j)) {
            constexpr int ma = c - (j1 + j2 - j) / 2;
            if constexpr (ma >= 0 && ma < row_na(j, (j - d) / 2)) {
              constexpr float w = (float)cgm(j1, j2, j, ma1, ma2);
              sr[k] += w * pr;
              si[k] += w * pi;
            }
          }
        });
      });
      static_for<0, NR>([&](auto kc) {
        constexpr int k = decltype(kc)::value, j = ps.pj[P][k];
        if constexpr (valid(J, j1, j2, j)) {
          constexpr int ma = c - (j1 + j2 - j) / 2;
          if constexpr (ma >= 0 && ma < row_na(j, (j - d) / 2)) {
            zr[k][ma] += cb[k] * sr[k];
            zi[k][ma] += cb[k] * si[k];
          }
        }
      });
    });
  }
  static_for<0, NR>([&](auto kc) {
    constexpr int k = decltype(kc)::value, j = ps.pj[P][k], mb = (j - d) / 2;
    if constexpr (valid(J, j1, j2, j)) {
      constexpr int NA = row_na(j, mb);
      constexpr BlockPos bp = block_pos(J, j1, j2, j);
      double z64r[NA], z64i[NA];
      static_for<0, NA>([&](auto mc) {
        z64r[decltype(mc)::value] = (double)zr[k][decltype(mc)::value];
        z64i[decltype(mc)::value] = (double)zi[k][decltype(mc)::value];
      });
      if constexpr (j >= j1) {
        // this half's share of B_(j1,j2,j) = 2 sum_half Re(conj(U_j) Z)
        const float2* uj = UL + (ub_of(j) + mb * (j + 1)) * AB + at;
        double bs = 0.0;
        static_for<0, NA>([&](auto mc) {
          constexpr int ma = decltype(mc)::value;
          constexpr double wgt = (2 * mb < j) ? 1.0 : (ma < mb ? 1.0 : 0.5);
          const float2 uu = uj[ma * AB];
          bs += wgt * ((double)uu.x * z64r[ma] + (double)uu.y * z64i[ma]);
        });
        double ev = 2.0 * bs;
        if constexpr (mb == 0)
          if (h == 0 && Pd.bzero) ev -= Pd.bzero_j[j];
        e += bts[bp.jjb] * ev;
      }
      const double cf = ycs[bp.blk];
      static_for<0, NA>([&](auto mc) {
        constexpr int ma = decltype(mc)::value;
        yr[k][ma] += cf * z64r[ma];
        yi[k][ma] += cf * z64i[ma];
        asm volatile("" : "+v"(yr[k][ma]), "+v"(yi[k][ma]));
      });
      asm volatile("" : "+v"(e));
    }
  });
}

// Piece P: every block, then the rows' Y (both halves' sums) stored with
// the force kernel's seed weights.
template <int J, int NW, int P>
__device__ __forceinline__ void piece(const float2* __restrict__ UL, const float* __restrict__ cgL, int h, int at,
                                      const double* __restrict__ ycs, const double* __restrict__ bts,
                                      const SnapDev& Pd, double& e, float2* __restrict__ yb, bool store) {
  constexpr auto& ps = PiecePlanOf<J, NW>::v.ps;
  constexpr int d = ps.pd[P], NR = ps.pn[P];
  double yr[PIECE_ROWS][J + 1], yi[PIECE_ROWS][J + 1];
  static_for<0, NR>([&](auto kc) {
    constexpr int k = decltype(kc)::value, j = ps.pj[P][k];
    static_for<0, row_na(j, (j - d) / 2)>([&](auto mc) {
      yr[k][decltype(mc)::value] = 0.0;
      yi[k][decltype(mc)::value] = 0.0;
    });
  });
  static_for<0, J + 1>([&](auto c1) {
    constexpr int j1 = decltype(c1)::value;
    static_for<0, j1 + 1>([&](auto c2) {
      constexpr int j2 = decltype(c2)::value;
      constexpr bool any = [] {
        for (int k = 0; k < NR; k++)
          if (valid(J, j1, j2, ps.pj[P][k])) return true;
        return false;
      }();
      if constexpr (any) piece_block<J, NW, P, j1, j2>(UL, cgL, h, at, ycs, bts, Pd, yr, yi, e);
    });
  });
  static_for<0, NR>([&](auto kc) {
    constexpr int k = decltype(kc)::value, j = ps.pj[P][k], mb = (j - d) / 2, NA = row_na(j, mb);
    static_for<0, NA>([&](auto mc) {
      constexpr int ma = decltype(mc)::value;
      yr[k][ma] = half_sum(yr[k][ma]);
      yi[k][ma] = half_sum(yi[k][ma]);
    });
    if (!store) return;
    float2* Yrow = yb + (ub_of(j) + mb * (j + 1)) * AB;
    // half h stores the columns ma = 2k + h, weighted as in unit()
    static_for<0, (NA + 1) / 2>([&](auto tc) {
      constexpr int m0 = 2 * decltype(tc)::value;
      constexpr double w0 = wadj::seed_weight(j, mb, m0);
      if constexpr (m0 + 1 < NA) {
        constexpr double w1 = wadj::seed_weight(j, mb, m0 + 1);
        const float2 v = h ? make_float2((float)(w1 * yr[k][m0 + 1]), (float)(w1 * yi[k][m0 + 1]))
                           : make_float2((float)(w0 * yr[k][m0]), (float)(w0 * yi[k][m0]));
        Yrow[(m0 + h) * AB] = v;
      } else {
        if (h == 0) Yrow[m0 * AB] = make_float2((float)(w0 * yr[k][m0]), (float)(w0 * yi[k][m0]));
      }
    });
  });
}

}  // namespace zy

// DESC = false: energy per atom and Y (adjoint).  DESC = true: Z in the
// per-atom layout Zout[a][idxz_max] (float2) and B in Bout[a][ncoeff] (FP64).
template <int J, int NW, bool DESC>
__global__ __launch_bounds__(64 * NW) void k_snap_zy32(SnapDev P, int64_t n, const int32_t* __restrict__ species,
                                                       const float2* __restrict__ Ub,
                                                       const double* __restrict__ beta, float2* __restrict__ Yb,
                                                       double* __restrict__ eatom, float2* __restrict__ Zout,
                                                       double* __restrict__ Bout) {
  constexpr int IDXU = ub_of(J + 1);
  constexpr zy::Plan<J, NW> plan{};
  constexpr zy::BSlots<J> bsl{};
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float2* UL = reinterpret_cast<float2*>(smem_raw);
  double* esum = reinterpret_cast<double*>(UL + IDXU * zy::AB);
  double* ycL = esum + NW * zy::AB;                       // Y coefficients [nspecies][nblk]
  double* btL = ycL + (DESC ? 0 : P.nspecies * P.nblk);   // beta [nspecies][ncoeff]
  double* BpL = btL + (DESC ? 0 : P.nspecies * P.ncoeff);  // B row sums [slot][atom] (DESC)
  float* cgL = reinterpret_cast<float*>(BpL + (DESC ? bsl.nslot * zy::AB : 0));
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int h = (tid >> 5) & 1, at = tid & 31;
  const int64_t batch = blockIdx.x;
  const int64_t a = batch * zy::AB + at;
  const bool valid = a < n;
  const float2* ug = Ub + batch * (int64_t)IDXU * zy::AB;
  for (int e = tid; e < IDXU * zy::AB; e += 64 * NW) UL[e] = ug[e];
  for (int e = tid; e < P.idxcg_max; e += 64 * NW) cgL[e] = P.cg_f[e];
  if constexpr (!DESC) {
    for (int e = tid; e < P.nspecies * P.ncoeff; e += 64 * NW) btL[e] = beta[e];
    // Y coefficient of each block and species: the beta entries whose dB/dr
    // half-sums use the block, with their multiplicities
    for (int e = tid; e < P.nspecies * P.nblk; e += 64 * NW) {
      const int sp = e / P.nblk;
      const YBlk* yb = P.yblk + e % P.nblk;
      double c = 0.0;
      for (int k = 0; k < yb->n; k++) c += beta[sp * P.ncoeff + yb->jjb[k]] * yb->fac[k];
      ycL[e] = c;
    }
  }
  __syncthreads();
  const int s = valid ? snap_species(species, a, P.nspecies) : 0;
  const double* ycs = ycL + s * P.nblk;
  const double* bts = btL + s * P.ncoeff;
  float2* yb = DESC ? nullptr : Yb + batch * (int64_t)IDXU * zy::AB + at;
  float2* zat = DESC ? Zout + (valid ? a : 0) * (int64_t)bsl.nz : nullptr;
  double e = 0.0;
  if constexpr (DESC) {
    static_for<0, zy::Plan<J, NW>::NU>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      constexpr int j = plan.uj[u], mb = plan.umb[u];
      if (w == plan.owner[u]) zy::unit<J, j, mb, DESC>(UL, cgL, h, at, ycs, bts, P, e, nullptr, valid, zat, BpL + at);
    });
  } else {
    static_for<0, zy::PiecePlanOf<J, NW>::v.ps.np>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      if (w == zy::PiecePlanOf<J, NW>::v.owner[p]) zy::piece<J, NW, p>(UL, cgL, h, at, ycs, bts, P, e, yb, valid);
    });
  }
  if constexpr (DESC) {
    __syncthreads();
    // B_jjb = sum over its row slots, in row order, minus bzero
    for (int t = tid; t < zy::AB * bsl.NBC; t += 64 * NW) {
      const int ab = t % zy::AB, jb = t / zy::AB;
      const int64_t aa = batch * zy::AB + ab;
      if (aa >= n) continue;
      double b = 0.0;
      for (int r = 0; r <= bsl.lev[jb] / 2; r++) b += BpL[(bsl.slot0[jb] + r) * zy::AB + ab];
      if (P.bzero) b -= P.bzero_j[bsl.lev[jb]];
      Bout[aa * bsl.NBC + jb] = b;
    }
    return;
  }
  e = zy::half_sum(e);
  if (h == 0) esum[w * zy::AB + at] = e;
  __syncthreads();
  if (w == 0 && h == 0 && valid) {
    double t = 0.0;
    for (int k = 0; k < NW; k++) t += esum[k * zy::AB + at];
    eatom[a] = t;
  }
}

template <int J, bool DESC>
void launch_zy_j(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const float* U, const double* beta,
                        float* Y, double* eatom, float* Z, double* B) {
  constexpr int NU = zy::num_units(J);
  constexpr int NW = NU < 12 ? NU : 12;
  constexpr zy::BSlots<J> bsl{};
  const size_t sm = sizeof(float2) * (size_t)ub_of(J + 1) * zy::AB + sizeof(double) * NW * zy::AB +
                    (DESC ? sizeof(double) * (size_t)bsl.nslot * zy::AB
                          : sizeof(double) * (size_t)P.nspecies * (P.nblk + P.ncoeff)) +
                    sizeof(float) * (size_t)P.idxcg_max;
  const unsigned grid = (unsigned)((R.n + zy::AB - 1) / zy::AB);
  if (sm > 64 * 1024)
    IPG_HIP(hipFuncSetAttribute((const void*)k_snap_zy32<J, NW, DESC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)sm));
  k_snap_zy32<J, NW, DESC><<<grid, 64 * NW, sm, ctx->stream>>>(
      P, R.n, R.species, reinterpret_cast<const float2*>(U), beta, reinterpret_cast<float2*>(Y), eatom,
      reinterpret_cast<float2*>(Z), B);
}

// launch_zy_j<J, DESC> is instantiated in ipg_snap_zy_j*.hip, split so that
// the largest kernels compile in parallel.
#define IPG_ZY_LAUNCHER(KW, J, DESC)                                                                             \
  KW void launch_zy_j<J, DESC>(ipg_ctx*, const SnapDev&, const SnapRun&, const float*, const double*, float*,     \
                               double*, float*, double*);

}  // namespace ipg
