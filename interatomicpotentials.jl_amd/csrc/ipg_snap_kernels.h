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

namespace ipg {

__host__ __device__ constexpr int ub_of(int j) { return j * (j + 1) * (2 * j + 1) / 6; }  // idxu_block[j]

template <typename T>
__device__ __forceinline__ T shfl_xor_s(T v, int o) {
  return __shfl_xor(v, o, 64);
}
template <typename T>
__device__ __forceinline__ Dual<T> shfl_xor_s(Dual<T> v, int o) {
  return Dual<T>(__shfl_xor(v.v, o, 64), __shfl_xor(v.t, o, 64));
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
__device__ __forceinline__ void cayley_klein(S x, S y, S z, double rcut, double wj, const SnapDev& P, CK<S>& c) {
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

// sqrt(p/q) (prebuilt_arrays.jl:314-321); the arguments are compile-time
// constants after unrolling, so this folds to an immediate.
template <typename T>
__device__ __forceinline__ T rootpq(int p, int q) {
  return (T)__builtin_sqrt((double)p / (double)q);
}

// One recursion step, in place: level j row mb from level j-1 row mb
// (compute_ui.jl:71-100, :227-267).  Descending ma keeps the old values of
// ma and ma-1 available.
template <int J, typename S, int NK>
__device__ __forceinline__ void wigner_step(const int j, const int mb, const CK<S>& c, S (&ur)[J + 1],
                                            S (&ui)[J + 1], S (&dur)[NK > 0 ? NK : 1][J + 1],
                                            S (&dui)[NK > 0 ? NK : 1][J + 1]) {
  using T = typename BaseOf<S>::type;
#pragma unroll
  for (int ma = J; ma >= 0; ma--) {
    if (ma > j) continue;
    S nr = S((T)0), ni = S((T)0);
    S ndr[NK > 0 ? NK : 1], ndi[NK > 0 ? NK : 1];
#pragma unroll
    for (int k = 0; k < NK; k++) {
      ndr[k] = S((T)0);
      ndi[k] = S((T)0);
    }
    if (ma >= 1) {
      const T rb = rootpq<T>(ma, j - mb);
      const S pr = ur[ma - 1], pi = ui[ma - 1];
      nr = -rb * (c.b_r * pr + c.b_i * pi);
      ni = -rb * (c.b_r * pi - c.b_i * pr);
#pragma unroll
      for (int k = 0; k < NK; k++) {
        ndr[k] = -rb * (c.db_r[k] * pr + c.db_i[k] * pi + c.b_r * dur[k][ma - 1] + c.b_i * dui[k][ma - 1]);
        ndi[k] = -rb * (c.db_r[k] * pi - c.db_i[k] * pr + c.b_r * dui[k][ma - 1] - c.b_i * dur[k][ma - 1]);
      }
    }
    if (ma <= j - 1) {
      const T ra = rootpq<T>(j - ma, j - mb);
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
  }
}

// Visits every unique left-half entry (level j, row mb, column ma with
// 2mb < j, or 2mb == j and ma <= mb) once, calling
//   emit(j, mb, ma, half, u_r, u_i, du_r[3], du_i[3])
// where half = 1 marks the centre of an even level (weight 1/2 in the
// half-column sums of compute_b.jl) and du is the raw recursion derivative
// (before the switching-function scaling of compute_ui.jl:300-325).
template <int J, typename S, int NK, typename Emit>
__device__ __forceinline__ void wigner_chain(const CK<S>& c, Emit&& emit) {
  using T = typename BaseOf<S>::type;
  constexpr int NKA = NK > 0 ? NK : 1;
  S ur[J + 1], ui[J + 1], dur[NKA][J + 1], dui[NKA][J + 1];
  S sr[J + 1], si[J + 1], sdr[NKA][J + 1], sdi[NKA][J + 1];
#pragma unroll
  for (int mb = 0; 2 * mb <= J; mb++) {
    if (mb == 0) {
      ur[0] = S((T)1);
      ui[0] = S((T)0);
#pragma unroll
      for (int k = 0; k < NK; k++) {
        dur[k][0] = S((T)0);
        dui[k][0] = S((T)0);
      }
      emit(0, 0, 0, 1, ur[0], ui[0], dur, dui, 0);
    } else {
      // row mb of level 2mb-1 = (-1)^(mb+x) conj(row mb-1 of level 2mb-1)[2mb-1-x]
#pragma unroll
      for (int x = 0; x < 2 * mb; x++) {
        const bool neg = (mb + x) & 1;
        const int y = 2 * mb - 1 - x;
        ur[x] = neg ? -sr[y] : sr[y];
        ui[x] = neg ? si[y] : -si[y];
#pragma unroll
        for (int k = 0; k < NK; k++) {
          dur[k][x] = neg ? -sdr[k][y] : sdr[k][y];
          dui[k][x] = neg ? sdi[k][y] : -sdi[k][y];
        }
      }
      wigner_step<J, S, NK>(2 * mb, mb, c, ur, ui, dur, dui);
      // middle row of the even level 2mb
#pragma unroll
      for (int ma = mb + 1; ma <= 2 * mb; ma++) {
        const bool neg = (mb + ma) & 1;
        const int y = 2 * mb - ma;
        ur[ma] = neg ? -ur[y] : ur[y];
        ui[ma] = neg ? ui[y] : -ui[y];
#pragma unroll
        for (int k = 0; k < NK; k++) {
          dur[k][ma] = neg ? -dur[k][y] : dur[k][y];
          dui[k][ma] = neg ? dui[k][y] : -dui[k][y];
        }
      }
      ui[mb] = -ui[mb];
#pragma unroll
      for (int k = 0; k < NK; k++) dui[k][mb] = -dui[k][mb];
#pragma unroll
      for (int ma = 0; ma <= mb; ma++) emit(2 * mb, mb, ma, ma == mb ? 1 : 0, ur[ma], ui[ma], dur, dui, ma);
    }
#pragma unroll
    for (int j = 2 * mb + 1; j <= J; j++) {
      wigner_step<J, S, NK>(j, mb, c, ur, ui, dur, dui);
#pragma unroll
      for (int ma = 0; ma <= j; ma++) emit(j, mb, ma, 0, ur[ma], ui[ma], dur, dui, ma);
      if (j == 2 * mb + 1 && 2 * (mb + 1) <= J) {
#pragma unroll
        for (int x = 0; x <= j; x++) {
          sr[x] = ur[x];
          si[x] = ui[x];
#pragma unroll
          for (int k = 0; k < NK; k++) {
            sdr[k][x] = dur[k][x];
            sdi[k][x] = dui[k][x];
          }
        }
      }
    }
  }
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

// ---------------------------------------------------------------------------
// K1: ulisttot (compute_ui.jl:5-44, zero_uarraytot :131-152, add_uarraytot
// :155-175).  One wave holds APW atoms; the G = 64/APW lanes of an atom take
// one neighbor each (looping over chunks of G neighbors) and every emitted
// entry is summed over the group with a fixed butterfly, so the result is
// deterministic.  The right half is filled from the left half by the
// inversion symmetry, and w_self sits on the diagonal.
// ---------------------------------------------------------------------------
template <int J, typename S, int APW>
__global__ __launch_bounds__(64) void k_snap_ui(SnapDev P, SnapRun R, S* __restrict__ ulisttot) {
  using T = typename BaseOf<S>::type;
  constexpr int G = 64 / APW;
  constexpr int IDXU = ub_of(J + 1);
  __shared__ S acc[APW][IDXU][2];
  const int lane = threadIdx.x;
  const int g = lane / G, l = lane % G;
  const int64_t a = (int64_t)blockIdx.x * APW + g;
  const bool valid = a < R.n;
  for (int e = l; e < IDXU; e += G) {
    acc[g][e][0] = S((T)0);
    acc[g][e][1] = S((T)0);
  }
  __syncthreads();
  const int64_t q0 = valid ? R.off[a] : 0, q1 = valid ? R.off[a + 1] : 0;
  int nch = (int)((q1 - q0 + G - 1) / G);
  for (int o = 1; o < 64; o <<= 1) nch = max(nch, __shfl_xor(nch, o, 64));
  const int si = valid && R.species ? R.species[a] : 0;
  for (int ch = 0; ch < nch; ch++) {
    const int64_t q = q0 + (int64_t)ch * G + l;
    const bool act = q < q1;
    CK<S> c;
    if (act) {
      const int32_t j = R.nbr[q];
      const int sj = R.species ? R.species[j] : 0;
      S x, y, z;
      load_pair<S>(R, q, (int32_t)a, j, x, y, z);
      const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
      cayley_klein<S, false>(x, y, z, rcut, P.weight[sj], P, c);
    } else {
      c.a_r = S((T)1);
      c.a_i = c.b_r = c.b_i = S((T)0);
      c.sfac = S((T)0);
    }
    const S w = c.sfac;
    S dz[1][J + 1];
    (void)dz;
    wigner_chain<J, S, 0>(c, [&](int j, int mb, int ma, int, S u_r, S u_i, auto&, auto&, int) {
      S vr = w * u_r, vi = w * u_i;
#pragma unroll
      for (int o = 1; o < G; o <<= 1) {
        vr += shfl_xor_s(vr, o);
        vi += shfl_xor_s(vi, o);
      }
      const int jju = ub_of(j) + (j + 1) * mb + ma;
      if (l == jju % G) {
        acc[g][jju][0] += vr;
        acc[g][jju][1] += vi;
      }
    });
  }
  __syncthreads();
  if (!valid) return;
  S* out = ulisttot + (int64_t)a * IDXU * 2;
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
    out[2 * e] = vr;
    out[2 * e + 1] = vi;
  }
}

// ---------------------------------------------------------------------------
// K2: Z (compute_z.jl:1-62), B (compute_b.jl:67-157) and the adjoint
// Y = dE/dU for one atom per 256-thread block.  U and Z live in LDS.
// Y_l(e) = sum over Z blocks (a,c,l) of coef(a,c,l) Z_(a,c,l)(e), with coef
// collecting the beta entries whose dB/dr half-sums use that block.
// ---------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_snap_zby(SnapDev P, SnapRun R, const S* __restrict__ ulisttot,
                                                  const double* __restrict__ beta,
                                                  const typename BaseOf<S>::type* __restrict__ ycoef,
                                                  double* __restrict__ Bout, double* __restrict__ eatom,
                                                  S* __restrict__ Yout, S* __restrict__ Zout,
                                                  double* __restrict__ tanB) {
  using T = typename BaseOf<S>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  S* U = reinterpret_cast<S*>(smem_raw);
  S* Z = U + 2 * P.idxu_max;
  __shared__ double red[4];
  const int tid = threadIdx.x;
  const int64_t a = blockIdx.x;
  const int s = R.species ? R.species[a] : 0;
  const S* ug = ulisttot + a * (int64_t)P.idxu_max * 2;
  for (int e = tid; e < 2 * P.idxu_max; e += 256) U[e] = ug[e];
  __syncthreads();
  const T* cgt;
  if constexpr (std::is_same<T, float>::value) cgt = P.cg_f;
  else cgt = P.cg_d;
  for (int jjz = tid; jjz < P.idxz_max; jjz += 256) {
    const ZEnt ze = P.zent[jjz];
    const T* cg = cgt + ze.cgoff;
    S zr = S((T)0), zi = S((T)0);
    int jju1 = P.idxu_block[ze.j1] + (ze.j1 + 1) * ze.mb1min;
    int jju2 = P.idxu_block[ze.j2] + (ze.j2 + 1) * ze.mb2max;
    int icgb = ze.mb1min * (ze.j2 + 1) + ze.mb2max;
    for (int ib = 0; ib < ze.nb; ib++) {
      S sr = S((T)0), si_ = S((T)0);
      int ma1 = ze.ma1min, ma2 = ze.ma2max;
      int icga = ze.ma1min * (ze.j2 + 1) + ze.ma2max;
      for (int ia = 0; ia < ze.na; ia++) {
        const S u1r = U[2 * (jju1 + ma1)], u1i = U[2 * (jju1 + ma1) + 1];
        const S u2r = U[2 * (jju2 + ma2)], u2i = U[2 * (jju2 + ma2) + 1];
        const T cga = cg[icga];
        sr += cga * (u1r * u2r - u1i * u2i);
        si_ += cga * (u1r * u2i + u1i * u2r);
        ma1++;
        ma2--;
        icga += ze.j2;
      }
      const T cgb = cg[icgb];
      zr += cgb * sr;
      zi += cgb * si_;
      jju1 += ze.j1 + 1;
      jju2 -= ze.j2 + 1;
      icgb += ze.j2;
    }
    if (P.bnorm) {
      zr = zr / (T)(ze.j + 1);
      zi = zi / (T)(ze.j + 1);
    }
    Z[2 * jjz] = zr;
    Z[2 * jjz + 1] = zi;
  }
  __syncthreads();
  if (Zout) {
    S* zg = Zout + a * (int64_t)P.idxz_max * 2;
    for (int e = tid; e < 2 * P.idxz_max; e += 256) zg[e] = Z[e];
  }
  // B and the per-atom energy beta_s . B
  double epart = 0.0;
  const int M = P.ncoeff * P.nspecies;
  if (tanB)
    for (int c = tid; c < M; c += 256) tanB[a * M + c] = 0.0;
  if (tanB) __syncthreads();
  for (int jjb = tid; jjb < P.idxb_max; jjb += 256) {
    const BTerm bt = P.bterm[jjb];
    const int j = bt.j;
    int jju = P.idxu_block[j], jjz = bt.jjz0[0];
    S sum = S((T)0);
    for (int mb = 0; 2 * mb < j; mb++)
      for (int ma = 0; ma <= j; ma++, jju++, jjz++) sum += U[2 * jju] * Z[2 * jjz] + U[2 * jju + 1] * Z[2 * jjz + 1];
    if (j % 2 == 0) {
      const int mb = j / 2;
      for (int ma = 0; ma < mb; ma++, jju++, jjz++) sum += U[2 * jju] * Z[2 * jjz] + U[2 * jju + 1] * Z[2 * jjz + 1];
      sum += (T)0.5 * (U[2 * jju] * Z[2 * jjz] + U[2 * jju + 1] * Z[2 * jjz + 1]);
    }
    S b = (T)2 * sum;
    if (P.bzero) b = b - (T)P.bzero_j[j];
    const double bv = value_of(b);
    if (Bout) Bout[a * P.ncoeff + jjb] = bv;
    if (beta) epart += beta[s * P.ncoeff + jjb] * bv;
    if (tanB) tanB[a * M + s * P.ncoeff + jjb] = tangent_of(b);
  }
  if (eatom) {
    for (int o = 32; o > 0; o >>= 1) epart += __shfl_xor(epart, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = epart;
    __syncthreads();
    if (tid == 0) eatom[a] = ((red[0] + red[1]) + red[2]) + red[3];
  }
  if (Yout) {
    S* yg = Yout + a * (int64_t)P.idxu_max * 2;
    const T* yc = ycoef + (int64_t)s * P.nblk;
    for (int h = tid; h < P.nhalf; h += 256) {
      S yr = S((T)0), yi = S((T)0);
      for (int k = P.yoff[h]; k < P.yoff[h + 1]; k++) {
        const int jjz = P.yjjz[k];
        const T cf = yc[P.zent[jjz].blk];
        yr += cf * Z[2 * jjz];
        yi += cf * Z[2 * jjz + 1];
      }
      const int jju = P.half_jju[h];
      yg[2 * jju] = yr;
      yg[2 * jju + 1] = yi;
    }
  }
}

// ---------------------------------------------------------------------------
// K3: adjoint force per pair.  dE_i/dr_ij = 2 w_j sum_half w(e)
// Re(conj(dsfac u e_r + sfac du) Y_i), one thread per pair, the u/du chain
// in registers (no per-pair dB is formed).
// ---------------------------------------------------------------------------
template <int J, typename S>
__global__ __launch_bounds__(256) void k_snap_dedr(SnapDev P, SnapRun R, const S* __restrict__ Y,
                                                   double* __restrict__ dedr, double* __restrict__ dedr_t) {
  using T = typename BaseOf<S>::type;
  constexpr int IDXU = ub_of(J + 1);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= R.npairs) return;
  const int32_t i = R.center[q], j = R.nbr[q];
  const int si = R.species ? R.species[i] : 0, sj = R.species ? R.species[j] : 0;
  S x, y, z;
  load_pair<S>(R, q, i, j, x, y, z);
  const double rcut = (P.radii[si] + P.radii[sj]) * P.rcutfac;
  CK<S> c;
  cayley_klein<S, true>(x, y, z, rcut, P.weight[sj], P, c);
  const S* Yi = Y + (int64_t)i * IDXU * 2;
  S s0 = S((T)0), s1 = S((T)0), s2 = S((T)0), s3 = S((T)0);
  wigner_chain<J, S, 3>(c, [&](int jl, int mb, int ma, int half, S u_r, S u_i, auto& dur, auto& dui, int col) {
    const int jju = ub_of(jl) + (jl + 1) * mb + ma;
    S yr = Yi[2 * jju], yi = Yi[2 * jju + 1];
    if (half) {
      yr = (T)0.5 * yr;
      yi = (T)0.5 * yi;
    }
    s0 += u_r * yr + u_i * yi;
    s1 += dur[0][col] * yr + dui[0][col] * yi;
    s2 += dur[1][col] * yr + dui[1][col] * yi;
    s3 += dur[2][col] * yr + dui[2][col] * yi;
  });
  const S g0 = c.dsfac * s0;
  S d0 = (T)2 * (g0 * c.u[0] + c.sfac * s1);
  S d1 = (T)2 * (g0 * c.u[1] + c.sfac * s2);
  S d2 = (T)2 * (g0 * c.u[2] + c.sfac * s3);
  dedr[3 * q] = value_of(d0);
  dedr[3 * q + 1] = value_of(d1);
  dedr[3 * q + 2] = value_of(d2);
  if (dedr_t) {
    dedr_t[3 * q] = tangent_of(d0);
    dedr_t[3 * q + 1] = tangent_of(d1);
    dedr_t[3 * q + 2] = tangent_of(d2);
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
  const double avg = (double)R.npairs / (double)R.n;
  const int apw = avg <= 24.0 ? 2 : 1;
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
  if (R.n == 0) return;
  size_t sm = sizeof(S) * 2 * (size_t)(P.idxu_max + P.idxz_max);
  const char* nm = IsDual<S>::value ? "snap_zby_dual" : (sizeof(S) == 4 ? "snap_zby_fp32" : "snap_zby_fp64");
  launch(ctx, nm, [&] {
    if (sm > 64 * 1024)
      IPG_HIP(hipFuncSetAttribute((const void*)k_snap_zby<S>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    k_snap_zby<S><<<(unsigned)R.n, 256, sm, ctx->stream>>>(P, R, ulisttot, beta, ycoef, Bout, eatom, Yout, Zout,
                                                           tanB);
  });
}

template <typename S>
void snap_launch_dedr(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const S* Y, double* dedr, double* dedr_t) {
  if (R.npairs == 0) return;
  unsigned g = grid_for(R.npairs, 256);
  const char* nm = IsDual<S>::value ? "snap_dedr_dual" : (sizeof(S) == 4 ? "snap_dedr_fp32" : "snap_dedr_fp64");
  launch(ctx, nm, [&] {
    switch (P.J) {
      case 0: k_snap_dedr<0, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 1: k_snap_dedr<1, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 2: k_snap_dedr<2, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 3: k_snap_dedr<3, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 4: k_snap_dedr<4, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 5: k_snap_dedr<5, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 6: k_snap_dedr<6, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      case 7: k_snap_dedr<7, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
      default: k_snap_dedr<8, S><<<g, 256, 0, ctx->stream>>>(P, R, Y, dedr, dedr_t); break;
    }
  });
}

#define IPG_SNAP_INSTANTIATE(S)                                                                                  \
  template void snap_launch_ui<S>(ipg_ctx*, const SnapDev&, const SnapRun&, S*);                                 \
  template void snap_launch_zby<S>(ipg_ctx*, const SnapDev&, const SnapRun&, const S*, const double*,            \
                                   const typename BaseOf<S>::type*, double*, double*, S*, S*, double*);          \
  template void snap_launch_dedr<S>(ipg_ctx*, const SnapDev&, const SnapRun&, const S*, double*, double*);

}  // namespace ipg
