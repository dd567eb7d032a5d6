/*
 * ipg_oracle.c -- CPU FP64 restatement of InteratomicPotentials.jl's hot path.
 *
 * TEST INFRASTRUCTURE ONLY (correctness oracle and CPU baseline).  The
 * product library never links or calls this file.  Each function names the
 * reference file:line it follows.  Build with -ffp-contract=off.
 */
#include "ipg_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ================================================================== */
/* distance semantics                                                 */
/* ================================================================== */

/* Julia Base.mod(x::Float64, y::Float64): r = rem(x, y) (exact, C fmod);
 * r == 0 -> copysign(r, y); sign(r) != sign(y) -> r + y. */
double orc_jl_mod(double x, double y) {
  double r = fmod(x, y);
  if (r == 0.0) return copysign(r, y);
  if ((r > 0) != (y > 0)) return r + y;
  return r;
}

/* Distances.jl 0.10 PeriodicEuclidean (used at src/nnlist.jl:30, utl.jl:65):
 * per dimension s1 = |a-b|, s2 = mod(s1, p), s3 = min(s2, p - s2), summed
 * as ((0 + s3x^2) + s3y^2) + s3z^2.  For p = Inf, mod(s1, Inf) = s1 and
 * min(s1, Inf) = s1. */
double orc_pe_sumsq(const double a[3], const double b[3], const double L[3]) {
  double s = 0.0;
  for (int d = 0; d < 3; d++) {
    double s1 = fabs(a[d] - b[d]);
    double s3;
    if (isinf(L[d])) {
      s3 = s1;
    } else {
      double s2 = orc_jl_mod(s1, L[d]);
      double s2c = L[d] - s2;
      s3 = (s2 < s2c) ? s2 : s2c;
    }
    s = s + s3 * s3;
  }
  return s;
}

/* src/nnlist.jl:18-21 get_displacement (identical to utl.jl:8-17
 * get_distance): d = mod(x - y, L); isinf(L) ? x - y : (2d < L ? d : d - L). */
void orc_get_displacement(const double L[3], const double x[3], const double y[3], double out[3]) {
  for (int k = 0; k < 3; k++) {
    double diff = x[k] - y[k];
    if (isinf(L[k])) {
      out[k] = diff;
    } else {
      double d = orc_jl_mod(diff, L[k]);
      out[k] = (2.0 * d < L[k]) ? d : d - L[k];
    }
  }
}

double orc_wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ================================================================== */
/* neighbor lists                                                     */
/* ================================================================== */

/* Membership of src/nnlist.jl:40: inrange(tree, X[n], rcut) keeps m with
 * evaluate(PeriodicEuclidean(L), X[n], X[m]) <= rcut, i.e. sqrt(sumsq) <= rcut. */
static int pe_within(const double* a, const double* b, const double L[3], double rcut) {
  return sqrt(orc_pe_sumsq(a, b, L)) <= rcut;
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* src/nnlist.jl:23-46 neighborlist(s, rcutoff): half list j > i (kind HALF)
 * or full list j != i (kind FULL), r = get_displacement(L, X[i], X[j]). */
int64_t orc_nlist_brute(const double* pos, int64_t n, const double L[3], double rcut, int kind,
                        int64_t* offsets, int32_t* nbr, double* disp, int64_t cap) {
  if (n < 0 || !(rcut > 0)) return -1;
  int64_t np = 0;
  for (int64_t i = 0; i < n; i++) {
    offsets[i] = np;
    int64_t j0 = (kind == ORC_LIST_HALF) ? i + 1 : 0;
    for (int64_t j = j0; j < n; j++) {
      if (j == i) continue;
      if (pe_within(pos + 3 * i, pos + 3 * j, L, rcut)) {
        if (np < cap) {
          if (nbr) nbr[np] = (int32_t)j;
          if (disp) orc_get_displacement(L, pos + 3 * i, pos + 3 * j, disp + 3 * np);
        }
        np++;
      }
    }
  }
  offsets[n] = np;
  return np;
}

/* Cell-binned restatement: atoms are binned on wrapped coordinates with cells
 * at least rcut*(1+1e-9) wide, candidates come from the 27 surrounding cells
 * (fewer when a dimension has < 3 cells), and membership uses the same
 * PeriodicEuclidean formula on the original coordinates.  Output equals
 * orc_nlist_brute (sorted j). */
/* Hits of atom i in the cell grid (the candidates of the 27 surrounding
 * cells, or every cell of a periodic dimension with fewer than 3), sorted;
 * tmp has room for n entries.  Returns the count. */
typedef struct {
  const double* pos;
  const double* L;
  double rcut;
  int kind, nc[3];
  const int32_t* cell;
  const int64_t* cstart;
  const int32_t* cat;
} orc_grid;

static int64_t cells_row(const orc_grid* g, int64_t i, int32_t* tmp) {
  const int* nc = g->nc;
  const double* L = g->L;
  int c0 = g->cell[i] % nc[0], c1 = (g->cell[i] / nc[0]) % nc[1], c2 = g->cell[i] / (nc[0] * nc[1]);
  int lo_[3], hi_[3], cc[3] = {c0, c1, c2};
  for (int d = 0; d < 3; d++) {
    if (nc[d] >= 3 || isinf(L[d])) { lo_[d] = cc[d] - 1; hi_[d] = cc[d] + 1; }
    else { lo_[d] = 0; hi_[d] = nc[d] - 1; }
  }
  int64_t cnt = 0;
  for (int z = lo_[2]; z <= hi_[2]; z++) {
    int zz = z;
    if (isinf(L[2])) { if (zz < 0 || zz >= nc[2]) continue; }
    else if (nc[2] >= 3) zz = (zz + nc[2]) % nc[2];
    for (int y = lo_[1]; y <= hi_[1]; y++) {
      int yy = y;
      if (isinf(L[1])) { if (yy < 0 || yy >= nc[1]) continue; }
      else if (nc[1] >= 3) yy = (yy + nc[1]) % nc[1];
      for (int x = lo_[0]; x <= hi_[0]; x++) {
        int xx = x;
        if (isinf(L[0])) { if (xx < 0 || xx >= nc[0]) continue; }
        else if (nc[0] >= 3) xx = (xx + nc[0]) % nc[0];
        int64_t cid = ((int64_t)zz * nc[1] + yy) * nc[0] + xx;
        for (int64_t s = g->cstart[cid]; s < g->cstart[cid + 1]; s++) {
          int32_t j = g->cat[s];
          if (j == i) continue;
          if (g->kind == ORC_LIST_HALF && j < i) continue;
          if (pe_within(g->pos + 3 * i, g->pos + 3 * j, L, g->rcut)) tmp[cnt++] = j;
        }
      }
    }
  }
  qsort(tmp, (size_t)cnt, sizeof(int32_t), cmp_i32);
  return cnt;
}

/* Cell-binned list; rows are independent, so nthreads > 1 (OpenMP) counts
 * and fills them in parallel with the same result. */
static int64_t nlist_cells_mt(const double* pos, int64_t n, const double L[3], double rcut, int kind,
                              int64_t* offsets, int32_t* nbr, double* disp, int64_t cap, int nthreads) {
  if (n < 0 || !(rcut > 0)) return -1;
  if (n == 0) { offsets[0] = 0; return 0; }
  if (nthreads < 1) nthreads = 1;
  double lo[3], w[3];
  int nc[3];
  for (int d = 0; d < 3; d++) {
    double ext, l0;
    if (isinf(L[d])) {
      double mn = pos[d], mx = pos[d];
      for (int64_t i = 1; i < n; i++) {
        double v = pos[3 * i + d];
        if (v < mn) mn = v;
        if (v > mx) mx = v;
      }
      l0 = mn; ext = mx - mn;
    } else {
      l0 = 0.0; ext = L[d];
    }
    double cw = rcut * (1.0 + 1e-9);
    int c = (int)floor(ext / cw);
    if (c < 1) c = 1;
    if (c > 1024) c = 1024;
    nc[d] = c; lo[d] = l0;
    w[d] = (c == 1) ? (ext > 0 ? ext : 1.0) : ext / c;
  }
  int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
  int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * n);
  int64_t* cstart = (int64_t*)calloc(ncell + 1, sizeof(int64_t));
  int32_t* cat = (int32_t*)malloc(sizeof(int32_t) * n);
  for (int64_t i = 0; i < n; i++) {
    int cc[3];
    for (int d = 0; d < 3; d++) {
      double x = pos[3 * i + d];
      double u = isinf(L[d]) ? x - lo[d] : x - floor(x / L[d]) * L[d];
      int c = (int)floor(u / w[d]);
      if (c < 0) c = 0;
      if (c >= nc[d]) c = nc[d] - 1;
      cc[d] = c;
    }
    cell[i] = (int32_t)((cc[2] * nc[1] + cc[1]) * nc[0] + cc[0]);
    cstart[cell[i] + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
  int64_t* fillp = (int64_t*)malloc(sizeof(int64_t) * ncell);
  memcpy(fillp, cstart, sizeof(int64_t) * ncell);
  for (int64_t i = 0; i < n; i++) cat[fillp[cell[i]]++] = (int32_t)i;
  free(fillp);

  orc_grid g = {pos, L, rcut, kind, {nc[0], nc[1], nc[2]}, cell, cstart, cat};
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * n * (size_t)nthreads);
  /* pass 1: row lengths */
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
  for (int64_t i = 0; i < n; i++) {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    offsets[i] = cells_row(&g, i, tmp + (size_t)t * n);
  }
  int64_t np = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t c = offsets[i];
    offsets[i] = np;
    np += c;
  }
  offsets[n] = np;
  /* pass 2: rows (entries past cap are dropped) */
  if (nbr || disp) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      int32_t* row = tmp + (size_t)t * n;
      int64_t cnt = cells_row(&g, i, row);
      for (int64_t k = 0; k < cnt; k++) {
        int64_t q = offsets[i] + k;
        if (q < cap) {
          if (nbr) nbr[q] = row[k];
          if (disp) orc_get_displacement(L, pos + 3 * i, pos + 3 * row[k], disp + 3 * q);
        }
      }
    }
  }
  free(tmp); free(cat); free(cstart); free(cell);
  return np;
}

int64_t orc_nlist_cells(const double* pos, int64_t n, const double L[3], double rcut, int kind,
                        int64_t* offsets, int32_t* nbr, double* disp, int64_t cap) {
  return nlist_cells_mt(pos, n, L, rcut, kind, offsets, nbr, disp, cap, 1);
}

/* ================================================================== */
/* pair potentials                                                    */
/* ================================================================== */

/* src/empirical/zbl.jl:27-28 */
static const double ZBL_C[4] = {0.18175, 0.50986, 0.28022, 0.02817};
static const double ZBL_X[4] = {3.19980, 0.94229, 0.40290, 0.20162};

static double zbl_phi(double d, double e) { /* zbl.jl:29 */
  double s = ZBL_C[0] * pow(e, -ZBL_X[0] * d);
  for (int k = 1; k < 4; k++) s = s + ZBL_C[k] * pow(e, -ZBL_X[k] * d);
  return s;
}
static double zbl_dphi(double d, double e) { /* zbl.jl:30 */
  double s = ZBL_X[0] * ZBL_C[0] * pow(e, -ZBL_X[0] * d);
  for (int k = 1; k < 4; k++) s = s + ZBL_X[k] * ZBL_C[k] * pow(e, -ZBL_X[k] * d);
  return -s;
}
static double zbl_coeff(const orc_pair_term* t) { /* zbl.jl:32, k_e = 1 (constants.jl:7) */
  return 1.0 * t->p[0] * t->p[1] * (t->p[2] * t->p[2]);
}
static double zbl_d(const orc_pair_term* t, double R) { /* zbl.jl:33 */
  return R / (0.8854 * 0.529 / (pow(t->p[0], 0.23) + pow(t->p[1], 0.23)));
}

double orc_pair_V(const orc_pair_term* t, double R) {
  const double* p = t->p;
  switch (t->kind) {
    case ORC_PAIR_LJ: { /* lj.jl:35 */
      double s = p[1] / R;
      return 4.0 * p[0] * (pow(s, 12) - pow(s, 6));
    }
    case ORC_PAIR_MORSE: { /* morse.jl:36,38 */
      double E = exp(-p[1] * (R - p[2]));
      double q = 1.0 - E;
      return p[0] * (q * q);
    }
    case ORC_PAIR_ZBL: /* zbl.jl:35 */
      return zbl_coeff(t) * zbl_phi(zbl_d(t, R), p[2]) / R;
    case ORC_PAIR_BM: /* bm.jl:38 */
      return p[0] * exp((p[2] - R) / p[1]) + pow(p[3] / R, 6) - pow(p[4] / R, 8);
    case ORC_PAIR_COULOMB: /* coulomb.jl:36 */
      return 1.0 * p[0] * p[1] / R;
    case ORC_PAIR_MOCK: /* test/mocks.jl:28-30 */
      return (R * R) * p[0];
  }
  return NAN;
}

void orc_pair_F(const orc_pair_term* t, double R, const double r[3], double out[3]) {
  const double* p = t->p;
  double s = NAN;
  switch (t->kind) {
    case ORC_PAIR_LJ: { /* lj.jl:36 */
      double q = p[1] / R;
      s = 24.0 * p[0] * (2.0 * pow(q, 12) - pow(q, 6)) / (R * R);
      break;
    }
    case ORC_PAIR_MORSE: { /* morse.jl:39 (sign as in the reference) */
      double E = exp(-p[1] * (R - p[2]));
      s = 2.0 * p[0] * p[1] * (1.0 - E) * E / R;
      break;
    }
    case ORC_PAIR_ZBL: { /* zbl.jl:36 */
      double d = zbl_d(t, R);
      s = zbl_coeff(t) * (zbl_phi(d, p[2]) + d * zbl_dphi(d, p[2])) / (R * R * R);
      break;
    }
    case ORC_PAIR_BM: { /* bm.jl:39 */
      s = p[0] * exp((p[2] - R) / p[1]) / (p[1] * R) + (pow(p[3] / R, 6) - pow(p[4] / R, 8)) / (R * R);
      break;
    }
    case ORC_PAIR_COULOMB: /* coulomb.jl:37 */
      s = 1.0 * p[0] * p[1] / (R * R * R);
      break;
    case ORC_PAIR_MOCK: /* test/mocks.jl:31-33 */
      s = (R * R) * p[0];
      break;
  }
  out[0] = s * r[0];
  out[1] = s * r[1];
  out[2] = s * r[2];
}

static double norm3(const double r[3]) { return sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]); }

/* src/types/empirical_potential.jl:20-67 (_perform_pairwise with the
 * energy_and_force and virial_stress closures) for each term, summed with
 * the term coefficient as linear_combination_potential.jl:27-39 does. */
/* empirical_potential.jl:20-67 for one term over its own half list: the
 * reference splits the atoms over threads with per-thread energy, force
 * and virial accumulators (:28-46); here thread t takes a contiguous block
 * of atoms and the per-thread sums are added in thread order, so a given
 * thread count always gives the same rounding. */
static void pair_term_mt(const orc_pair_term* t, int64_t n, const int32_t* species, const int64_t* off,
                         const int32_t* nbr, const double* disp, int nthreads, double* et_out, double* ft_out,
                         double vt_out[6]) {
  double* ftt = (double*)calloc((size_t)3 * n * nthreads, sizeof(double));
  double* etv = (double*)calloc((size_t)7 * nthreads, sizeof(double));
#pragma omp parallel num_threads(nthreads)
  {
    int tid = 0, nt = 1;
#ifdef _OPENMP
    tid = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    int64_t a0 = n * tid / nt, a1 = n * (tid + 1) / nt;
    double* ft = ftt + (size_t)3 * n * tid;
    double et = 0.0, vt[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = a0; i < a1; i++) {
      /* :36-38: a pair counts iff both species are in get_species(p) (the
       * fast path of :26 gives the same set). */
      if (!((t->species_mask >> species[i]) & 1ull)) continue;
      for (int64_t q = off[i]; q < off[i + 1]; q++) {
        int32_t j = nbr[q];
        if (!((t->species_mask >> species[j]) & 1ull)) continue;
        const double* r = disp + 3 * q;
        double R = norm3(r);             /* :17-18 potential_energy/force(r) */
        double fo[3];
        orc_pair_F(t, R, r, fo);
        et = et + orc_pair_V(t, R);      /* :51-56 */
        for (int d = 0; d < 3; d++) {
          ft[3 * i + d] += fo[d];
          ft[3 * j + d] -= fo[d];
        }
        /* :60-67 virial: r * F' in Voigt order [xx, yy, zz, zy, zx, yx] */
        vt[0] += r[0] * fo[0];
        vt[1] += r[1] * fo[1];
        vt[2] += r[2] * fo[2];
        vt[3] += r[2] * fo[1];
        vt[4] += r[2] * fo[0];
        vt[5] += r[1] * fo[0];
      }
    }
    etv[7 * tid] = et;
    for (int k = 0; k < 6; k++) etv[7 * tid + 1 + k] = vt[k];
  }
  *et_out = 0.0;
  for (int k = 0; k < 6; k++) vt_out[k] = 0.0;
  memset(ft_out, 0, sizeof(double) * 3 * n);
  for (int tid = 0; tid < nthreads; tid++) {
    *et_out += etv[7 * tid];
    for (int k = 0; k < 6; k++) vt_out[k] += etv[7 * tid + 1 + k];
    const double* ft = ftt + (size_t)3 * n * tid;
    for (int64_t a = 0; a < 3 * n; a++) ft_out[a] += ft[a];
  }
  free(ftt);
  free(etv);
}

int orc_pair_energy_force_virial_mt(const double* pos, int64_t n, const double L[3], const int32_t* species,
                                    const orc_pair_term* terms, int nterms, int use_cells, int nthreads,
                                    double* e, double* f, double* v) {
  if (n < 0 || nterms < 0) return -1;
  if (nthreads < 1) nthreads = 1;
  if (e) *e = 0.0;
  if (f) memset(f, 0, sizeof(double) * 3 * n);
  if (v) memset(v, 0, sizeof(double) * 6);
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
  double* ft = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
  for (int k = 0; k < nterms; k++) {
    /* linear_combination_potential.jl:27-39: every term rebuilds its list */
    const orc_pair_term* t = &terms[k];
    int64_t np = use_cells ? nlist_cells_mt(pos, n, L, t->rcut, ORC_LIST_HALF, off, NULL, NULL, 0, nthreads)
                           : orc_nlist_brute(pos, n, L, t->rcut, ORC_LIST_HALF, off, NULL, NULL, 0);
    if (np < 0) { free(off); free(ft); return -1; }
    int32_t* nbr = (int32_t*)malloc(sizeof(int32_t) * (np > 0 ? np : 1));
    double* disp = (double*)malloc(sizeof(double) * 3 * (np > 0 ? np : 1));
    if (use_cells) nlist_cells_mt(pos, n, L, t->rcut, ORC_LIST_HALF, off, nbr, disp, np, nthreads);
    else orc_nlist_brute(pos, n, L, t->rcut, ORC_LIST_HALF, off, nbr, disp, np);
    double et = 0.0, vt[6];
    pair_term_mt(t, n, species, off, nbr, disp, nthreads, &et, ft, vt);
    if (e) *e += t->coeff * et;
    if (f) for (int64_t a = 0; a < 3 * n; a++) f[a] += t->coeff * ft[a];
    if (v) for (int a = 0; a < 6; a++) v[a] += t->coeff * vt[a];
    free(nbr); free(disp);
  }
  free(off); free(ft);
  return 0;
}

int orc_pair_energy_force_virial(const double* pos, int64_t n, const double L[3], const int32_t* species,
                                 const orc_pair_term* terms, int nterms, int use_cells,
                                 double* e, double* f, double* v) {
  return orc_pair_energy_force_virial_mt(pos, n, L, species, terms, nterms, use_cells, 1, e, f, v);
}

/* ================================================================== */
/* SNAP tables (types/prebuilt_arrays.jl)                              */
/* ================================================================== */

int orc_snap_num_coeffs(int twojmax, int nspecies, int chem) { /* utl.jl:19-37 */
  int J = twojmax, nc;
  if (J % 2 == 0) { int m = J / 2 + 1; nc = m * (m + 1) * (2 * m + 1) / 6; }
  else { int m = (J + 1) / 2; nc = m * (m + 1) * (m + 2) / 3; }
  if (chem) nc *= nspecies * nspecies * nspecies;
  return nc;
}

static double factorial_d(int n) { /* Julia factorial(Int) -> exact Int64 (n <= 20) */
  int64_t r = 1;
  for (int k = 2; k <= n; k++) r *= k;
  return (double)r;
}
static int64_t factorial_i(int n) {
  int64_t r = 1;
  for (int k = 2; k <= n; k++) r *= k;
  return r;
}

/* prebuilt_arrays.jl:250-257 */
static double deltacg(int j1, int j2, int j) {
  int64_t sfaccg = factorial_i((j1 + j2 + j) / 2 + 1);
  int64_t num = factorial_i((j1 + j2 - j) / 2) * factorial_i((j1 - j2 + j) / 2) * factorial_i((-j1 + j2 + j) / 2);
  return sqrt((double)num / (double)sfaccg);
}

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

#define B3(t, a, b, c) ((((a) * ((t)->twojmax + 1)) + (b)) * ((t)->twojmax + 1) + (c))

/* prebuilt_arrays.jl:48-247 initialize_prebuilt_arrays */
orc_snap_tables* orc_snap_tables_build(int twojmax, int nspecies, int chem, int bzero_flag, int bnorm) {
  if (twojmax < 0 || twojmax > 16) return NULL;
  orc_snap_tables* t = (orc_snap_tables*)calloc(1, sizeof(orc_snap_tables));
  int J = twojmax, jdim = J + 1;
  t->twojmax = J;
  t->wself = 1.0;
  t->bzero = (double*)calloc(jdim, sizeof(double));
  if (bzero_flag) { /* :54-63 */
    double www = t->wself * t->wself * t->wself;
    for (int j = 0; j <= J; j++) t->bzero[j] = bnorm ? www : www * (j + 1);
  }
  t->num_elements = chem ? nspecies : 1;
  t->idxcg_block = (int*)calloc(jdim * jdim * jdim, sizeof(int));
  t->idxu_block = (int*)calloc(jdim, sizeof(int));
  t->idxb_block = (int*)calloc(jdim * jdim * jdim, sizeof(int));
  t->idxz_block = (int*)calloc(jdim * jdim * jdim, sizeof(int));
  int cnt = 0;
  for (int j1 = 0; j1 <= J; j1++) /* :76-91 */
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = j1 - j2; j <= imin(J, j1 + j2); j += 2) {
        t->idxcg_block[B3(t, j1, j2, j)] = cnt;
        cnt += (j1 + 1) * (j2 + 1);
      }
  t->idxcg_max = cnt;
  cnt = 0;
  for (int j = 0; j <= J; j++) { t->idxu_block[j] = cnt; cnt += (j + 1) * (j + 1); } /* :95-106 */
  t->idxu_max = cnt;
  cnt = 0;
  for (int j1 = 0; j1 <= J; j1++) /* :110-123 */
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = j1 - j2; j <= imin(J, j1 + j2); j += 2)
        if (j >= j1) cnt++;
  t->idxb_max = cnt;
  t->idxb = (int*)calloc(3 * (cnt > 0 ? cnt : 1), sizeof(int));
  cnt = 0;
  for (int j1 = 0; j1 <= J; j1++) /* :128-156 */
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = j1 - j2; j <= imin(J, j1 + j2); j += 2)
        if (j >= j1) {
          t->idxb[3 * cnt] = j1; t->idxb[3 * cnt + 1] = j2; t->idxb[3 * cnt + 2] = j;
          t->idxb_block[B3(t, j1, j2, j)] = cnt;
          cnt++;
        }
  cnt = 0;
  for (int j1 = 0; j1 <= J; j1++) /* :159-177 */
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = j1 - j2; j <= imin(J, j1 + j2); j += 2)
        for (int mb = 0; 2 * mb <= j; mb++) cnt += j + 1;
  t->idxz_max = cnt;
  t->idxz = (orc_zidx*)calloc(cnt > 0 ? cnt : 1, sizeof(orc_zidx));
  cnt = 0;
  for (int j1 = 0; j1 <= J; j1++) /* :183-216 */
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = j1 - j2; j <= imin(J, j1 + j2); j += 2) {
        t->idxz_block[B3(t, j1, j2, j)] = cnt;
        for (int mb = 0; 2 * mb <= j; mb++)
          for (int ma = 0; ma <= j; ma++) {
            orc_zidx* z = &t->idxz[cnt];
            z->j1 = j1; z->j2 = j2; z->j = j;
            /* Julia ÷ truncates toward zero, as C integer division does */
            z->ma1min = imax(0, (2 * ma - j - j2 + j1) / 2);
            z->ma2max = (2 * ma - j - (2 * z->ma1min - j1) + j2) / 2;
            z->na = imin(j1, (2 * ma - j + j2 + j1) / 2) - z->ma1min + 1;
            z->mb1min = imax(0, (2 * mb - j - j2 + j1) / 2);
            z->mb2max = (2 * mb - j - (2 * z->mb1min - j1) + j2) / 2;
            z->nb = imin(j1, (2 * mb - j + j2 + j1) / 2) - z->mb1min + 1;
            z->jju = t->idxu_block[j] + (j + 1) * mb + ma;
            cnt++;
          }
      }
  /* :314-321 init_rootpqarray */
  t->rootpq = (double*)calloc((J + 2) * (J + 2), sizeof(double));
  for (int p = 1; p <= J; p++)
    for (int q = 1; q <= J; q++) t->rootpq[p * (J + 2) + q] = sqrt((double)p / (double)q);
  /* :259-312 init_clebsch_gordan */
  t->cglist = (double*)calloc(t->idxcg_max > 0 ? t->idxcg_max : 1, sizeof(double));
  cnt = 0;
  for (int j1 = 0; j1 <= J; j1++)
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = j1 - j2; j <= imin(J, j1 + j2); j += 2)
        for (int m1 = 0; m1 <= j1; m1++) {
          int aa2 = 2 * m1 - j1;
          for (int m2 = 0; m2 <= j2; m2++) {
            int bb2 = 2 * m2 - j2;
            int m = (aa2 + bb2 + j) / 2;
            if (m < 0 || m > j) { t->cglist[cnt++] = 0.0; continue; }
            double sum = 0.0;
            int min_z = imax(0, imax(-(j - j2 + aa2) / 2, -(j - j1 - bb2) / 2));
            int max_z = imin((j1 + j2 - j) / 2, imin((j1 - aa2) / 2, (j2 + bb2) / 2));
            for (int z = min_z; z <= max_z; z++) {
              double ifac = (z % 2 == 0) ? 1.0 : -1.0;
              int64_t den = factorial_i(z) * factorial_i((j1 + j2 - j) / 2 - z) * factorial_i((j1 - aa2) / 2 - z) *
                            factorial_i((j2 + bb2) / 2 - z) * factorial_i((j - j2 + aa2) / 2 + z) *
                            factorial_i((j - j1 - bb2) / 2 + z);
              sum += ifac / (double)den;
            }
            int cc2 = 2 * m - j;
            double dcg = deltacg(j1, j2, j);
            int64_t fp = factorial_i((j1 + aa2) / 2) * factorial_i((j1 - aa2) / 2) * factorial_i((j2 + bb2) / 2) *
                         factorial_i((j2 - bb2) / 2) * factorial_i((j + cc2) / 2) * factorial_i((j - cc2) / 2) *
                         (int64_t)(j + 1);
            double sfaccg = sqrt((double)fp);
            t->cglist[cnt++] = sum * dcg * sfaccg;
          }
        }
  (void)factorial_d;
  return t;
}

void orc_snap_tables_free(orc_snap_tables* t) {
  if (!t) return;
  free(t->idxcg_block); free(t->idxu_block); free(t->idxb_block); free(t->idxz_block);
  free(t->idxb); free(t->idxz); free(t->cglist); free(t->rootpq); free(t->bzero);
  free(t);
}

/* ================================================================== */
/* SNAP neighbor list (utilities/utl.jl:38-101)                        */
/* ================================================================== */

int64_t orc_snap_neighborlist(const double* pos, int64_t n, const double L[3], const int32_t* species,
                              const int32_t* appear, const orc_snap_params* p, int mode,
                              int64_t* offsets, int32_t* nbr, double* R, double* rvec, int64_t cap) {
  int ns = p->nspecies;
  if (n < 0 || ns < 1) return -1;
  double* cutsq = (double*)calloc(ns * ns, sizeof(double));
  double cutmax = 0.0;
  for (int i = 0; i < ns; i++) { /* :42-56 */
    double cut = 2.0 * p->radii[i] * p->rcutfac;
    if (cut > cutmax) cutmax = cut;
    cutsq[i * ns + i] = cut * cut;
    for (int j = i + 1; j < ns; j++) {
      cut = (p->radii[i] + p->radii[j]) * p->rcutfac;
      cutsq[i * ns + j] = cut * cut;
      cutsq[j * ns + i] = cut * cut;
    }
  }
  /* :81 inrange radius: sqrt(cutmax) as-is (REFERENCE), cutmax (CUTOFF) */
  double q = (mode == ORC_SNAP_LIST_REFERENCE) ? sqrt(cutmax) : cutmax;
  int64_t np = 0;
  int64_t rc = 0;
  for (int64_t a = 0; a < n; a++) {
    offsets[a] = np;
    /* :70 n_element from unique(atomic_symbol(A)) order (REFERENCE) */
    int ne = (mode == ORC_SNAP_LIST_REFERENCE) ? appear[a] : species[a];
    int64_t m = 0; /* enumerate index over the sorted inrange result */
    for (int64_t b = 0; b < n; b++) {
      if (!(sqrt(orc_pe_sumsq(pos + 3 * a, pos + 3 * b, L)) <= q)) continue;
      m++;
      /* :86 m_element = element of atom m (enumerate index), as-is in
       * REFERENCE mode; CUTOFF mode uses the neighbor's own species. */
      int me;
      if (mode == ORC_SNAP_LIST_REFERENCE) {
        if (m - 1 >= n) { rc = -2; continue; }
        me = appear[m - 1];
      } else {
        me = species[b];
      }
      if (ne >= ns || me >= ns) { rc = -3; continue; }
      double d[3];
      orc_get_displacement(L, pos + 3 * a, pos + 3 * b, d);
      double rr[3] = {-1.0 * d[0], -1.0 * d[1], -1.0 * d[2]}; /* :87 */
      double rsq = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
      if (b != a && rsq <= cutsq[ne * ns + me]) { /* :89 */
        if (np < cap) {
          if (nbr) nbr[np] = (int32_t)b;
          if (R) R[np] = sqrt(rsq);
          if (rvec) { rvec[3 * np] = rr[0]; rvec[3 * np + 1] = rr[1]; rvec[3 * np + 2] = rr[2]; }
        }
        np++;
      }
    }
  }
  offsets[n] = np;
  free(cutsq);
  return rc < 0 ? rc : np;
}

/* ================================================================== */
/* SNAP per-atom kernels                                              */
/* ================================================================== */

#define RPQ(t, p_, q_) ((t)->rootpq[(p_) * ((t)->twojmax + 2) + (q_)])

/* compute_ui.jl:50-129 compute_uarray; ur/ui are one ulist row. */
static void uarray(const orc_snap_tables* t, double x, double y, double z, double z0, double r, double* ur,
                   double* ui) {
  double r0inv = 1.0 / sqrt(r * r + z0 * z0);
  double a_r = r0inv * z0;
  double a_i = -r0inv * z;
  double b_r = r0inv * y;
  double b_i = -r0inv * x;
  ur[0] = 1.0;
  ui[0] = 0.0;
  for (int j = 1; j <= t->twojmax; j++) {
    int jju = t->idxu_block[j];
    int jjup = t->idxu_block[j - 1];
    for (int mb = 0; 2 * mb <= j; mb++) {
      ur[jju] = 0.0;
      ui[jju] = 0.0;
      for (int ma = 0; ma < j; ma++) {
        double rootpq = RPQ(t, j - ma, j - mb);
        ur[jju] += rootpq * (a_r * ur[jjup] + a_i * ui[jjup]);
        ui[jju] += rootpq * (a_r * ui[jjup] - a_i * ur[jjup]);
        rootpq = RPQ(t, ma + 1, j - mb);
        ur[jju + 1] = -rootpq * (b_r * ur[jjup] + b_i * ui[jjup]);
        ui[jju + 1] = -rootpq * (b_r * ui[jjup] - b_i * ur[jjup]);
        jju++;
        jjup++;
      }
      jju++;
    }
    /* inversion symmetry, :103-125 */
    jju = t->idxu_block[j];
    jjup = jju + (j + 1) * (j + 1) - 1;
    int mbpar = 1;
    for (int mb = 0; 2 * mb <= j; mb++) {
      int mapar = mbpar;
      for (int ma = 0; ma <= j; ma++) {
        if (mapar == 1) { ur[jjup] = ur[jju]; ui[jjup] = -ui[jju]; }
        else { ur[jjup] = -ur[jju]; ui[jjup] = ui[jju]; }
        mapar = -mapar;
        jju++;
        jjup--;
      }
      mbpar = -mbpar;
    }
  }
}

/* compute_ui.jl:350-364 */
static double sfac_of(const orc_snap_params* p, double r, double rcut) {
  if (!p->switch_flag) return 1.0;
  if (r <= p->rmin0) return 1.0;
  if (r > rcut) return 0.0;
  double rcutfac = M_PI / (rcut - p->rmin0);
  return 0.5 * (cos((r - p->rmin0) * rcutfac) + 1.0);
}
/* compute_ui.jl:366-380 (the `rcut > rcut` branch of :372 never fires) */
static double dsfac_of(const orc_snap_params* p, double r, double rcut) {
  if (!p->switch_flag) return 0.0;
  if (r <= p->rmin0) return 0.0;
  double rcutfac = M_PI / (rcut - p->rmin0);
  return -0.5 * (sin((r - p->rmin0) * rcutfac) * rcutfac);
}

typedef struct {
  int nnb, nel;
  const orc_snap_tables* t;
  double *ulisttot_r, *ulisttot_i;  /* nel x idxu_max */
  double *ulist_r, *ulist_i;        /* (nnb+1) x idxu_max */
  double *dulist_r, *dulist_i;      /* idxu_max x 3 */
  double *zlist_r, *zlist_i;        /* nel^2 x idxz_max */
  double* blist;                    /* nel^3 x idxb_max */
  double* dblist;                   /* nel^3 x idxb_max x 3 */
} orc_rt;

static void rt_alloc(orc_rt* rt, const orc_snap_tables* t, int nnb) {
  int nel = t->num_elements;
  rt->nnb = nnb; rt->nel = nel; rt->t = t;
  rt->ulisttot_r = (double*)calloc(nel * t->idxu_max, sizeof(double));
  rt->ulisttot_i = (double*)calloc(nel * t->idxu_max, sizeof(double));
  rt->ulist_r = (double*)calloc((nnb + 1) * t->idxu_max, sizeof(double));
  rt->ulist_i = (double*)calloc((nnb + 1) * t->idxu_max, sizeof(double));
  rt->dulist_r = (double*)calloc(t->idxu_max * 3, sizeof(double));
  rt->dulist_i = (double*)calloc(t->idxu_max * 3, sizeof(double));
  rt->zlist_r = (double*)calloc(nel * nel * t->idxz_max, sizeof(double));
  rt->zlist_i = (double*)calloc(nel * nel * t->idxz_max, sizeof(double));
  rt->blist = (double*)calloc(nel * nel * nel * t->idxb_max, sizeof(double));
  rt->dblist = (double*)calloc(nel * nel * nel * t->idxb_max * 3, sizeof(double));
}
static void rt_free(orc_rt* rt) {
  free(rt->ulisttot_r); free(rt->ulisttot_i); free(rt->ulist_r); free(rt->ulist_i);
  free(rt->dulist_r); free(rt->dulist_i); free(rt->zlist_r); free(rt->zlist_i);
  free(rt->blist); free(rt->dblist);
}

/* compute_ui.jl:5-44 compute_ui (+ zero_uarraytot :131-152, add_uarraytot
 * :155-175).  ielem is 1-based as in the reference. elem: per-neighbor
 * element (1-based). */
static void compute_ui(const orc_snap_params* p, orc_rt* rt, int ielem, const double* rij, const double* wj,
                       const double* rcutij, const int* elem) {
  const orc_snap_tables* t = rt->t;
  if (!p->chem_flag) ielem = 1;
  for (int jelem = 0; jelem < t->num_elements; jelem++)
    for (int j = 0; j <= t->twojmax; j++) {
      int jju = t->idxu_block[j];
      for (int mb = 0; mb <= j; mb++)
        for (int ma = 0; ma <= j; ma++) {
          rt->ulisttot_r[jelem * t->idxu_max + jju] = 0.0;
          rt->ulisttot_i[jelem * t->idxu_max + jju] = 0.0;
          if ((jelem + 1) == ielem || p->wselfall_flag)
            if (ma == mb) rt->ulisttot_r[jelem * t->idxu_max + jju] = t->wself;
          jju++;
        }
    }
  for (int ind = 0; ind < rt->nnb; ind++) {
    double x = rij[3 * ind], y = rij[3 * ind + 1], z = rij[3 * ind + 2];
    double rsq = x * x + y * y + z * z;
    double r = sqrt(rsq);
    double theta0 = (r - p->rmin0) * p->rfac0 * M_PI / (rcutij[ind] - p->rmin0);
    double z0 = r / tan(theta0);
    double* ur = rt->ulist_r + (ind + 1) * t->idxu_max;
    double* ui = rt->ulist_i + (ind + 1) * t->idxu_max;
    uarray(t, x, y, z, z0, r, ur, ui);
    int je = p->chem_flag ? elem[ind] : 1;
    double sfac = sfac_of(p, r, rcutij[ind]);
    sfac *= wj[ind];
    for (int jju = 0; jju < t->idxu_max; jju++) {
      rt->ulisttot_r[(je - 1) * t->idxu_max + jju] += sfac * ur[jju];
      rt->ulisttot_i[(je - 1) * t->idxu_max + jju] += sfac * ui[jju];
    }
  }
}

/* compute_z.jl:1-62 compute_zi, including the as-is accumulation into
 * zlist[jjz] without the idouble offset (:46-47). */
static void compute_zi(const orc_snap_params* p, orc_rt* rt) {
  const orc_snap_tables* t = rt->t;
  int idouble = 0;
  int nel = t->num_elements;
  for (int e1 = 0; e1 < nel; e1++)
    for (int e2 = 0; e2 < nel; e2++) {
      for (int jjz = 0; jjz < t->idxz_max; jjz++) {
        const orc_zidx* zi = &t->idxz[jjz];
        int j1 = zi->j1, j2 = zi->j2, j = zi->j;
        const double* cgblock = t->cglist + t->idxcg_block[B3(t, j1, j2, j)];
        rt->zlist_r[idouble * t->idxz_max + jjz] = 0.0;
        rt->zlist_i[idouble * t->idxz_max + jjz] = 0.0;
        int jju1 = t->idxu_block[j1] + (j1 + 1) * zi->mb1min;
        int jju2 = t->idxu_block[j2] + (j2 + 1) * zi->mb2max;
        int icgb = zi->mb1min * (j2 + 1) + zi->mb2max;
        for (int ib = 0; ib < zi->nb; ib++) {
          double suma1_r = 0.0, suma1_i = 0.0;
          const double* u1_r = rt->ulisttot_r + e1 * t->idxu_max + jju1;
          const double* u1_i = rt->ulisttot_i + e1 * t->idxu_max + jju1;
          const double* u2_r = rt->ulisttot_r + e2 * t->idxu_max + jju2;
          const double* u2_i = rt->ulisttot_i + e2 * t->idxu_max + jju2;
          int ma1 = zi->ma1min, ma2 = zi->ma2max;
          int icga = zi->ma1min * (j2 + 1) + zi->ma2max;
          for (int ia = 0; ia < zi->na; ia++) {
            suma1_r += cgblock[icga] * (u1_r[ma1] * u2_r[ma2] - u1_i[ma1] * u2_i[ma2]);
            suma1_i += cgblock[icga] * (u1_r[ma1] * u2_i[ma2] + u1_i[ma1] * u2_r[ma2]);
            ma1++;
            ma2--;
            icga += j2;
          }
          rt->zlist_r[jjz] += cgblock[icgb] * suma1_r;
          rt->zlist_i[jjz] += cgblock[icgb] * suma1_i;
          jju1 += j1 + 1;
          jju2 -= j2 + 1;
          icgb += j2;
        }
        if (p->bnorm_flag) {
          rt->zlist_r[jjz] /= (j + 1);
          rt->zlist_i[jjz] /= (j + 1);
        }
      }
      idouble++;
    }
}

/* compute_b.jl:5-95 compute_bi (half-column sums :36-62, bzero shift :70-93) */
static void compute_bi(const orc_snap_params* p, orc_rt* rt) {
  const orc_snap_tables* t = rt->t;
  int nel = t->num_elements;
  for (int ielem = 1; ielem <= nel; ielem++) {
    int itriple = 0, idouble = 0;
    for (int e1 = 0; e1 < nel; e1++)
      for (int e2 = 0; e2 < nel; e2++) {
        const double* zr = rt->zlist_r + idouble * t->idxz_max;
        const double* zi = rt->zlist_i + idouble * t->idxz_max;
        for (int e3 = 0; e3 < nel; e3++) {
          const double* ur = rt->ulisttot_r + e3 * t->idxu_max;
          const double* ui = rt->ulisttot_i + e3 * t->idxu_max;
          for (int jjb = 0; jjb < t->idxb_max; jjb++) {
            int j1 = t->idxb[3 * jjb], j2 = t->idxb[3 * jjb + 1], j = t->idxb[3 * jjb + 2];
            int jjz = t->idxz_block[B3(t, j1, j2, j)];
            int jju = t->idxu_block[j];
            double sumzu = 0.0;
            for (int mb = 0; 2 * mb < j; mb++)
              for (int ma = 0; ma <= j; ma++) {
                sumzu += (ur[jju] * zr[jjz] + ui[jju] * zi[jjz]);
                jjz++;
                jju++;
              }
            if (j % 2 == 0) {
              int mb = j / 2;
              for (int ma = 0; ma < mb; ma++) {
                sumzu += ur[jju] * zr[jjz] + ui[jju] * zi[jjz];
                jjz++;
                jju++;
              }
              sumzu += 0.5 * (ur[jju] * zr[jjz] + ui[jju] * zi[jjz]);
            }
            rt->blist[itriple * t->idxb_max + jjb] = 2.0 * sumzu;
          }
          itriple++;
        }
        idouble++;
      }
    if (p->bzero_flag) {
      if (!p->wselfall_flag) {
        int it = (ielem - 1) * (nel + 1) * nel + (ielem - 1);
        for (int jjb = 0; jjb < t->idxb_max; jjb++) rt->blist[it * t->idxb_max + jjb] -= t->bzero[t->idxb[3 * jjb + 2]];
      } else {
        int it = 0;
        for (int a = 0; a < nel * nel * nel; a++, it++)
          for (int jjb = 0; jjb < t->idxb_max; jjb++) rt->blist[it * t->idxb_max + jjb] -= t->bzero[t->idxb[3 * jjb + 2]];
      }
    }
  }
}

/* compute_ui.jl:186-326 compute_duarray; ur/ui are the neighbor's ulist row.
 * dulist layout: [jju*3 + k]. */
static void duarray(const orc_snap_params* p, const orc_snap_tables* t, double x, double y, double z, double z0,
                    double r, double dz0dr, double wj, double rcut, const double* ur, const double* ui, double* dr,
                    double* di) {
  double rinv = 1.0 / r;
  double ux = x * rinv, uy = y * rinv, uz = z * rinv;
  double r0inv = 1.0 / sqrt(r * r + z0 * z0);
  double a_r = z0 * r0inv, a_i = -z * r0inv, b_r = y * r0inv, b_i = -x * r0inv;
  double dr0invdr = -r0inv * r0inv * r0inv * (r + z0 * dz0dr);
  double dr0inv[3] = {dr0invdr * ux, dr0invdr * uy, dr0invdr * uz};
  double dz0[3] = {dz0dr * ux, dz0dr * uy, dz0dr * uz};
  double da_r[3], da_i[3], db_r[3], db_i[3];
  for (int k = 0; k < 3; k++) {
    da_r[k] = dz0[k] * r0inv + z0 * dr0inv[k];
    da_i[k] = -z * dr0inv[k];
    db_r[k] = y * dr0inv[k];
    db_i[k] = -x * dr0inv[k];
  }
  da_i[2] += -r0inv;
  db_i[0] += -r0inv;
  db_r[1] += r0inv;
  for (int k = 0; k < 3; k++) { dr[k] = 0.0; di[k] = 0.0; }
  for (int j = 1; j <= t->twojmax; j++) {
    int jju = t->idxu_block[j], jjup = t->idxu_block[j - 1];
    for (int mb = 0; 2 * mb <= j; mb++) {
      for (int k = 0; k < 3; k++) { dr[3 * jju + k] = 0.0; di[3 * jju + k] = 0.0; }
      for (int ma = 0; ma < j; ma++) {
        double rootpq = RPQ(t, j - ma, j - mb);
        for (int k = 0; k < 3; k++) {
          dr[3 * jju + k] += rootpq * (da_r[k] * ur[jjup] + da_i[k] * ui[jjup] + a_r * dr[3 * jjup + k] +
                                       a_i * di[3 * jjup + k]);
          di[3 * jju + k] += rootpq * (da_r[k] * ui[jjup] - da_i[k] * ur[jjup] + a_r * di[3 * jjup + k] -
                                       a_i * dr[3 * jjup + k]);
        }
        rootpq = RPQ(t, ma + 1, j - mb);
        for (int k = 0; k < 3; k++) {
          dr[3 * (jju + 1) + k] = -rootpq * (db_r[k] * ur[jjup] + db_i[k] * ui[jjup] + b_r * dr[3 * jjup + k] +
                                             b_i * di[3 * jjup + k]);
          di[3 * (jju + 1) + k] = -rootpq * (db_r[k] * ui[jjup] - db_i[k] * ur[jjup] + b_r * di[3 * jjup + k] -
                                             b_i * dr[3 * jjup + k]);
        }
        jju++;
        jjup++;
      }
      jju++;
    }
    jju = t->idxu_block[j];
    jjup = jju + (j + 1) * (j + 1) - 1;
    int mbpar = 1;
    for (int mb = 0; 2 * mb <= j; mb++) {
      int mapar = mbpar;
      for (int ma = 0; ma <= j; ma++) {
        for (int k = 0; k < 3; k++) {
          if (mapar == 1) { dr[3 * jjup + k] = dr[3 * jju + k]; di[3 * jjup + k] = -di[3 * jju + k]; }
          else { dr[3 * jjup + k] = -dr[3 * jju + k]; di[3 * jjup + k] = di[3 * jju + k]; }
        }
        mapar = -mapar;
        jju++;
        jjup--;
      }
      mbpar = -mbpar;
    }
  }
  double sfac = sfac_of(p, r, rcut), dsfac = dsfac_of(p, r, rcut);
  sfac *= wj;
  dsfac *= wj;
  for (int j = 0; j <= t->twojmax; j++) {
    int jju = t->idxu_block[j];
    for (int mb = 0; 2 * mb <= j; mb++)
      for (int ma = 0; ma <= j; ma++) {
        dr[3 * jju + 0] = dsfac * ur[jju] * ux + sfac * dr[3 * jju + 0];
        di[3 * jju + 0] = dsfac * ui[jju] * ux + sfac * di[3 * jju + 0];
        dr[3 * jju + 1] = dsfac * ur[jju] * uy + sfac * dr[3 * jju + 1];
        di[3 * jju + 1] = dsfac * ui[jju] * uy + sfac * di[3 * jju + 1];
        dr[3 * jju + 2] = dsfac * ur[jju] * uz + sfac * dr[3 * jju + 2];
        di[3 * jju + 2] = dsfac * ui[jju] * uz + sfac * di[3 * jju + 2];
        jju++;
      }
  }
}

/* compute_ui.jl:328-347 compute_duidrj */
static void compute_duidrj(const orc_snap_params* p, const orc_snap_tables* t, const double* rij, double wj,
                           double rcut, const double* ur, const double* ui, double* dr, double* di) {
  double x = rij[0], y = rij[1], z = rij[2];
  double rsq = x * x + y * y + z * z;
  double r = sqrt(rsq);
  double rscale0 = p->rfac0 * M_PI / (rcut - p->rmin0);
  double theta0 = (r - p->rmin0) * rscale0;
  double cs = cos(theta0), sn = sin(theta0);
  double z0 = r * cs / sn;
  double dz0dr = z0 / r - (r * rscale0) * (rsq + z0 * z0) / rsq;
  duarray(p, t, x, y, z, z0, r, dz0dr, wj, rcut, ur, ui, dr, di);
}

/* half-column sum of Re(conj(du_k) * z) over level jl, as in
 * compute_b.jl:150-208 (and the j1, j2 variants :210-272, :274-337). */
static void half_sum_du_z(const orc_snap_tables* t, int jl, int jju, int jjz, const double* dr, const double* di,
                          const double* zr, const double* zi, double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
  for (int mb = 0; 2 * mb < jl; mb++)
    for (int ma = 0; ma <= jl; ma++) {
      for (int k = 0; k < 3; k++) out[k] += dr[3 * jju + k] * zr[jjz] + di[3 * jju + k] * zi[jjz];
      jjz++;
      jju++;
    }
  if (jl % 2 == 0) {
    int mb = jl / 2;
    for (int ma = 0; ma < mb; ma++) {
      for (int k = 0; k < 3; k++) out[k] += dr[3 * jju + k] * zr[jjz] + di[3 * jju + k] * zi[jjz];
      jjz++;
      jju++;
    }
    for (int k = 0; k < 3; k++) out[k] += (dr[3 * jju + k] * zr[jjz] + di[3 * jju + k] * zi[jjz]) * 0.5;
  }
  (void)t;
}

/* compute_b.jl:104-342 compute_dbidrj; dblist layout [(itriple*idxb_max+jjb)*3 + k] */
static void compute_dbidrj(const orc_snap_params* p, orc_rt* rt, int jelem) {
  const orc_snap_tables* t = rt->t;
  int nel = t->num_elements;
  memset(rt->dblist, 0, sizeof(double) * nel * nel * nel * t->idxb_max * 3);
  int elem3 = p->chem_flag ? jelem - 1 : 0;
  for (int jjb = 0; jjb < t->idxb_max; jjb++) {
    int j1 = t->idxb[3 * jjb], j2 = t->idxb[3 * jjb + 1], j = t->idxb[3 * jjb + 2];
    for (int e1 = 0; e1 < nel; e1++)
      for (int e2 = 0; e2 < nel; e2++) {
        double s[3];
        int idouble = e1 * nel + e2;
        int itriple = (e1 * nel + e2) * nel + elem3;
        half_sum_du_z(t, j, t->idxu_block[j], t->idxz_block[B3(t, j1, j2, j)], rt->dulist_r, rt->dulist_i,
                      rt->zlist_r + idouble * t->idxz_max, rt->zlist_i + idouble * t->idxz_max, s);
        for (int k = 0; k < 3; k++) rt->dblist[(itriple * t->idxb_max + jjb) * 3 + k] = 2.0 * s[k];

        double j1fac = (j + 1) / (j1 + 1.0);
        idouble = e1 * nel + e2;
        itriple = (elem3 * nel + e2) * nel + e1;
        half_sum_du_z(t, j1, t->idxu_block[j1], t->idxz_block[B3(t, j, j2, j1)], rt->dulist_r, rt->dulist_i,
                      rt->zlist_r + idouble * t->idxz_max, rt->zlist_i + idouble * t->idxz_max, s);
        for (int k = 0; k < 3; k++)
          rt->dblist[(itriple * t->idxb_max + jjb) * 3 + k] += p->bnorm_flag ? 2.0 * s[k] : 2.0 * s[k] * j1fac;

        double j2fac = (j + 1) / (j2 + 1.0);
        idouble = e2 * nel + e1;
        itriple = (e1 * nel + elem3) * nel + e2;
        half_sum_du_z(t, j2, t->idxu_block[j2], t->idxz_block[B3(t, j, j1, j2)], rt->dulist_r, rt->dulist_i,
                      rt->zlist_r + idouble * t->idxz_max, rt->zlist_i + idouble * t->idxz_max, s);
        for (int k = 0; k < 3; k++)
          rt->dblist[(itriple * t->idxb_max + jjb) * 3 + k] += p->bnorm_flag ? 2.0 * s[k] : 2.0 * s[k] * j2fac;
      }
  }
}

/* Per-atom neighbor data of runtime_arrays.jl:32-81 initialize_runtime_arrays. */
static int atom_pairs(const orc_snap_params* p, int64_t i, const int32_t* species, const int64_t* offsets,
                      const int32_t* nbr, const double* rvec, double** rij, double** wj, double** rcutij,
                      int** elem) {
  int nnb = (int)(offsets[i + 1] - offsets[i]);
  int si = species[i];
  *rij = (double*)malloc(sizeof(double) * 3 * (nnb > 0 ? nnb : 1));
  *wj = (double*)malloc(sizeof(double) * (nnb > 0 ? nnb : 1));
  *rcutij = (double*)malloc(sizeof(double) * (nnb > 0 ? nnb : 1));
  *elem = (int*)malloc(sizeof(int) * (nnb > 0 ? nnb : 1));
  for (int q = 0; q < nnb; q++) {
    int64_t pq = offsets[i] + q;
    int sj = species[nbr[pq]];
    (*elem)[q] = sj + 1;
    (*rcutij)[q] = (p->radii[si] + p->radii[sj]) * p->rcutfac;
    (*wj)[q] = p->weight[sj];
    (*rij)[3 * q] = rvec[3 * pq];
    (*rij)[3 * q + 1] = rvec[3 * pq + 1];
    (*rij)[3 * q + 2] = rvec[3 * pq + 2];
  }
  return nnb;
}

int orc_snap_atom_uzb(const orc_snap_params* p, const orc_snap_tables* t, int ielem, int nnb, const double* rij,
                      const double* wj, const double* rcutij, double* ulisttot, double* zlist, double* blist) {
  orc_rt rt;
  rt_alloc(&rt, t, nnb);
  int* elem = (int*)malloc(sizeof(int) * (nnb > 0 ? nnb : 1));
  for (int q = 0; q < nnb; q++) elem[q] = 1;
  compute_ui(p, &rt, ielem, rij, wj, rcutij, elem);
  compute_zi(p, &rt);
  compute_bi(p, &rt);
  for (int k = 0; k < t->idxu_max; k++) { ulisttot[2 * k] = rt.ulisttot_r[k]; ulisttot[2 * k + 1] = rt.ulisttot_i[k]; }
  if (zlist)
    for (int k = 0; k < t->idxz_max; k++) { zlist[2 * k] = rt.zlist_r[k]; zlist[2 * k + 1] = rt.zlist_i[k]; }
  if (blist) memcpy(blist, rt.blist, sizeof(double) * t->idxb_max);
  free(elem);
  rt_free(&rt);
  return 0;
}

int orc_snap_uarray(const orc_snap_tables* t, double rfac0, double rmin0, double rcut, const double r3[3],
                    double* ulist) {
  double x = r3[0], y = r3[1], z = r3[2];
  double rsq = x * x + y * y + z * z;
  double r = sqrt(rsq);
  double theta0 = (r - rmin0) * rfac0 * M_PI / (rcut - rmin0);
  double z0 = r / tan(theta0);
  double* ur = (double*)malloc(sizeof(double) * t->idxu_max);
  double* ui = (double*)malloc(sizeof(double) * t->idxu_max);
  uarray(t, x, y, z, z0, r, ur, ui);
  for (int k = 0; k < t->idxu_max; k++) { ulist[2 * k] = ur[k]; ulist[2 * k + 1] = ui[k]; }
  free(ur); free(ui);
  return 0;
}

int orc_snap_duarray(const orc_snap_params* p, const orc_snap_tables* t, double rcut, double wj, const double r3[3],
                     double* dulist) {
  double x = r3[0], y = r3[1], z = r3[2];
  double rsq = x * x + y * y + z * z;
  double r = sqrt(rsq);
  double theta0 = (r - p->rmin0) * p->rfac0 * M_PI / (rcut - p->rmin0);
  double z0 = r / tan(theta0);
  double* ur = (double*)malloc(sizeof(double) * t->idxu_max);
  double* ui = (double*)malloc(sizeof(double) * t->idxu_max);
  double* dr = (double*)calloc(3 * t->idxu_max, sizeof(double));
  double* di = (double*)calloc(3 * t->idxu_max, sizeof(double));
  uarray(t, x, y, z, z0, r, ur, ui);
  compute_duidrj(p, t, r3, wj, rcut, ur, ui, dr, di);
  for (int k = 0; k < 3 * t->idxu_max; k++) { dulist[2 * k] = dr[k]; dulist[2 * k + 1] = di[k]; }
  free(ur); free(ui); free(dr); free(di);
  return 0;
}

int orc_snap_local_descriptors(const orc_snap_params* p, const orc_snap_tables* t, int64_t n, const int32_t* species,
                               const int64_t* offsets, const int32_t* nbr, const double* rvec, double* B) {
  int ncoeff = orc_snap_num_coeffs(p->twojmax, p->nspecies, p->chem_flag);
  for (int64_t i = 0; i < n; i++) { /* snap.jl:112-122 */
    double *rij, *wj, *rcutij;
    int* elem;
    int nnb = atom_pairs(p, i, species, offsets, nbr, rvec, &rij, &wj, &rcutij, &elem);
    orc_rt rt;
    rt_alloc(&rt, t, nnb);
    compute_ui(p, &rt, species[i] + 1, rij, wj, rcutij, elem);
    compute_zi(p, &rt);
    compute_bi(p, &rt);
    memcpy(B + i * ncoeff, rt.blist, sizeof(double) * ncoeff);
    rt_free(&rt);
    free(rij); free(wj); free(rcutij); free(elem);
  }
  return 0;
}

int orc_snap_all_descriptors(const orc_snap_params* p, const orc_snap_tables* t, int64_t n, const double* pos,
                             const int32_t* species, const int64_t* offsets, const int32_t* nbr, const double* rvec,
                             double* B, double* dB, double* W) {
  int ncoeff = orc_snap_num_coeffs(p->twojmax, p->nspecies, p->chem_flag);
  int M = ncoeff * p->nspecies;
  if (dB) memset(dB, 0, sizeof(double) * n * M * 3);
  if (W) memset(W, 0, sizeof(double) * M * 6);
  for (int64_t i = 0; i < n; i++) { /* snap.jl:237-283 */
    double *rij, *wj, *rcutij;
    int* elem;
    int nnb = atom_pairs(p, i, species, offsets, nbr, rvec, &rij, &wj, &rcutij, &elem);
    orc_rt rt;
    rt_alloc(&rt, t, nnb);
    int ie = species[i] + 1;
    compute_ui(p, &rt, ie, rij, wj, rcutij, elem);
    compute_zi(p, &rt);
    compute_bi(p, &rt);
    if (B) memcpy(B + i * ncoeff, rt.blist, sizeof(double) * ncoeff);
    int ioff = ncoeff * (ie - 1);
    for (int ind = 0; ind < nnb; ind++) {
      int64_t jj = nbr[offsets[i] + ind];
      /* :261 uses U row ind+1 (the neighbor's own row) */
      compute_duidrj(p, t, rij + 3 * ind, wj[ind], rcutij[ind], rt.ulist_r + (ind + 1) * t->idxu_max,
                     rt.ulist_i + (ind + 1) * t->idxu_max, rt.dulist_r, rt.dulist_i);
      compute_dbidrj(p, &rt, elem[ind]);
      const double* xi = pos + 3 * i;
      const double* xj = pos + 3 * jj;
      for (int c = 0; c < ncoeff; c++) {
        const double* db = rt.dblist + 3 * c;
        if (dB)
          for (int k = 0; k < 3; k++) {
            dB[(i * M + ioff + c) * 3 + k] += db[k];   /* :264 */
            dB[(jj * M + ioff + c) * 3 + k] -= db[k];  /* :265 */
          }
        if (W) { /* :270-281 rows [x dBx, y dBy, z dBz, z dBy, z dBx, y dBx] */
          double* w = W + (ioff + c) * 6;
          w[0] += db[0] * xi[0]; w[1] += db[1] * xi[1]; w[2] += db[2] * xi[2];
          w[3] += db[1] * xi[2]; w[4] += db[0] * xi[2]; w[5] += db[0] * xi[1];
          w[0] -= db[0] * xj[0]; w[1] -= db[1] * xj[1]; w[2] -= db[2] * xj[2];
          w[3] -= db[1] * xj[2]; w[4] -= db[0] * xj[2]; w[5] -= db[0] * xj[1];
        }
      }
    }
    rt_free(&rt);
    free(rij); free(wj); free(rcutij); free(elem);
  }
  return 0;
}

int orc_snap_plain_energy_forces(const orc_snap_params* p, const orc_snap_tables* t, int64_t n, int64_t nloop,
                                 const int32_t* species, const int64_t* offsets, const int32_t* nbr,
                                 const double* rvec, const double* beta, int nthreads, double* energy,
                                 double* forces) {
  int ncoeff = orc_snap_num_coeffs(p->twojmax, p->nspecies, p->chem_flag);
  double E = 0.0;
  memset(forces, 0, sizeof(double) * 3 * n);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : E) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < nloop; i++) {
    double *rij, *wj, *rcutij;
    int* elem;
    int nnb = atom_pairs(p, i, species, offsets, nbr, rvec, &rij, &wj, &rcutij, &elem);
    orc_rt rt;
    rt_alloc(&rt, t, nnb);
    int ie = species[i] + 1;
    const double* bet = beta + (int64_t)ncoeff * (ie - 1);
    compute_ui(p, &rt, ie, rij, wj, rcutij, elem);
    compute_zi(p, &rt);
    compute_bi(p, &rt);
    for (int c = 0; c < ncoeff; c++) E += bet[c] * rt.blist[c];
    for (int ind = 0; ind < nnb; ind++) {
      int64_t jj = nbr[offsets[i] + ind];
      compute_duidrj(p, t, rij + 3 * ind, wj[ind], rcutij[ind], rt.ulist_r + (ind + 1) * t->idxu_max,
                     rt.ulist_i + (ind + 1) * t->idxu_max, rt.dulist_r, rt.dulist_i);
      compute_dbidrj(p, &rt, elem[ind]);
      double f[3] = {0.0, 0.0, 0.0};
      for (int c = 0; c < ncoeff; c++)
        for (int k = 0; k < 3; k++) f[k] += bet[c] * rt.dblist[3 * c + k];
      for (int k = 0; k < 3; k++) {
#pragma omp atomic
        forces[3 * i + k] += f[k];
#pragma omp atomic
        forces[3 * jj + k] -= f[k];
      }
    }
    rt_free(&rt);
    free(rij); free(wj); free(rcutij); free(elem);
  }
  *energy = E;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Adjoint (derived from the B formula; compute_yi.jl:6-122 and         */
/* compute_dij.jl:5-74 are the specification only).                    */
/*                                                                      */
/* E_i = sum_b beta_b B_b with B_(j1,j2,j) = sum_half w Re(conj U_j Z_(j1,j2,j)). */
/* compute_dbidrj shows dB_b/dr = half-sums of Re(conj dU_l . Z) over the */
/* three levels l = j, j1, j2 with blocks (j1,j2,j), (j,j2,j1), (j,j1,j2) */
/* and factors 1, (j+1)/(j1+1), (j+1)/(j2+1) (1 if bnorm).  Collecting    */
/* by level gives dE/dr = 2 sum_half w Re(conj dU . Y) with              */
/*   Y_l = sum over Z blocks (a,c,l) of coef(a,c,l) Z_(a,c,l),           */
/* coef from the (triple, role) pairs mapping to that block.            */
/* ------------------------------------------------------------------ */

/* Coefficient multiplying Z block (j1,j2,j) (j1 >= j2) in Y_j. */
static double yi_coef(const orc_snap_params* p, const orc_snap_tables* t, const double* beta, int j1, int j2, int j) {
  double c = 0.0;
  /* role j of b = (j1,j2,j): needs j >= j1 */
  if (j >= j1) c += beta[t->idxb_block[B3(t, j1, j2, j)]];
  /* role j1 of b = (j, j2, j1) [triple (j1'=j, j2'=j2, j'=j1)]: needs j1 >= j >= j2 */
  if (j1 >= j && j >= j2) {
    double f = p->bnorm_flag ? 1.0 : (j1 + 1) / (j + 1.0);
    c += beta[t->idxb_block[B3(t, j, j2, j1)]] * f;
  }
  /* role j2 of b = (j2, j, j1) [triple (j1'=j2, j2'=j, j'=j1)]: needs j1 >= j2 >= j */
  if (j1 >= j2 && j2 >= j) {
    double f = p->bnorm_flag ? 1.0 : (j1 + 1) / (j + 1.0);
    c += beta[t->idxb_block[B3(t, j2, j, j1)]] * f;
  }
  return c;
}

int orc_snap_adjoint_energy_forces(const orc_snap_params* p, const orc_snap_tables* t, int64_t n,
                                   const int32_t* species, const int64_t* offsets, const int32_t* nbr,
                                   const double* rvec, const double* beta, double beta0, double* energy,
                                   double* forces, double* Ylist, double* dEdr) {
  if (p->chem_flag) return -1; /* the adjoint covers num_elements == 1 */
  int ncoeff = orc_snap_num_coeffs(p->twojmax, p->nspecies, 0);
  double E = beta0;
  if (forces) memset(forces, 0, sizeof(double) * 3 * n);
  double* yr = (double*)calloc(t->idxu_max, sizeof(double));
  double* yi = (double*)calloc(t->idxu_max, sizeof(double));
  for (int64_t i = 0; i < n; i++) {
    double *rij, *wj, *rcutij;
    int* elem;
    int nnb = atom_pairs(p, i, species, offsets, nbr, rvec, &rij, &wj, &rcutij, &elem);
    orc_rt rt;
    rt_alloc(&rt, t, nnb);
    int ie = species[i] + 1;
    const double* bet = beta + (size_t)(ie - 1) * ncoeff;
    compute_ui(p, &rt, ie, rij, wj, rcutij, elem);
    compute_zi(p, &rt);
    compute_bi(p, &rt);
    for (int c = 0; c < ncoeff; c++) E += bet[c] * rt.blist[c];
    memset(yr, 0, sizeof(double) * t->idxu_max);
    memset(yi, 0, sizeof(double) * t->idxu_max);
    for (int jjz = 0; jjz < t->idxz_max; jjz++) {
      const orc_zidx* z = &t->idxz[jjz];
      double c = yi_coef(p, t, bet, z->j1, z->j2, z->j);
      yr[z->jju] += c * rt.zlist_r[jjz];
      yi[z->jju] += c * rt.zlist_i[jjz];
    }
    if (Ylist)
      for (int k = 0; k < t->idxu_max; k++) {
        Ylist[(i * t->idxu_max + k) * 2] = yr[k];
        Ylist[(i * t->idxu_max + k) * 2 + 1] = yi[k];
      }
    for (int ind = 0; ind < nnb; ind++) {
      int64_t jj = nbr[offsets[i] + ind];
      compute_duidrj(p, t, rij + 3 * ind, wj[ind], rcutij[ind], rt.ulist_r + (ind + 1) * t->idxu_max,
                     rt.ulist_i + (ind + 1) * t->idxu_max, rt.dulist_r, rt.dulist_i);
      double s[3] = {0, 0, 0};
      for (int j = 0; j <= t->twojmax; j++) {
        double h[3];
        half_sum_du_z(t, j, t->idxu_block[j], t->idxu_block[j], rt.dulist_r, rt.dulist_i, yr, yi, h);
        for (int k = 0; k < 3; k++) s[k] += h[k];
      }
      double de[3] = {2.0 * s[0], 2.0 * s[1], 2.0 * s[2]};
      if (dEdr) for (int k = 0; k < 3; k++) dEdr[(offsets[i] + ind) * 3 + k] = de[k];
      if (forces)
        for (int k = 0; k < 3; k++) {
          forces[3 * i + k] += de[k];
          forces[3 * jj + k] -= de[k];
        }
    }
    rt_free(&rt);
    free(rij); free(wj); free(rcutij); free(elem);
  }
  if (energy) *energy = E;
  free(yr); free(yi);
  return 0;
}
