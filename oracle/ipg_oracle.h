/*
 * ipg_oracle.h -- CPU FP64 restatement of InteratomicPotentials.jl's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (libipgpu.so) never calls it.
 *
 * Every function restates one reference function; the file:line it follows is
 * given beside it (paths relative to the reference repository root).
 * Arithmetic is plain IEEE double, compiled with -ffp-contract=off so that
 * a*b+c is never fused (Julia does not fuse either).
 *
 * Pinning status (how each function is checked, see tests/):
 *   orc_jl_mod / orc_pe_sumsq / orc_get_displacement  -- test/unit/nnlist.jl
 *       golden vectors, brute-force identities.
 *   orc_nlist_*                   -- test/unit/nnlist.jl golden list, brute force
 *                                    vs cell-binned equality.
 *   orc_pair_*                    -- test/unit/types/empirical_potential.jl
 *                                    (mock potential, species filter, virial),
 *                                    LJ 150/1000 cluster known answers, closed
 *                                    forms, finite differences of V; Born-Mayer
 *                                    by FD of its exponential part and an exact
 *                                    value of its power part; ZBL with e != 1
 *                                    against the textbook screening function
 *                                    at d ln e and its derivative.
 *   orc_snap_tables_*             -- sympy exact Clebsch-Gordan, closed-form
 *                                    table sizes and rootpq.
 *   orc_snap_atom_*, descriptors  -- Wigner-D unitarity and closed form,
 *                                    rotation/permutation invariance of B,
 *                                    numpy brute-force bispectrum, central
 *                                    finite differences for dB; bzero as the
 *                                    isolated atom's B (B - bzero = 0); the
 *                                    as-is chem_flag path on isolated atoms
 *                                    against its hand-derived closed form.
 *   orc_snap_adjoint_*            -- equality with beta . dB (plain path) and
 *                                    finite differences of E.
 *   orc_snap_neighborlist         -- a hand-derived list of a 6-atom,
 *                                    two-species, mixed-boundary system in
 *                                    both modes (tests/golden/
 *                                    snap_list_hand.json).
 */
#ifndef IPG_ORACLE_H
#define IPG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* distance semantics                                                 */
/* ------------------------------------------------------------------ */
double orc_jl_mod(double x, double y);
/* Distances.PeriodicEuclidean, squared sum before sqrt. L[d] = +INF for a
 * non-periodic dimension. */
double orc_pe_sumsq(const double a[3], const double b[3], const double L[3]);
void orc_get_displacement(const double L[3], const double x[3], const double y[3], double out[3]);

/* ------------------------------------------------------------------ */
/* neighbor lists (src/nnlist.jl:23-46)                               */
/* ------------------------------------------------------------------ */
enum { ORC_LIST_HALF = 0, ORC_LIST_FULL = 1 };
/* O(N^2) brute force.  Returns the number of pairs found (even when it is
 * larger than cap, in which case nbr/disp are only partially written), or -1
 * on invalid input.  offsets has n+1 entries.  j are 0-based and sorted. */
int64_t orc_nlist_brute(const double* pos, int64_t n, const double L[3], double rcut, int kind,
                        int64_t* offsets, int32_t* nbr, double* disp, int64_t cap);
/* Cell-binned O(N) variant with identical output (sorted j per atom). */
int64_t orc_nlist_cells(const double* pos, int64_t n, const double L[3], double rcut, int kind,
                        int64_t* offsets, int32_t* nbr, double* disp, int64_t cap);

/* ------------------------------------------------------------------ */
/* pair potentials (src/empirical, src/types/empirical_potential.jl)        */
/* ------------------------------------------------------------------ */
enum {
  ORC_PAIR_LJ = 0,       /* p = {eps, sigma}            lj.jl:35-36      */
  ORC_PAIR_MORSE = 1,    /* p = {D, alpha, sigma}       morse.jl:36-39   */
  ORC_PAIR_ZBL = 2,      /* p = {Z1, Z2, e}             zbl.jl:27-36     */
  ORC_PAIR_BM = 3,       /* p = {A, rho, sigma, C, D}   bm.jl:38-39      */
  ORC_PAIR_COULOMB = 4,  /* p = {q1, q2}                coulomb.jl:36-37 */
  ORC_PAIR_MOCK = 5      /* p = {seed}                  test/mocks.jl:22-34 */
};
typedef struct {
  int32_t kind;
  int32_t pad_;
  double coeff;          /* LinearCombinationPotential coefficient */
  double rcut;           /* atomic units (bohr) */
  double p[6];           /* atomic units */
  uint64_t species_mask; /* bit s set <=> species s in get_species(p); ~0 = missing */
} orc_pair_term;

double orc_pair_V(const orc_pair_term* t, double R);
void orc_pair_F(const orc_pair_term* t, double R, const double r[3], double out[3]);

/* energy_and_force / virial_stress for a LinearCombinationPotential of
 * nterms terms: each term rebuilds its own half list, exactly as
 * linear_combination_potential.jl:27-39 does.  e, f (3n), v (6) may be NULL.
 * use_cells selects orc_nlist_cells instead of the brute-force list. */
int orc_pair_energy_force_virial(const double* pos, int64_t n, const double L[3], const int32_t* species,
                                 const orc_pair_term* terms, int nterms, int use_cells,
                                 double* e, double* f, double* v);
/* The same over nthreads OpenMP threads, as the reference threads
 * _perform_pairwise (empirical_potential.jl:28-46): contiguous blocks of
 * atoms, per-thread sums added in thread order (the rounding depends on
 * the thread count only).  The cell list rows are built in parallel too
 * (the same list).  Used to time the CPU baseline. */
int orc_pair_energy_force_virial_mt(const double* pos, int64_t n, const double L[3], const int32_t* species,
                                    const orc_pair_term* terms, int nterms, int use_cells, int nthreads,
                                    double* e, double* f, double* v);

/* ------------------------------------------------------------------ */
/* SNAP (src/BasisSystems/SNAP)                                        */
/* ------------------------------------------------------------------ */
typedef struct {
  int j1, j2, j, ma1min, ma2max, mb1min, mb2max, na, nb, jju;
} orc_zidx;

typedef struct {
  int twojmax, num_elements;
  int idxcg_max, idxu_max, idxb_max, idxz_max;
  int* idxcg_block; /* jdim^3, [j1][j2][j] */
  int* idxu_block;  /* jdim */
  int* idxb_block;  /* jdim^3 */
  int* idxz_block;  /* jdim^3 */
  int* idxb;        /* idxb_max x 3 (j1,j2,j) */
  orc_zidx* idxz;   /* idxz_max */
  double* cglist;   /* idxcg_max */
  double* rootpq;   /* (twojmax+2)^2, [p][q] */
  double wself;
  double* bzero;    /* jdim */
} orc_snap_tables;

typedef struct {
  int32_t twojmax;
  int32_t nspecies;
  double rcutfac, rmin0, rfac0;
  const double* radii;   /* nspecies */
  const double* weight;  /* nspecies */
  int32_t chem_flag, bzero_flag, bnorm_flag, switch_flag, wselfall_flag;
} orc_snap_params;

orc_snap_tables* orc_snap_tables_build(int twojmax, int nspecies, int chem, int bzero, int bnorm);
void orc_snap_tables_free(orc_snap_tables* t);
int orc_snap_num_coeffs(int twojmax, int nspecies, int chem);

enum { ORC_SNAP_LIST_REFERENCE = 0, ORC_SNAP_LIST_CUTOFF = 1 };
/* utl.jl:38-101.  species: index into snap.species per atom (0-based);
 * appear: index into unique(atomic_symbol(A)) (first appearance order). */
int64_t orc_snap_neighborlist(const double* pos, int64_t n, const double L[3], const int32_t* species,
                              const int32_t* appear, const orc_snap_params* p, int mode,
                              int64_t* offsets, int32_t* nbr, double* R, double* rvec, int64_t cap);

/* compute_local_descriptors (snap.jl:103-124): B is n x ncoeff. */
int orc_snap_local_descriptors(const orc_snap_params* p, const orc_snap_tables* t, int64_t n,
                               const int32_t* species, const int64_t* offsets, const int32_t* nbr,
                               const double* rvec, double* B);
/* compute_all_descriptors (snap.jl:226-285), including the U-row fix of
 * line 261.  B: n x ncoeff; dB: n x (ncoeff*nspecies) x 3 (Julia 3 x M
 * column-major per atom); W: (ncoeff*nspecies) x 6 (Julia 6 x M
 * column-major), positions are the raw (unwrapped) ones as in :267-268.
 * Any output may be NULL. */
int orc_snap_all_descriptors(const orc_snap_params* p, const orc_snap_tables* t, int64_t n,
                             const double* pos, const int32_t* species, const int64_t* offsets,
                             const int32_t* nbr, const double* rvec, double* B, double* dB, double* W);
/* Energy and forces of a linear SNAP potential by the reference's plain
 * route (snap.jl:237-265 then linear_potential.jl:20-35): per atom U, Z, B
 * and E_i = beta_s . B_i; per neighbor dB/dr_ij, f = beta_s . dB/dr_ij,
 * F_i += f, F_j -= f, over the first nloop of the n atoms (forces has 3n
 * entries).  nthreads > 1 splits the atoms over OpenMP threads (forces then
 * accumulate with atomic adds, so their rounding depends on the schedule).
 * Used to time the CPU baseline and checked against the adjoint. */
int orc_snap_plain_energy_forces(const orc_snap_params* p, const orc_snap_tables* t, int64_t n, int64_t nloop,
                                 const int32_t* species, const int64_t* offsets, const int32_t* nbr,
                                 const double* rvec, const double* beta, int nthreads, double* energy,
                                 double* forces);
/* Energy and forces of a linear SNAP potential through the adjoint
 * Y = dE/dU (derived, compute_yi.jl is a specification only).  beta has
 * ncoeff*nspecies entries (block s is used for atoms of species s).
 * energy = beta0 + sum_i beta_{s(i)} . B_i;  F_k = -dE/dx_k.
 * Ylist (optional): n x idxu_max complex (re, im interleaved), zero outside
 * the left half. dEdr (optional): per pair, 3 doubles. */
int orc_snap_adjoint_energy_forces(const orc_snap_params* p, const orc_snap_tables* t, int64_t n,
                                   const int32_t* species, const int64_t* offsets, const int32_t* nbr,
                                   const double* rvec, const double* beta, double beta0,
                                   double* energy, double* forces, double* Ylist, double* dEdr);
/* Per-atom building blocks, exposed for the tests.  For one atom with nnb
 * neighbors: rij (nnb x 3), wj, rcutij (nnb).  ulisttot: idxu_max complex
 * (interleaved), zlist: idxz_max complex, blist: idxb_max. */
int orc_snap_atom_uzb(const orc_snap_params* p, const orc_snap_tables* t, int ielem, int nnb,
                      const double* rij, const double* wj, const double* rcutij,
                      double* ulisttot, double* zlist, double* blist);
/* Wigner U for one neighbor (compute_uarray, compute_ui.jl:50-129):
 * ulist idxu_max complex. */
int orc_snap_uarray(const orc_snap_tables* t, double rfac0, double rmin0, double rcut,
                    const double r3[3], double* ulist);
/* dU for one neighbor (compute_duidrj/compute_duarray, compute_ui.jl:186-347)
 * dulist: idxu_max x 3 complex, layout [jju][k] (re, im). */
int orc_snap_duarray(const orc_snap_params* p, const orc_snap_tables* t, double rcut, double wj,
                     const double r3[3], double* dulist);

/* Wall-clock helper for the CPU baseline. */
double orc_wtime(void);

#ifdef __cplusplus
}
#endif
#endif
