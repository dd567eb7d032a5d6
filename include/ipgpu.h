/*
 * ipgpu.h -- C ABI of libipgpu.so, the MI355X (gfx950) force-field engine
 * behind InteratomicPotentials.jl's neighbor-list, pair-potential and SNAP
 * calls.
 *
 * Conventions
 *  - Every entry point returns an int status: IPG_OK (0) or a negative
 *    IPG_ERR_* code.  ipg_last_error() returns a thread-local message for the
 *    last failure on the calling thread.  No C++ exception or abort crosses
 *    the ABI.
 *  - Host-pointer entry points copy inputs to the device, run on the
 *    context's HIP stream and copy results back before returning (they are
 *    synchronous).  The caller owns all host buffers; the library keeps no
 *    host pointer after return.
 *  - *_dev entry points take device pointers for every array (inputs and
 *    outputs) and enqueue work on the context's stream.  Those whose
 *    comment says so synchronize the stream once, at their end, to read a
 *    status word (a list outgrew its buffer, a species index is out of
 *    range); the others return without synchronizing, and
 *    ipg_ctx_synchronize() waits for them.
 *  - Species indices must lie in [0, 64) for pair potentials and in
 *    [0, nspecies) for SNAP.  Host entry points check them before any
 *    work; device entry points check them on the device and return
 *    IPG_ERR_INVALID (the outputs of that call are then unspecified, and no
 *    kernel reads outside its tables).
 *  - Atom indices are 0-based (the Julia wrapper adds 1).
 *  - Positions are AoS x,y,z per atom (Julia Vector{SVector{3,Float64}}
 *    reinterpreted).  Boxes are orthorhombic: box[d] is bounding_box(s)[d][d].
 *  - Calls on one context are serialized by the caller; different contexts
 *    may be used from different threads.  Reductions are deterministic:
 *    repeating a call gives bit-identical results.
 */
#ifndef IPGPU_H
#define IPGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPG_VERSION 100

/* status codes */
#define IPG_OK 0
#define IPG_ERR_INVALID (-1)     /* invalid argument (n < 0, rcut <= 0, NaN, bad box, ...) */
#define IPG_ERR_CAPACITY (-2)    /* output capacity too small; *npairs holds the size needed */
#define IPG_ERR_HIP (-3)         /* HIP runtime error */
#define IPG_ERR_UNSUPPORTED (-4) /* option not available (e.g. twojmax > 10, an adjoint path with chem_flag) */
#define IPG_ERR_NOMEM (-5)       /* device allocation failed */

typedef struct ipg_ctx ipg_ctx;
typedef struct ipg_nlist ipg_nlist;
typedef struct ipg_snap ipg_snap;

/* ------------------------------------------------------------------ */
/* context, errors, device memory, profiling                          */
/* ------------------------------------------------------------------ */
int ipg_get_version(void);
const char* ipg_last_error(void);
int ipg_ctx_create(int device, ipg_ctx** out);
int ipg_ctx_destroy(ipg_ctx* ctx);
int ipg_ctx_synchronize(ipg_ctx* ctx);
int ipg_device_malloc(ipg_ctx* ctx, size_t bytes, void** dptr);
int ipg_device_free(ipg_ctx* ctx, void* dptr);
int ipg_memcpy_h2d(ipg_ctx* ctx, void* dst, const void* src, size_t bytes);
int ipg_memcpy_d2h(ipg_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a plain vector-copy kernel ("device_copy"),
 * enqueued on the context stream; bytes and both pointers multiples of 16.
 * No counterpart in the reference: a kernel of known HBM traffic for
 * calibrating hardware byte counters. */
int ipg_device_copy(ipg_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Per-kernel timing with HIP events recorded on the context stream around
 * every launch (off by default). */
int ipg_prof_enable(ipg_ctx* ctx, int on);
int ipg_prof_reset(ipg_ctx* ctx);
/* Accumulated launches and milliseconds of the kernel named `kernel`
 * (flushes pending events, synchronizing the stream). */
int ipg_prof_query(ipg_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms);
/* Restrict the timing to the kernels named in `names` (comma-separated;
 * NULL or "" times every kernel).  Fewer event pairs disturb a run less,
 * inside a captured graph in particular. */
int ipg_prof_select(ipg_ctx* ctx, const char* names);
/* Enable (default) or disable the replay of captured HIP graphs by the
 * entry points that support it (ipg_snap_energy_forces*,
 * ipg_snap_batch_descriptors*).  With graphs off
 * every call launches its kernels one by one.  No counterpart in the
 * reference (a launch-overhead control of this implementation). */
int ipg_ctx_set_graphs(ipg_ctx* ctx, int on);

/* ------------------------------------------------------------------ */
/* systems (AtomsBase AbstractSystem, reduced to what the path reads)  */
/* ------------------------------------------------------------------ */
typedef struct {
  int64_t natoms;
  const double* positions; /* 3*natoms; host or device pointer per entry point */
  const int32_t* species;  /* natoms species indices into the potential's table; NULL = all 0 */
  double box[3];           /* box edge lengths (only read for periodic dimensions) */
  int32_t periodic[3];     /* periodicity(s): nonzero = periodic */
} ipg_system;

/* ------------------------------------------------------------------ */
/* neighbor lists                                                      */
/* ------------------------------------------------------------------ */
enum { IPG_LIST_HALF = 0, IPG_LIST_FULL = 1 };

/* Replaces src/nnlist.jl:23-46 neighborlist(s, rcutoff) (half list, j > i)
 * and its full (j != i) variant.  Membership is the reference's:
 * PeriodicEuclidean(L) distance <= rcut.  disp[q] = get_displacement(L,
 * x_i, x_j) (src/nnlist.jl:18-21).  offsets has natoms+1 entries; j are
 * sorted per atom.  If cap < the pair count, returns IPG_ERR_CAPACITY with
 * *npairs = count (offsets are still written). disp may be NULL. */
int ipg_neighborlist(ipg_ctx* ctx, const ipg_system* sys, double rcut, int kind, int64_t* offsets, int32_t* nbr,
                     double* disp, int64_t cap, int64_t* npairs);

/* Device-resident list for repeated evaluations (sys holds device pointers). */
int ipg_nlist_build_dev(ipg_ctx* ctx, const ipg_system* sys, double rcut, int kind, ipg_nlist** out);
int ipg_nlist_info(const ipg_nlist* list, int64_t* natoms, int64_t* npairs, const int64_t** d_offsets,
                   const int32_t** d_nbr);
int ipg_nlist_destroy(ipg_nlist* list);
/* Verlet-skin list for MD callers, which the reference serves by rebuilding
 * the list on every call (src/types/empirical_potential.jl:21): the list
 * holds every pair within rcut + skin of the positions at its last build.
 * ipg_nlist_update_dev measures the largest PeriodicEuclidean displacement
 * of any atom since then (one synchronization) and rebuilds when it exceeds
 * skin/2 or the box changed, so every pair within rcut of the current
 * positions is in the list; *rebuilt (may be NULL) says whether it rebuilt.
 * ipg_pair_eval_dev over such a list tests each pair's membership against
 * the terms' own cutoffs with the reference's distance formula, so its
 * result equals a fresh list's up to the summation order.  No counterpart
 * in the reference. */
int ipg_nlist_build_skin_dev(ipg_ctx* ctx, const ipg_system* sys, double rcut, double skin, int kind,
                             ipg_nlist** out);
int ipg_nlist_update_dev(ipg_nlist* list, const ipg_system* sys, int32_t* rebuilt);

/* ------------------------------------------------------------------ */
/* pair potentials                                                     */
/* ------------------------------------------------------------------ */
enum {
  IPG_PAIR_LJ = 0,          /* p = {eps, sigma}           src/empirical/lj.jl:35-36      */
  IPG_PAIR_MORSE = 1,       /* p = {D, alpha, sigma}      src/empirical/morse.jl:36-39   */
  IPG_PAIR_ZBL = 2,         /* p = {Z1, Z2, e}            src/empirical/zbl.jl:27-36     */
  IPG_PAIR_BORN_MAYER = 3,  /* p = {A, rho, sigma, C, D}  src/empirical/bm.jl:38-39      */
  IPG_PAIR_COULOMB = 4,     /* p = {q1, q2}               src/empirical/coulomb.jl:36-37 */
  IPG_PAIR_MOCK = 5         /* p = {seed}: V = R^2 seed   test/mocks.jl:22-34            */
};
enum { IPG_FP64 = 0, IPG_FP32 = 1 };

typedef struct {
  int32_t kind;
  int32_t reserved;
  double coeff;          /* LinearCombinationPotential coefficient (1 for a plain potential) */
  double rcut;           /* get_rcutoff(p), atomic units */
  double p[6];           /* parameters, atomic units (austrip'ed as the constructors do) */
  uint64_t species_mask; /* bit s set <=> species s in get_species(p); all ones = missing */
} ipg_pair_term;

/* Replaces energy_and_force / virial_stress (src/types/empirical_potential.jl
 * :49-67) and their LinearCombinationPotential sums
 * (src/mixed/linear_combination_potential.jl:27-39), fused: one list at the
 * largest rcut, every term evaluated per pair with its own rcut and species
 * filter.  Energy in hartree, forces hartree/bohr (3*natoms), virial in
 * Voigt order [xx,yy,zz,zy,zx,yx]; any output may be NULL.  precision
 * IPG_FP32 evaluates the pair terms in float with double reductions. */
int ipg_pair_energy_forces(ipg_ctx* ctx, const ipg_system* sys, const ipg_pair_term* terms, int nterms,
                           int precision, double* energy, double* forces, double* virial);
/* Device variant: every pointer (sys arrays, energy[1], forces[3N],
 * virial[6]) is a device pointer; terms is a host array.  Synchronizes once
 * at its end to read the status word (neighbor-row overflow, which repeats
 * the evaluation with a counted list, and out-of-range species). */
int ipg_pair_energy_forces_dev(ipg_ctx* ctx, const ipg_system* sys, const ipg_pair_term* terms, int nterms,
                               int precision, double* energy, double* forces, double* virial);
/* Pair kernel alone over a prebuilt device full list (IPG_LIST_FULL at a
 * cutoff >= every term's rcut); pointers as in the _dev variant.
 * Synchronizes once at its end to read the status word (out-of-range
 * species). */
int ipg_pair_eval_dev(ipg_ctx* ctx, const ipg_nlist* full_list, const ipg_system* sys, const ipg_pair_term* terms,
                      int nterms, int precision, double* energy, double* forces, double* virial);

/* The pair kernel over a prebuilt device full list, as ipg_pair_eval_dev,
 * without any device allocation: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Its partial sums and
 * status word live in a static device array of 1 MiB per GPU, held until
 * the process ends; a list whose pair kernel needs more than
 * (1 MiB - 512) / 56 = 18715 blocks (about 299k atoms at 48 or more
 * neighbors each, 1.19M atoms below 12) returns IPG_ERR_INVALID.  The
 * results equal ipg_pair_eval_dev's bit for bit.  It does not synchronize
 * and reports no device status: species indices outside [0, 64) make an
 * atom take part in no term (no kernel reads outside its tables).  Every
 * call on a GPU shares that GPU's array, so calls that overlap on one GPU
 * (other threads, other streams) race; calls on one stream are ordered.
 * Profiling events are not recorded. */
int ipg_pair_eval_noalloc_dev(ipg_ctx* ctx, const ipg_nlist* full_list, const ipg_system* sys,
                              const ipg_pair_term* terms, int nterms, int precision, double* energy, double* forces,
                              double* virial);

/* ------------------------------------------------------------------ */
/* SNAP                                                                */
/* ------------------------------------------------------------------ */
typedef struct {
  int32_t twojmax;
  int32_t nspecies;
  double rcutfac, rmin0, rfac0;
  const double* radii;  /* host, nspecies */
  const double* weight; /* host, nspecies */
  int32_t chem_flag, bzero_flag, bnorm_flag, switch_flag, wselfall_flag;
} ipg_snap_params;

/* SNAP neighbor-list modes.  REFERENCE reproduces utl.jl:38-101 as-is
 * (inrange radius sqrt(cutmax) :81, element lookup by first appearance :71
 * and by enumerate index :86).  CUTOFF queries radius cutmax and uses each
 * neighbor's own species. */
enum { IPG_SNAP_LIST_REFERENCE = 0, IPG_SNAP_LIST_CUTOFF = 1 };
enum { IPG_SNAP_PLAIN_FP64 = 0, IPG_SNAP_ADJOINT_FP64 = 1, IPG_SNAP_ADJOINT_FP32 = 2 };

/* Replaces SNAP(...) construction + initialize_prebuilt_arrays
 * (src/BasisSystems/SNAP/snap.jl:20-36, types/prebuilt_arrays.jl:48-321):
 * builds the index, Clebsch-Gordan and rootpq tables in FP64 and uploads
 * them.  twojmax in [0, 10] (from 11 on the reference's Int64 factorial
 * products in init_clebsch_gordan overflow and it stops with an error:
 * IPG_ERR_UNSUPPORTED).  chem_flag = true (bnorm_flag must equal it,
 * snap.jl:24-27) and twojmax 9, 10 run runtime-table FP64 kernels that
 * reproduce the reference's multi-element arithmetic as written
 * (compute_zi accumulates every element pair into the first Z block,
 * compute_z.jl:46-47); for those configurations the energy/force/virial
 * calls accept IPG_SNAP_PLAIN_FP64 only (the reference's beta . dB route;
 * other paths: IPG_ERR_UNSUPPORTED), the batched descriptors IPG_FP64
 * only, and ipg_snap_force_vjp returns IPG_ERR_UNSUPPORTED. */
int ipg_snap_create(ipg_ctx* ctx, const ipg_snap_params* params, ipg_snap** out);
int ipg_snap_destroy(ipg_snap* snap);
/* ncoeff = get_num_snap_coeffs (utl.jl:19-37); table sizes. */
int ipg_snap_info(const ipg_snap* snap, int32_t* ncoeff, int32_t* idxu_max, int32_t* idxz_max, int32_t* idxb_max);

/* Replaces neighborlist(A, snap) (utl.jl:38-101): full list, sorted j,
 * R = |r|, r = x_j - x_i (min image).  appear (may be NULL = species) gives
 * each atom's index in unique(atomic_symbol(A)), used by REFERENCE mode. */
int ipg_snap_neighborlist(ipg_snap* snap, const ipg_system* sys, int mode, const int32_t* appear, int64_t* offsets,
                          int32_t* nbr, double* R, double* r, int64_t cap, int64_t* npairs);

/* Replaces compute_local_descriptors (snap.jl:103-124): B natoms x ncoeff. */
int ipg_snap_local_descriptors(ipg_snap* snap, const ipg_system* sys, int mode, const int32_t* appear, double* B);
/* Replaces compute_force_descriptors (snap.jl:126-167) with the row fix of
 * compute_all_descriptors (snap.jl:261): dB natoms x (ncoeff*nspecies) x 3,
 * dB[k] = -d(sum_i B_i)/dx_k in the block of the species of i. */
int ipg_snap_force_descriptors(ipg_snap* snap, const ipg_system* sys, int mode, const int32_t* appear, double* dB);
/* Replaces compute_all_descriptors (snap.jl:226-285): B, dB as above and
 * W (ncoeff*nspecies) x 6 summed over atoms, rows [x dBx, y dBy, z dBz,
 * z dBy, z dBx, y dBx] with raw positions (:267-281).  Outputs may be NULL. */
int ipg_snap_all_descriptors(ipg_snap* snap, const ipg_system* sys, int mode, const int32_t* appear, double* B,
                             double* dB, double* W);
/* Replaces energy_and_force(s, LBasisPotential(beta, beta0, snap))
 * (src/BasisPotentials/basis_potentials.jl:32-41, linear_potential.jl:20-35):
 * E = beta0 + sum_i beta_{s(i)} . B_i, F_k = sum_c beta_c dB[k][:, c].
 * nbeta = ncoeff (one beta for every species, as the reference) or
 * ncoeff*nspecies (one block per species).  path selects the plain FP64
 * route (dB then beta, as the reference does) or the adjoint route
 * (Y = dE/dU, no per-pair dB) in FP64 or FP32 (FP64 reductions).  forces
 * 3*natoms; either output may be NULL. */
int ipg_snap_energy_forces(ipg_snap* snap, const ipg_system* sys, int mode, const int32_t* appear,
                           const double* beta, int64_t nbeta, double beta0, int path, double* energy,
                           double* forces);
/* Replaces virial_stress(s, ::BasisPotential) and virial(s, ...)
 * (src/BasisPotentials/basis_potentials.jl:43-57) for the linear potential,
 * W * beta (linear_potential.jl:37-49) with W of compute_all_descriptors
 * (snap.jl:267-281: raw positions, rows [xx, yy, zz, zy, zx, yx]); the
 * reference's own call passes the potential where compute_virial_descriptors
 * expects the basis (a MethodError), so this is its intended result.
 * virial[6]; virial(s, lbp) is the sum of the six entries.  path as in
 * ipg_snap_energy_forces: the plain route forms W from the per-pair dB and
 * contracts it with beta on the device; the adjoint routes contract
 * dE_i/dr_ij = beta_s(i) . dB_i/dr_ij with x_i - x_j directly. */
int ipg_snap_virial_stress(ipg_snap* snap, const ipg_system* sys, int mode, const int32_t* appear, const double* beta,
                           int64_t nbeta, int path, double* virial);
/* Energy, forces and virial stress of the linear SNAP potential from one
 * evaluation (ipg_snap_energy_forces and ipg_snap_virial_stress in one
 * pass over the per-pair derivatives).  Any output may be NULL. */
int ipg_snap_energy_forces_virial(ipg_snap* snap, const ipg_system* sys, int mode, const int32_t* appear,
                                  const double* beta, int64_t nbeta, double beta0, int path, double* energy,
                                  double* forces, double* virial);
/* Device variant: sys arrays, beta (ncoeff*nspecies, expanded), energy[1]
 * and forces[3N] are device pointers.  For a fully periodic system the
 * step is captured into a HIP graph on the first call and replayed while
 * the same buffers, box, mode, path and beta0 are passed (an MD loop);
 * neighbor lists then run at capacities sized from the first call's pair
 * counts (an overflowing call is recomputed with exact sizes).  Every call
 * synchronizes the context stream once, at its end, to read the list's
 * flag word (capacity overflow, and in CUTOFF mode a pair whose reverse pair
 * is missing, which is then fixed up), so the outputs are complete when it
 * returns.  Results equal the uncaptured call's bit for bit. */
int ipg_snap_energy_forces_dev(ipg_snap* snap, const ipg_system* sys, int mode, const double* beta, double beta0,
                               int path, double* energy, double* forces);
/* Captured-step statistics of a snap object: state 0 = no graph, 1 = graph
 * ready, 2 = capture failed (calls run uncaptured); captures, replays and
 * overflow recomputations so far.  reason (may be NULL) receives the
 * capture failure message, truncated to reason_len bytes.  No counterpart
 * in the reference. */
int ipg_snap_graph_info(const ipg_snap* snap, int32_t* state, int64_t* captures, int64_t* replays,
                        int64_t* overflows, char* reason, int64_t reason_len);
/* Descriptors of many configurations in one call: replaces a fitting loop
 * that maps compute_local_descriptors and compute_force_descriptors
 * (src/BasisSystems/SNAP/snap.jl:103-124, :126-224, computed as
 * compute_all_descriptors does, :226-285; generic functions at
 * src/api.jl:151-173) over a data set, in CUTOFF list mode.  systems[c]
 * are nconf host-pointer systems.  Outputs are packed in configuration
 * order, atoms of configuration c following those of c-1:
 * B [natoms_total][ncoeff], dB [natoms_total][ncoeff*nspecies][3] (the
 * Julia 3 x M matrix of each atom), erow [nconf][nspecies][ncoeff], the
 * per-configuration sums of B in species blocks (the row the energy of
 * linear_potential.jl:20-26 contracts with beta), and wrow
 * [nconf][ncoeff*nspecies][6], each configuration's W of
 * compute_all_descriptors (snap.jl:267-281, raw positions, rows
 * [x dBx, y dBy, z dBz, z dBy, z dBx, y dBx]).  Any output may be NULL.
 * precision IPG_FP64 runs the FP64 kernels; IPG_FP32 computes U, Z and the
 * per-pair dB in FP32 with FP64 sums for B, dB and erow.  The first call of
 * each precision times candidate launch shapes on the batch (autotuning);
 * later calls with the same layout replay one captured HIP graph. */
int ipg_snap_batch_descriptors(ipg_snap* snap, int64_t nconf, const ipg_system* systems, int precision, double* B,
                               double* dB, double* erow, double* wrow);
/* Device variant: the layout (natoms[nconf], box[nconf][3],
 * periodic[nconf][3]) is on the host; pos (3*natoms_total), species
 * (natoms_total, may be NULL) and the outputs are device pointers in the
 * packed order above.  Species indices are checked on the device.  The
 * call synchronizes the context stream once at its end (it reads the
 * neighbor-list flag word and recomputes exactly on a capacity overflow). */
int ipg_snap_batch_descriptors_dev(ipg_snap* snap, int64_t nconf, const int64_t* natoms, const double* box,
                                   const int32_t* periodic, const double* pos, const int32_t* species, int precision,
                                   double* B, double* dB, double* erow, double* wrow);
/* Launch shapes the autotuner chose for a precision (0 before the first
 * call): lanes per atom of the neighbor scan, atoms per wave of the U
 * kernel and block size of the per-pair dB kernel; graph_state as in
 * ipg_snap_graph_info, the graph replays so far and the SNAP pair count of
 * the last exact (uncaptured) batched run.  Any output may be NULL.  No
 * counterpart in the reference. */
int ipg_snap_batch_info(const ipg_snap* snap, int precision, int32_t* scan_lanes, int32_t* ui_atoms_per_wave,
                        int32_t* dbdr_threads, int32_t* graph_state, int64_t* replays, int64_t* npairs);
/* Reverse mode of F(x, beta) for force matching (a ChainRulesCore rrule of
 * the force call): given the cotangent gF = dL/dF (3*natoms), returns
 * gbeta = gF^T dF/dbeta (nbeta entries, same layout as beta) and
 * gpos = gF^T dF/dx (3*natoms), exactly (forward-mode tangents through the
 * adjoint, no finite differences).  precision IPG_FP64 or IPG_FP32.  Either
 * output may be NULL.  This has no counterpart in the reference (its NN/
 * fitting stack differentiates through Zygote on the CPU). */
int ipg_snap_force_vjp(ipg_snap* snap, const ipg_system* sys, int mode, const int32_t* appear, const double* beta,
                       int64_t nbeta, const double* gF, int precision, double* gbeta, double* gpos);

#ifdef __cplusplus
}
#endif
#endif
