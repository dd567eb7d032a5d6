// ipg_nlist.h -- device neighbor lists shared by the pair and SNAP paths.
#pragma once
#include "ipg_internal.h"

namespace ipg {

enum ListKind { KIND_HALF = 0, KIND_FULL = 1, KIND_FULL_SELF = 2 };

// Device-resident system (pointers are device pointers).
struct DevSystem {
  int64_t n = 0;
  const double* pos = nullptr;
  const int32_t* species = nullptr;  // may be null (all species 0)
  IpgBox box{};
};

// A CSR neighbor list on the device.
struct DevList {
  int64_t n = 0, npairs = 0;
  int64_t* off = nullptr;  // n+1
  int32_t* nbr = nullptr;  // npairs
  double* disp = nullptr;  // 3*npairs or null
  bool exact = true;       // npairs is the pair count (false: a capacity, the count is off[n])
  int ellK = 0;            // > 0: nbr not yet written, rows of width ellK are in ell (counts in cnt)
  int ellK_req = 0;        // row width the first pass used (0: no row buffer)
  const int32_t* ell = nullptr;
  const int32_t* cnt = nullptr;
  bool ell_sorted = false; // the rows in ell are sorted by j
  bool deferred = false;   // rows returned before the overflow bit was read: npairs is an estimate, off is
                           // unset, and the caller reads bit 8 of its flag after using the rows
};

DevSystem make_dev_system(const ipg_system* s);  // s holds device pointers
DevSystem upload_system(ipg_ctx* ctx, const ipg_system* s, const char* tag);  // s holds host pointers

// Builds a list with membership PeriodicEuclidean distance <= rcut.
// Storage comes from context scratch buffers named after `tag`.  sorted:
// j ascending per atom.  want_disp: store get_displacement(x_i, x_j).
// ellK != 0 (unsorted lists, with capflag): the scan writes each atom's
// neighbors into rows of width ellK (-1: from the density) in the order the
// CSR list would hold them; when no row overflowed the list comes back with
// ellK > 0, the rows in ell and their lengths in cnt, and nbr unfilled.
// defer (exact sizing with rows only): return the rows right after the
// scan, with no offsets and no host synchronization (DevList::deferred).
DevList build_list(ipg_ctx* ctx, const DevSystem& s, double rcut, int kind, bool sorted, bool want_disp,
                   const char* tag, int64_t cap = 0, int32_t* capflag = nullptr, int ellK = 0,
                   bool defer = false);

// Same, into caller-owned device buffers allocated by the function (for
// ipg_nlist objects); frees nothing.
DevList build_list_owned(ipg_ctx* ctx, const DevSystem& s, double rcut, int kind, bool sorted, bool want_disp);

// SNAP list of utl.jl:38-101.  rvec = x_j - x_i (min image), R = |rvec|.
struct SnapListDev {
  DevList list;       // nbr/off of the SNAP list
  double* R = nullptr;
  double* rvec = nullptr;
  int32_t* center = nullptr;   // center atom of each pair
  int64_t* roff = nullptr;     // reverse list offsets (n+1): pairs whose neighbor is the atom
  int32_t* rpair = nullptr;    // reverse list pair indices, ascending
  int32_t* rev = nullptr;      // index of the reverse pair (j -> i) of each pair (CUTOFF mode)
  int64_t npairs_a = 0;        // size of the stage-A list (count or capacity)
  int ellK = 0;                // CUTOFF mode: row width of the first pass (0: none)
};
// What build_snap_list forms besides nbr/off/rvec/center.
enum SnapListNeed { SNAPL_R = 1, SNAPL_REVERSE = 2, SNAPL_SYM = 4 };
struct SnapCut {
  int ns;
  double cutsq[64];  // ns x ns (ns <= 8)
  double query;      // inrange radius (sqrt(cutmax) in REFERENCE mode, cutmax otherwise)
  int reference;     // 1: element lookup as utl.jl:70,86 (appear[] of atom and of rank m)
};
// need: SnapListNeed bits (R: distances; REVERSE: roff/rpair; SYM: rev,
// CUTOFF mode only -- REFERENCE-mode lists are not symmetric, so there SYM
// gives roff/rpair).  CUTOFF mode builds the list with one species-aware
// cell scan; REFERENCE mode runs the two stages of utl.jl.  capA/capB > 0
// select capacity mode (no host readback): the lists hold at most that many
// pairs (CUTOFF mode uses capB), their offsets are clamped to it and *flag
// gets bit 1 (stage A / CUTOFF list) or 2 (stage B) when the true count was
// larger; bit 4 marks a missing reverse pair.  CUTOFF mode's first pass
// also writes rows of width ellK (0: estimated from the density) and builds
// the sorted list from them; bit 8 marks a row wider than ellK.  With
// clear_flag, CUTOFF mode zeroes *flag in the list's first launch (the
// caller's own clear is then not needed); REFERENCE mode ignores it.
SnapListDev build_snap_list(ipg_ctx* ctx, const DevSystem& s, const int32_t* d_appear, const SnapCut& cut, int need,
                            const char* tag, int64_t capA = 0, int64_t capB = 0, int32_t* flag = nullptr,
                            int ellK = 0, bool clear_flag = false);

// Packed configurations on the device: atoms of configuration c are
// start[c] .. start[c+1]-1, cfg[i] is the configuration of atom i, boxes[c]
// its box; expect is the largest expected neighbor count (row width hint).
struct BatchSys {
  int64_t n = 0, nconf = 0;
  const double* pos = nullptr;
  const int32_t* species = nullptr;
  const int32_t* cfg = nullptr;
  const int64_t* start = nullptr;
  const IpgBox* boxes = nullptr;
  double expect = 0.0;
};
// CUTOFF-mode SNAP list of packed configurations (rows sorted by j).  cap,
// flag and ellK as build_snap_list; lp: lanes per atom (16, 32 or 64).
// need: SNAPL_SYM gives rev (bit 4 of *flag marks a missing reverse pair),
// SNAPL_REVERSE the general reverse rows.
SnapListDev build_snap_list_batch(ipg_ctx* ctx, const BatchSys& b, const SnapCut& cut, int need, const char* tag,
                                  int64_t cap, int32_t* flag, int ellK, int lp);

// center[q] and the reverse list of any CSR list (npairs may be a
// capacity: the count used is off[n] on the device).
void build_center_and_reverse(ipg_ctx* ctx, int64_t n, int64_t npairs, const int64_t* off, const int32_t* nbr,
                              int32_t* center, int64_t* roff, int32_t* rpair, const char* tag);

}  // namespace ipg

struct ipg_nlist {
  ipg_ctx* ctx = nullptr;
  ipg::DevList list;
  double rcut = 0;  // radius the list was built at (cut + skin)
  double cut = 0;   // interaction cutoff the list serves
  double skin = 0;  // Verlet skin (0: a plain list)
  int kind = 0;
  double* ref = nullptr;  // positions at the last build (skin lists; device)
  IpgBox box{};           // box of the last build
  uint64_t* dmax = nullptr;  // device word: largest squared displacement (as bits)
  int64_t builds = 0, updates = 0;
};
