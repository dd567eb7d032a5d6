// ipg_snap_driver.h -- host-side state of a SNAP object: the captured
// energy+force step, the batched-descriptor pipeline's tuning and layout,
// and struct ipg_snap itself.  Only the drivers include it, so kernel
// sources do not depend on it.
#pragma once
#include <functional>

#include "ipg_snap.h"

namespace ipg {

// A captured energy+force step (lists in capacity mode, no host
// synchronization inside) replayed while its inputs stay the same buffers.
struct SnapGraph {
  hipGraphExec_t exec = nullptr;
  // what the captured work depends on
  int64_t n = -1;
  const double* pos = nullptr;
  const int32_t* species = nullptr;
  const int32_t* appear = nullptr;
  IpgBox box{};
  int mode = -1, path = -1;
  const double* beta = nullptr;
  double beta0 = 0.0;
  double* energy = nullptr;
  double* forces = nullptr;
  double* aux = nullptr;  // a further output buffer (the batched virial rows)
  bool prof = false;
  uint64_t prof_gen = 0;  // profiling setup the captured event nodes belong to
  uint64_t gen = 0;  // scratch generation the captured pointers belong to
  int64_t capA = 0, capB = 0;
  int ellK = 0;  // neighbor row width of the CUTOFF list's first pass
  double avg_nbr = -1.0;
  int32_t* flag = nullptr;  // bits of the last replay: 1, 2, 8 capacity; 4 missing reverse pair; 16 species
  bool pinned = false;      // the step's last kernel copies the flag to the snap object's host word
  std::vector<ipg_ctx::Pending> events;  // profiling event pairs inside the graph
  bool disabled[2] = {false, false};     // capture failed (without / with profiling events)
  std::string reason;
  int64_t captures = 0, replays = 0, overflows = 0;
};

// Shared by the single-system and batched drivers (ipg_snap.hip).
SnapCut snap_make_cut(const SnapHost& h, int mode);
// dB per atom from the per-pair dB (dB may be NULL); over a symmetric list
// (R.rev) optionally also the per-atom virial rows Wa [n][ncoeff][6].
template <typename T>
void snap_gather_db(ipg_ctx* ctx, const SnapDev& P, const SnapRun& R, const T* dbl, double* dB,
                    const double* pos = nullptr, double* Wa = nullptr);
bool graph_capture(ipg_ctx* ctx, SnapGraph& g, const std::function<void()>& step);
void graph_release(ipg_ctx* ctx, SnapGraph& g);

// Launch shapes of the batched descriptor pipeline, chosen by the autotuner
// (0: not tuned yet).
struct BatchTune {
  int scan_lp = 0, ui_apw = 0, dbdr_threads = 0;
  double ms_scan[3] = {0, 0, 0}, ms_ui[2] = {0, 0}, ms_dbdr[2] = {0, 0};
};
// Host copy of the batch layout last uploaded (skips re-uploading it).
struct BatchLayout {
  std::vector<int64_t> natoms;
  std::vector<double> box;
  std::vector<int32_t> per;
  uint64_t gen = 0;  // scratch generation after the call that uploaded it
};

}  // namespace ipg

struct ipg_snap {
  ipg_ctx* ctx = nullptr;
  ipg::SnapHost host;
  ipg::SnapDev dev{};
  void* dmem = nullptr;  // tables
  double* beta_dev = nullptr;  // expanded beta (ncoeff * nspecies), owned
  int64_t beta_cap = 0;
  ipg::SnapGraph graph;
  ipg::SnapGraph bgraph;       // captured batched-descriptor pipeline
  ipg::BatchTune btune[2];     // per precision (FP64, FP32)
  ipg::BatchLayout blayout;
  int bprec = -1;              // precision the captured batch pipeline was built for
  int64_t bpairs = 0;          // SNAP pairs of the last exact batched run
  int32_t* host_flag = nullptr;  // pinned host word: the flag of the last step that ran the epilogue
};
