// Engine facade of the reference API (src/Engine.h:16-70): owns the
// algorithm / mode / pattern choice and forwards to the solvers.
// Additions: RunCF (the reference's CFSolver is not wired into its Engine)
// and motif_counts() (the reference declares motif_result but never fills it,
// Engine.h:28).
#pragma once

#include <vector>

#include "graph.h"
#include "io.h"
#include "sm_pattern.h"
#include "solvers.h"

namespace project_AntiRF {

template <typename VID, typename VLABEL>
class Engine {
 public:
  Engine(Graph<VID, VLABEL>& hg, int n_dev, algorithms al, modes cal_m)
      : hg_(hg), n_devices_(n_dev), algo_(al), cal_mode_(cal_m) {}
  Engine(Graph<VID, VLABEL>& hg, int n_dev, algorithms al, modes cal_m, int k)
      : hg_(hg), n_devices_(n_dev), algo_(al), cal_mode_(cal_m), motif_k_(k) {}
  Engine(Graph<VID, VLABEL>& hg, int n_dev, algorithms al, modes cal_m, Pattern p)
      : hg_(hg), n_devices_(n_dev), algo_(al), cal_mode_(cal_m), s_pattern_(p) {}

  // Engine::RunTC (Engine.h:47-56): orients the graph (and, for v_centric,
  // relabels it by out-degree) before counting; the graph is mutated.
  uint64_t RunTC() {
    if (cal_mode_ == v_centric) {
      hg_.orientation(false);
      hg_.SortCSRGraph(true);
    } else {
      hg_.orientation();
    }
    TCSolver(hg_, result_, n_devices_, cal_mode_);
    return result_;
  }

  uint64_t RunPatternEnum() {
    PatternSolver(hg_, result_, n_devices_, cal_mode_, s_pattern_, &motif_result_);
    return result_;
  }

  uint64_t RunG2Miner() {
    G2MinerSolver(hg_, result_, n_devices_, cal_mode_, s_pattern_, &motif_result_);
    return result_;
  }

  // k-clique count with k from the constructor; orients the graph first.
  uint64_t RunCF(const gpm_sched* sched = nullptr) {
    if (!hg_.is_oriented()) hg_.orientation();
    CFSolver(hg_, motif_k_, result_, n_devices_, cal_mode_, sched);
    return result_;
  }

  // Motif4: induced {3-star, 4-path, tailed-triangle, 4-cycle, diamond,
  // 4-clique} of the last RunG2Miner / RunPatternEnum.
  const std::vector<uint64_t>& motif_counts() const { return motif_result_; }

 private:
  Graph<VID, VLABEL>& hg_;
  uint64_t result_ = 0;
  int n_devices_ = 1;
  algorithms algo_;
  modes cal_mode_;
  int motif_k_ = 0;
  std::vector<uint64_t> motif_result_;
  Pattern s_pattern_;
};

}  // namespace project_AntiRF
