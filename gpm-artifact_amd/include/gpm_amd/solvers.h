// Free solver functions of the reference API, each a call into the C ABI:
//   TCSolver       src/TCSolver.cuh:10-121
//   CFSolver       src/CFSolver.cuh:99-157
//   G2MinerSolver  src/G2MinerSolver.cuh:18-175
//   PatternSolver  src/PatternSolver.cuh:17-264
// Same signatures; n_dev is accepted and ignored, as in the reference
// (device 0 is used).
#pragma once

#include <vector>

#include "graph.h"
#include "sm_pattern.h"

namespace project_AntiRF {

template <typename VID, typename VLABEL>
void TCSolver(Graph<VID, VLABEL>& hg, uint64_t& result, int n_dev, modes cal_m) {
  (void)n_dev;
  if (cal_m != v_centric && cal_m != e_centric) fatal("Wrong Calculation Mode.");
  if (!hg.is_oriented()) hg.orientation(false);
  uint64_t out = 0;
  check_status(gpm_tc(hg.DeviceObject(), cal_m == v_centric ? GPM_TC_VERTEX : GPM_TC_EDGE, &out),
               "TCSolver");
  result = out;
}

// The reference asserts k > 3 and supports k = 4, 5 (CFSolver.cuh:102,
// 297-305); the engine also accepts k = 3.  The graph must be oriented
// first, as in the reference.
template <typename VID, typename VLABEL>
void CFSolver(Graph<VID, VLABEL>& hg, int k, uint64_t& result, int n_dev, modes cal_m,
              const gpm_sched* sched = nullptr) {
  (void)n_dev;
  if (k < 3 || k > 5) fatal("Not supported right now");
  if (cal_m != e_centric && cal_m != v_centric) fatal("Wrong Calculation Mode.");
  if (!hg.is_oriented()) fatal("CFSolver needs an oriented graph (call orientation() first)");
  uint64_t out = 0;
  check_status(gpm_clique(hg.DeviceObject(), k, sched, &out), "CFSolver");
  result = out;
}

template <typename VID, typename VLABEL>
void G2MinerSolver(Graph<VID, VLABEL>& hg, uint64_t& result, int n_dev, modes cal_m, Pattern p,
                   std::vector<uint64_t>* motif = nullptr) {
  (void)n_dev;
  (void)cal_m;
  if (!p.id()) fatal(p.get_name() + " not support!");
  uint64_t out[6] = {0, 0, 0, 0, 0, 0}, scalar = 0;
  check_status(gpm_pattern(hg.DeviceObject(), p.id(), GPM_FORM_G2MINER, out, &scalar),
               "G2MinerSolver");
  result = scalar;
  if (motif) motif->assign(out, out + 6);
}

template <typename VID, typename VLABEL>
void PatternSolver(Graph<VID, VLABEL>& hg, uint64_t& result, int n_dev, modes cal_m, Pattern p,
                   std::vector<uint64_t>* motif = nullptr) {
  (void)n_dev;
  (void)cal_m;
  if (!p.id()) fatal(p.get_name() + " not support!");
  uint64_t out[6] = {0, 0, 0, 0, 0, 0}, scalar = 0;
  check_status(gpm_pattern(hg.DeviceObject(), p.id(), GPM_FORM_PATTERN, out, &scalar),
               "PatternSolver");
  result = scalar;
  if (motif) motif->assign(out, out + 6);
}

}  // namespace project_AntiRF
