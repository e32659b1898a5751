// Device graph handle of libgpm.  Holds the CSR resident in HBM, the
// graph's HIP stream and timing events, a host copy of the offsets (for
// scheduling decisions) and reusable scratch.
#pragma once

#include <hip/hip_runtime.h>

#include <deque>
#include <vector>

#include "device_common.hpp"
#include "tiers.hpp"

constexpr uint32_t kSigLabels = 8;

struct gpm_graph {
  int device = 0;
  uint32_t n = 0;
  uint64_t m = 0;
  uint32_t max_deg = 0;
  uint32_t flags = 0;
  uint32_t n_labels = 0;
  // CSR (uint32 offsets and ids), resident on the device.
  uint32_t* off = nullptr;
  uint32_t* adj = nullptr;
  // COO source of every adjacency entry, built on first use.
  uint32_t* src = nullptr;
  // Labels and the label-partitioned index: ladj holds every list sorted by
  // (label, id); loff[v * (L + 1) + l] is where label l starts in v's list.
  uint32_t* labels = nullptr;
  uint32_t* ladj = nullptr;
  uint32_t* loff = nullptr;
  // Label alphabets of at most kSigLabels labels: sig[v] = degree << 32 |
  // per label l, min(15, neighbours of label l) in bits 4l..4l+3 -- the
  // matcher's degree and neighbour-label filter in one load.
  unsigned long long* sig = nullptr;
  // Vertices grouped by label: lverts[h_lvoff[l] .. h_lvoff[l+1]) have label l.
  uint32_t* lverts = nullptr;
  std::vector<uint32_t> h_lvoff;
  // Adjacency entries per (label of source, label of target), L x L; empty
  // when L > 1024.  Feeds the matcher's cost estimates.
  std::vector<double> h_lpair;
  std::vector<uint32_t> h_off;
  std::vector<uint32_t> h_labels;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  // Per-kernel marks of the last call: an event after each launch.
  static constexpr int kMaxMarks = 8;
  hipEvent_t mark_ev[kMaxMarks] = {};
  const char* mark_name[kMaxMarks] = {};
  double mark_ms[kMaxMarks] = {};
  int n_marks = 0;
  // Scratch: one growable device buffer plus a block of 64 counters.
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  unsigned long long* counters = nullptr;
  // Root tiers built for earlier calls, keyed by the request that made them.
  // A deque keeps pointers to cached tiers valid while entries are added.
  std::deque<std::pair<gpm::TierKey, gpm::RootTiers>> tier_cache;
  // Cached algorithmic-byte totals per k (0 = not computed yet).
  uint64_t plan_bytes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // Symmetric graphs: the graph relabelled by degree rank (position in
  // (degree, id) order) with sorted lists (rk_off/rk_adj), and the
  // degree-oriented DAG, both built on first use by the motif counters.
  uint32_t* rk_off = nullptr;
  uint32_t* rk_adj = nullptr;
  gpm_graph* dag = nullptr;
  // Last prepared matching batch (match.hip), replayed on identical calls.
  void* match_cache = nullptr;
  // Instantiated hipGraph of repeated edge-parallel triangle counts, and the
  // per-block partial sums of the edge kernel (tc.hip).
  hipGraphExec_t tc_exec = nullptr;
  hipGraph_t tc_graph = nullptr;
  unsigned long long* tc_partial = nullptr;
  // ORIENTED graphs: the same DAG relabelled in a topological order
  // (toff/tadj, every edge from a lower to a higher id; h_rank[v] = new id
  // of v), built on first use by the k-clique kernels.
  uint32_t* toff = nullptr;
  uint32_t* tadj = nullptr;
  std::vector<uint32_t> h_rank;
  // SYMMETRIC graphs: block mask of the counting formulation's prefilter,
  // bmask[v] bit b set when v has a neighbour w with (w >> bmask_shift) == b,
  // built on first use.
  unsigned long long* bmask = nullptr;
  uint32_t bmask_shift = 0;
};

namespace gpm {

// Host-to-device copy ordered on the handle's stream and complete on return.
// (A plain hipMemcpy runs on the null stream, which does not order against
// the handle's non-blocking stream, and may return before a pageable
// source's DMA lands: a kernel launched next on the handle's stream could
// read the old contents.)
int h2d(gpm_graph* g, void* dst, const void* src, size_t bytes);
// Grows g->scratch to at least `bytes`; returns status.
int ensure_scratch(gpm_graph* g, size_t bytes);
// Builds g->src (COO sources) if missing.
int ensure_src(gpm_graph* g);
// Zeroes the 64-entry counter block with a kernel (no memset node, so the
// sequence stays valid inside a captured graph).
int zero_counters(gpm_graph* g, int count);
// Starts / stops event timing around the kernels of one mining call.
int timing_begin(gpm_graph* g);
// Records an event after the kernel just launched; `name` must be static.
int timing_mark(gpm_graph* g, const char* name);
int timing_end(gpm_graph* g);
// Number of compute units of the handle's device.
int device_cus(const gpm_graph* g);
// Builds g->toff/tadj/h_rank if missing; GPM_EINVAL if the graph has a cycle.
int ensure_topo(gpm_graph* g);
// Wraps a CSR already resident on the device (d_off: n + 1 entries, d_adj:
// m) into a new handle that owns d_off, d_adj and `stream`; h_off is the host
// copy of the offsets, h_adj the host copy of the adjacency (needed only
// with labels, for the label index).  On failure everything given is freed.
int adopt_device_csr(uint32_t n, uint64_t m, uint32_t* d_off, uint32_t* d_adj,
                     std::vector<uint32_t>&& h_off, const uint32_t* h_adj, const uint32_t* labels,
                     uint32_t n_labels, uint32_t flags, hipStream_t stream, gpm_graph** out);
// Floats of a packed MLP: mean[4], std[4], W1[64x4], b1[64], W2[64x64],
// b2[64], W3[64], b3[1] (the order of gpm_mlp_predict's arguments).
constexpr size_t kMlpWeights = 4 + 4 + 64 * 4 + 64 + 64 * 64 + 64 + 64 + 1;
// log1p-domain outputs of the packed MLP at d_packed for n feature rows
// (4 floats each, raw -- log1p and normalisation happen inside), enqueued on
// `stream` (mlp.hip).
int mlp_forward(const float* d_feat, uint32_t n, const float* d_packed, float* d_out, int cus,
                hipStream_t stream);
// Releases the cached matching batch of g (match.hip).
void free_match_cache(gpm_graph* g);

}  // namespace gpm
