// Labeled subgraph matching, many queries per launch (gpm_match_batch).
//
// Counts, for every query, the injective label- and edge-preserving maps
// into the data graph (non-induced, automorphisms counted): the number the
// reference's CPU CECI matcher prints as "#Embeddings"
// (sub_ceci/myceci.cpp:540-668).  The GPU path is new work; its pieces:
//
//  * QueryingOrder (host).  Plain policy: root = the query vertex minimising
//    |V(label)| / degree (the CECI start rule, myceci.cpp:15-31) and the BFS
//    order from it (myceci.cpp:65-103).  Learned policy: a cost-based order
//    from the data graph's label statistics -- for every start vertex the
//    order grows by the vertex that keeps the expected number of partial
//    embeddings smallest, and the start with the least expected total work
//    wins.  The same estimate sizes the autotuner's chunks.  With the
//    autotuner on, a query's candidate plans (the cost-based one, the BFS
//    one, the cost-based orders from other starts) are run on a
//    hash-selected sample of their tasks (light roots and parts of heavy
//    roots) in one trial launch that counts wave steps, and the plan with
//    the fewest estimated steps is kept.
//  * Candidate filter at every depth: right label, enough degree and, per
//    neighbour label, enough neighbours of that label (the NLF filter of
//    myceci.cpp:120-169), read in O(1) per label from the label-partitioned
//    adjacency.  Without it the search explores far more dead partial
//    embeddings (on the reference's HPRD queries, the largest root needed
//    1.8M serial steps without it and 60K with it).
//  * Enumeration (device): one wave per task (a chunk of root candidates of
//    one query), tasks pulled from one queue by a persistent grid.  The DFS
//    keeps no candidate lists: at depth d it streams the label-l slice of the
//    shortest backward neighbour's list 64 entries at a time, keeps a 64-bit
//    ballot of the entries that pass (degree, injectivity, membership in the
//    other backward neighbours' label slices by binary search) and descends
//    into the set bits one by one; the last depth only counts ballots.  All
//    DFS state is wave-uniform and lives in LDS.
//  * Autotuner (host): the work one root of a query costs is estimated from
//    the data graph's label statistics (expected label-slice lengths and
//    pass rates per depth of the plan); each query's chunk (roots per task)
//    is set so a task costs about the same (capped so the batch keeps two
//    tasks per resident wave), and queries are laid out heaviest first.
//  * With a fixed chunk, the launch (counter reset + enumeration) can be
//    built once as a two-node hipGraph and replayed.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <queue>
#include <vector>

#include "graph_dev.hpp"

using namespace gpm;

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxQ = 16;
constexpr uint32_t kMaxGroup = 3;

struct QPlan {
  uint32_t n;               // query vertices
  uint32_t label[kMaxQ];    // label per depth
  uint32_t qdeg[kMaxQ];     // query degree per depth
  uint32_t bn[kMaxQ];       // backward-neighbour depths (bitmask) per depth
  uint32_t same[kMaxQ];     // earlier depths with the same label (bitmask)
  uint32_t root_off;        // first root of this query in the root array
  uint32_t n_roots;
  uint32_t chunk;           // roots per task
  uint32_t task_off;        // first task of this query
  uint32_t qid;             // index of the query in the caller's array
  uint32_t setv;            // membership probes: kSetBinary or kSetMerge
  // Neighbour-label filter per depth: a data vertex matched at depth d needs
  // at least nlf_count[d][k] neighbours of label nlf_label[d][k].
  uint32_t nlf_n[kMaxQ];
  uint32_t nlf_label[kMaxQ][kMaxQ];
  uint32_t nlf_count[kMaxQ][kMaxQ];
  // Independent tail: depths tail..n-1 have no query edges among them, so
  // once depths 0..tail-1 are matched their candidate sets are fixed and the
  // injective assignments are counted, not enumerated.  The tail is split
  // into groups of equal label (at most kMaxGroup depths each; different
  // labels cannot collide).
  uint32_t tail;
  uint32_t n_groups;
  uint32_t grp_size[kMaxQ];
  uint32_t grp_dep[kMaxQ][3];
};

struct DevGraph {
  const uint32_t* off;
  const uint32_t* ladj;
  const uint32_t* loff;
  uint32_t L;
  // Packed degree + neighbour-label counts (label alphabets <= kSigLabels),
  // else nullptr.
  const unsigned long long* sig;
};

// Degree and neighbour-label filter of a candidate x for depth d: one load of
// x's packed signature when the graph has one, else the degree and the
// label-slice bounds of every needed label.
__device__ __forceinline__ bool nlf_ok(const DevGraph& g, const QPlan& p, uint32_t d, uint32_t x);

__device__ __forceinline__ void label_slice(const DevGraph& g, uint32_t w, uint32_t l,
                                            uint32_t* lo, uint32_t* len) {
  const uint32_t* p = g.loff + (size_t)w * (g.L + 1) + l;
  const uint32_t a = p[0], b = p[1];
  *lo = a;
  *len = b - a;
}

// Root candidates: qplan p, scan vertices of label p.label[0].
__global__ __launch_bounds__(kBlock) void roots_kernel(DevGraph g, const uint32_t* __restrict__ lverts,
                                                       const uint32_t* __restrict__ lvoff,
                                                       const QPlan* __restrict__ plans,
                                                       uint32_t* __restrict__ roots,
                                                       uint32_t* __restrict__ rlen,
                                                       unsigned int* __restrict__ counts) {
  const QPlan& p = plans[blockIdx.y];
  const uint32_t l0 = p.label[0];
  const uint32_t b = lvoff[l0], e = lvoff[l0 + 1];
  for (uint32_t i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e;
       i += gridDim.x * blockDim.x) {
    const uint32_t v = lverts[i];
    bool ok = g.off[v + 1] - g.off[v] >= p.qdeg[0];
    for (uint32_t k = 0; ok && k < p.nlf_n[0]; ++k) {
      uint32_t lo, len;
      label_slice(g, v, p.nlf_label[0][k], &lo, &len);
      ok = len >= p.nlf_count[0][k];
    }
    if (ok) {
      const uint32_t at = atomicAdd(&counts[blockIdx.y], 1u);
      roots[p.root_off + at] = v;
      // Depth-1 candidates stream v's label-l1 slice (its only backward
      // neighbour is the root): its length sizes the split of heavy roots.
      uint32_t lo = 0, len = 0;
      if (p.n >= 2) label_slice(g, v, p.label[1], &lo, &len);
      rlen[p.root_off + at] = len;
    }
  }
}

// Features of the root candidates for the learned root-cost model, one wave
// per root (plan = blockIdx.y): deg(v), the length of v's depth-1 label
// slice, the degrees summed over that slice (the 2-hop entries a DFS below v
// can reach) and the query size.  Rows of absent roots stay as they are.
__global__ __launch_bounds__(kBlock) void root_features_kernel(
    DevGraph g, const QPlan* __restrict__ plans, const uint32_t* __restrict__ roots,
    const unsigned int* __restrict__ rootcnt, float* __restrict__ feat) {
  const QPlan& p = plans[blockIdx.y];
  const uint32_t nr = rootcnt[blockIdx.y];
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  for (uint32_t r = wave; r < nr; r += nwaves) {
    const uint32_t v = roots[p.root_off + r];
    uint32_t lo = 0, len = 0;
    if (p.n >= 2) label_slice(g, v, p.label[1], &lo, &len);
    uint64_t hop = 0;
    for (uint32_t i = lane; i < len; i += kWave) {
      const uint32_t u = g.ladj[lo + i];
      hop += g.off[u + 1] - g.off[u];
    }
    hop = wave_sum_u64(hop);
    if (lane == 0) {
      float* f = feat + 4 * (size_t)(p.root_off + r);
      f[0] = (float)(g.off[v + 1] - g.off[v]);
      f[1] = (float)len;
      f[2] = (float)hop;
      f[3] = (float)p.n;
    }
  }
}

struct WaveState {
  uint32_t emb[kMaxQ];
  uint32_t cur[kMaxQ];   // next slice position to load
  uint32_t base[kMaxQ];  // slice position of the current window
  uint32_t sb[kMaxQ];    // slice start (in ladj)
  uint32_t sl[kMaxQ];    // slice length
  uint32_t anc[kMaxQ];   // anchor depth (its slice is the one streamed)
  unsigned long long mask[kMaxQ];
  // Slices of depth d's other backward neighbours (the ones an entry of the
  // anchor slice is probed against), shortest first: fixed while depth d is
  // streamed, so they are looked up once per DFS node, not once per window.
  uint32_t np[kMaxQ];
  uint32_t plo[kMaxQ][kMaxQ - 1];
  uint32_t plen[kMaxQ][kMaxQ - 1];
  // MERGE probes: per probed slice, a position before which every entry is
  // below the rest of the anchor stream (advanced window by window).
  uint32_t pcur[kMaxQ][kMaxQ - 1];
  // Range cut of the depth-d streams of the current task (make_cut).
  uint32_t cut[kMaxQ];
};

// Sets up depth d: picks the backward neighbour with the shortest label slice.
__device__ __forceinline__ void init_depth(const DevGraph& g, const QPlan& p, WaveState& s,
                                           uint32_t d) {
  uint32_t bn = p.bn[d], best = 0xFFFFFFFFu, blo = 0, bdep = 0;
  while (bn) {
    const uint32_t b = __builtin_ctz(bn);
    bn &= bn - 1;
    uint32_t lo, len;
    label_slice(g, s.emb[b], p.label[d], &lo, &len);
    if (len < best) {
      best = len;
      blo = lo;
      bdep = b;
    }
  }
  s.sb[d] = blo;
  s.sl[d] = best;
  s.anc[d] = bdep;
  s.cur[d] = 0;
  s.mask[d] = 0ull;
  uint32_t np = 0;
  for (uint32_t bm = p.bn[d] & ~(1u << bdep); bm; bm &= bm - 1) {
    uint32_t lo, len;
    label_slice(g, s.emb[__builtin_ctz(bm)], p.label[d], &lo, &len);
    uint32_t at = np++;  // insertion sort by length
    while (at > 0 && s.plen[d][at - 1] > len) {
      s.plen[d][at] = s.plen[d][at - 1];
      s.plo[d][at] = s.plo[d][at - 1];
      --at;
    }
    s.plen[d][at] = len;
    s.plo[d][at] = lo;
  }
  for (uint32_t k = 0; k < np; ++k) s.pcur[d][k] = 0u;
  s.np[d] = np;
}

__device__ __forceinline__ bool nlf_ok(const DevGraph& g, const QPlan& p, uint32_t d, uint32_t x) {
  if (g.sig) {
    const unsigned long long sg = g.sig[x];
    if ((uint32_t)(sg >> 32) < p.qdeg[d]) return false;
    for (uint32_t k = 0; k < p.nlf_n[d]; ++k)
      if (((uint32_t)(sg >> (4 * p.nlf_label[d][k])) & 15u) < p.nlf_count[d][k]) return false;
    return true;
  }
  if (g.off[x + 1] - g.off[x] < p.qdeg[d]) return false;
  for (uint32_t k = 0; k < p.nlf_n[d]; ++k) {
    uint32_t lo, len;
    label_slice(g, x, p.nlf_label[d][k], &lo, &len);
    if (len < p.nlf_count[d][k]) return false;
  }
  return true;
}

// Membership probes of a window.  BINARY: every lane binary-searches its
// entry in the probed slice (log2 |P| dependent loads per lane, 64 distinct
// addresses per step).  MERGE: the wave walks the probed slice once per
// window in 64-entry chunks (one coalesced load each) and every lane finds
// its entry among the chunk's values held by the other lanes.
constexpr uint32_t kSetBinary = 0, kSetMerge = 1;
// MERGE walks at most this many chunks per window; entries still unresolved
// then binary-search the rest of the slice (a cap: the choice in probe()
// keeps the expected walk shorter).
constexpr uint32_t kMergeChunks = 4;

// First index i in [lo, hi) with a[i] >= key, else hi; every argument is
// the same in all lanes, so the loads are wave-uniform.
__device__ __forceinline__ uint32_t lower_index_uniform(const uint32_t* __restrict__ a, uint32_t lo,
                                                        uint32_t hi, uint32_t key) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (uniform(a[mid]) < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// First index i in [c, len) with a[i] >= key (len if none), searched from
// c by doubling steps: a walk that resumes near its last position costs one
// or two loads.  Wave-uniform arguments.
__device__ __forceinline__ uint32_t lower_index_gallop(const uint32_t* __restrict__ a, uint32_t c,
                                                       uint32_t len, uint32_t key) {
  if (c >= len || uniform(a[c]) >= key) return c;
  uint32_t lo = c + 1, step = 1, hi;
  while (true) {
    hi = c + step;
    if (hi >= len) {
      hi = len;
      break;
    }
    if (uniform(a[hi]) >= key) break;
    lo = hi + 1;
    step <<= 1;
  }
  return lower_index_uniform(a, lo, hi, key);
}

// Is this lane's x in the sorted slice a[0, len)?  Asked for the lanes with
// `live` set, whose x must increase with the lane index (a window of a
// sorted slice); the others get false.  Whole-wave call: every lane must
// reach it with the same a, len and cursor.  *cursor (LDS, may be null)
// holds a position before which every entry is below the live x; it is
// advanced past the entries <= the largest live x, for the next window of
// the same stream.  The walk starts at the first entry >= the smallest
// live x; after each chunk the lanes whose x is <= the chunk's last value
// are resolved (x lies in the chunk or in no chunk), and the walk jumps to
// the first entry >= the smallest unresolved x.
__device__ __noinline__ bool member_merge(const uint32_t* __restrict__ a, uint32_t len, uint32_t x, bool live,
                             uint32_t* cursor) {
  const uint32_t lane = lane_id();
  const uint64_t live_mask = wave_ballot(live);
  uint64_t open = live_mask;
  bool found = false;
  if (open == 0ull || len == 0u) return false;
  const uint32_t xmax =
      uniform(__builtin_amdgcn_readlane(x, (int)uniform(63 - __builtin_clzll(live_mask))));
  uint32_t pos = lower_index_gallop(
      a, cursor ? uniform(*cursor) : 0u, len,
      uniform(__builtin_amdgcn_readlane(x, (int)uniform(__builtin_ctzll(open)))));
  uint32_t next = pos;  // cursor for the next window
  for (uint32_t round = 0; pos < len; ++round) {
    if (round == kMergeChunks) {
      if ((open >> lane) & 1ull) found = sorted_contains(a + pos, len - pos, x);
      break;
    }
    const uint32_t i = pos + lane;
    const uint32_t v = i < len ? a[i] : 0xFFFFFFFFu;  // ids are < 2^32 - 1
    const uint32_t vmax = uniform(__builtin_amdgcn_readlane(v, kWave - 1));
    // Largest j with v[j] <= x (0 when none): six steps over the lanes.
    uint32_t j = 0;
#pragma unroll
    for (uint32_t half = 32; half > 0; half >>= 1)
      if ((uint32_t)__shfl((int)v, (int)(j + half), kWave) <= x) j += half;
    const bool hit = (uint32_t)__shfl((int)v, (int)j, kWave) == x;
    const bool mine = ((open >> lane) & 1ull) && x <= vmax;
    if (mine) found = hit;
    next = pos + (uint32_t)__builtin_popcountll(wave_ballot(i < len && v <= xmax));
    open &= ~wave_ballot(mine);
    if (open == 0ull || vmax == 0xFFFFFFFFu) break;
    pos = lower_index_gallop(
        a, pos + kWave, len,
        uniform(__builtin_amdgcn_readlane(x, (int)uniform(__builtin_ctzll(open)))));
    next = pos;
  }
  if (cursor) *cursor = next;
  return found;
}

// One membership probe of a window streamed from an anchor slice of alen
// entries, by the plan's variant.  MERGE pays one gallop from the stream's
// cursor (a full lower-bound search on the first window) and about
// c = len / alen chunk loads per window, against log2(len) + 1 dependent
// loads for every lane's binary search; the first search only pays off
// over many windows, so a probe uses MERGE where the stream has at least
// kMergeMinWindows windows and 2c + 3 <= floor(log2 len), and binary search
// otherwise.  Without a cursor (probes re-derived per window) it always
// binary-searches.  Whole-wave call.
constexpr uint32_t kMergeMinWindows = 4;
template <bool kMerge>
__device__ __forceinline__ bool probe(const DevGraph& g, uint32_t setv, uint32_t lo, uint32_t len,
                                      uint32_t alen, uint32_t x, bool ok, uint32_t* cursor) {
  if (kMerge && setv == kSetMerge && cursor && alen >= kMergeMinWindows * kWave && len > 1u) {
    const uint32_t c = (len + alen - 1) / alen;
    if (2 * c + 3 <= (uint32_t)(31 - __builtin_clz(len)))
      return member_merge(g.ladj + lo, len, x, ok, cursor);
  }
  return ok && sorted_contains(g.ladj + lo, len, x);
}

// Filters of depth d for this lane's slice entry x (live: the lane holds an
// entry of the window): injectivity, degree and NLF, and membership in the
// other backward neighbours' label slices.  Whole-wave call.  BINARY checks
// degree and NLF first (one load, before the searches); MERGE probes first,
// so only the entries in every slice pay the signature load.
template <bool kMerge>
__device__ __forceinline__ bool window_passes(const DevGraph& g, const QPlan& p, WaveState& s,
                                              uint32_t d, uint32_t x, bool live) {
  bool ok = live;
  for (uint32_t same = p.same[d]; ok && same; same &= same - 1)
    if (s.emb[__builtin_ctz(same)] == x) ok = false;
  const bool merge = kMerge && p.setv == kSetMerge;
  if (!merge && ok) ok = nlf_ok(g, p, d, x);
  const uint32_t np = uniform(s.np[d]), alen = uniform(s.sl[d]);
  for (uint32_t k = 0; k < np; ++k)
    ok = probe<kMerge>(g, p.setv, uniform(s.plo[d][k]), uniform(s.plen[d][k]), alen, x, ok, &s.pcur[d][k]);
  if (merge && ok) ok = nlf_ok(g, p, d, x);
  return ok;
}

// Filters of tail depth d for this lane's entry x (label already right) of
// a window streamed from an anchor slice of alen entries: degree, NLF,
// injectivity against the matched depths below the tail, and membership in
// the lists of d's backward neighbours -- all of them when `full`, all but
// the streamed anchor otherwise.  Whole-wave call; the live lanes' x
// increase with the lane index.
template <bool kMerge>
__device__ __forceinline__ bool window_passes_tail(const DevGraph& g, const QPlan& p,
                                                   WaveState& s, uint32_t d, uint32_t x,
                                                   bool live, bool full, uint32_t alen) {
  bool ok = live;
  for (uint32_t same = p.same[d] & ((1u << p.tail) - 1u); ok && same; same &= same - 1)
    if (s.emb[__builtin_ctz(same)] == x) ok = false;
  const bool merge = kMerge && p.setv == kSetMerge;
  if (!merge && ok) ok = nlf_ok(g, p, d, x);
  if (!full) {
    const uint32_t np = uniform(s.np[d]);
    for (uint32_t k = 0; k < np; ++k)
      ok = probe<kMerge>(g, p.setv, uniform(s.plo[d][k]), uniform(s.plen[d][k]), alen, x, ok,
                 &s.pcur[d][k]);
  } else {
    for (uint32_t bn = p.bn[d]; bn; bn &= bn - 1) {
      uint32_t lo, len;
      label_slice(g, s.emb[__builtin_ctz(bn)], p.label[d], &lo, &len);
      ok = probe<kMerge>(g, p.setv, uniform(lo), uniform(len), alen, x, ok, nullptr);
    }
  }
  if (merge && ok) ok = nlf_ok(g, p, d, x);
  return ok;
}

// A trial task stops once its weighted steps pass this budget: a plan with
// such a task is already far slower than one without, and the trial's
// duration stays bounded whatever the candidates cost.
constexpr uint64_t kTrialBudget = 1ull << 20;

// Kernel modes: count embeddings; plan trial (weighted steps per plan,
// budget-limited); root log (weighted steps of every task, one root per
// task, no budget -- the training data of the learned root-cost model);
// budgeted count (a task that reaches the pass budget of unweighted steps
// stops, adds nothing and is listed, to be counted again cut finer).
constexpr int kModeCount = 0, kModeTrial = 1, kModeLog = 2, kModeBudget = 3;

// Step accounting: trial and log weigh a window or descent by its dependent
// loads; counting modes add 1.
__host__ __device__ constexpr bool weighted_steps(int mode) {
  return mode == kModeTrial || mode == kModeLog;
}

// Budget of a task in a budgeted count: the task stops once its steps reach
// `limit` while the launch's queue holds no task for an idle wave; while
// the queue still has tasks, the limit is raised by `quantum` instead (so
// tasks stop only in the launch's tail, when idle waves could take their
// parts in the next launch).
struct TaskBudget {
  uint64_t limit, quantum;
  const unsigned int* queue;
  uint32_t n_tasks;
};

__device__ __forceinline__ bool over_budget(uint64_t steps, TaskBudget& b) {
  if (steps < b.limit) return false;
  const uint32_t q =
      uniform(__hip_atomic_load(b.queue, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  if (q < b.n_tasks) {
    b.limit = steps + b.quantum;
    return false;
  }
  return true;
}

// Trial cost of one 64-entry window at depth d: the dependent loads one
// lane makes -- the entry, its degree and NLF counts, and per probed
// backward neighbour its slice bounds plus a binary search of that slice
// (all neighbours when `full`, all but the streamed anchor otherwise).
__device__ uint64_t window_cost(const DevGraph& g, const QPlan& p, const WaveState& s, uint32_t d,
                                bool full) {
  uint64_t c = 2 + p.nlf_n[d];
  for (uint32_t bm = full ? p.bn[d] : (p.bn[d] & ~(1u << s.anc[d])); bm; bm &= bm - 1) {
    uint32_t lo, len;
    label_slice(g, s.emb[__builtin_ctz(bm)], p.label[d], &lo, &len);
    c += 2 + (32 - __builtin_clz(len | 1u));
  }
  return c;
}

// Injective assignments of the tail depths given emb[0..tail-1]: the product
// over equal-label groups of the inclusion-exclusion count over the group's
// candidate sets A_1..A_k (k <= 3):
//   k=1: |A1|;  k=2: |A1||A2| - |A12|;
//   k=3: |A1||A2||A3| - |A12||A3| - |A13||A2| - |A23||A1| + 2|A123|.
// |A_B| for a subset B is counted while streaming the candidates of B's
// first member (each x tested against the later members' full filters).
// Arithmetic is modulo 2^64, exact when the true count fits.
template <int kMode, bool kMerge>
__device__ uint64_t count_tail(const DevGraph& g, const QPlan& p, WaveState& s, uint64_t& steps,
                               TaskBudget& budget) {
  const uint32_t lane = lane_id();
  uint64_t prod = 1;
  for (uint32_t gi = 0; gi < p.n_groups; ++gi) {
    const uint32_t k = p.grp_size[gi];
    uint64_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // a[B]: |intersection over members in B|
    for (uint32_t m = 0; m < k; ++m) {
      const uint32_t d = p.grp_dep[gi][m];
      init_depth(g, p, s, d);
      const uint32_t sl = uniform(s.sl[d]), sb = uniform(s.sb[d]);
      uint64_t wcost = 1;
      if (weighted_steps(kMode)) {
        wcost = window_cost(g, p, s, d, false);
        for (uint32_t m2 = m + 1; m2 < k; ++m2) wcost += window_cost(g, p, s, p.grp_dep[gi][m2], true);
      }
      for (uint32_t cur = 0; cur < sl; cur += kWave) {
        if (kMode == kModeTrial && steps >= kTrialBudget) break;
        if (kMode == kModeBudget && over_budget(steps, budget)) break;
        steps += wcost;
        const uint32_t i = cur + lane;
        const bool live = i < sl;
        const uint32_t x = live ? g.ladj[sb + i] : 0u;
        const bool ok = window_passes_tail<kMerge>(g, p, s, d, x, live, false, sl);
        uint32_t mask = ok ? 1u << m : 0u;
        for (uint32_t m2 = m + 1; m2 < k; ++m2)
          if (window_passes_tail<kMerge>(g, p, s, p.grp_dep[gi][m2], x, ok, true, sl)) mask |= 1u << m2;
        // Subsets B with lowest member m and B within mask.
        for (uint32_t B = 1u << m; B < 8u; ++B)
          if ((B & ((1u << m) - 1u)) == 0 && (B & ~mask) == 0 && (B >> m) & 1u) a[B] += 1;
      }
    }
    for (uint32_t B = 1; B < 8; ++B) a[B] = wave_sum_u64(a[B]);
    uint64_t c;
    if (k == 1)
      c = a[1];
    else if (k == 2)
      c = a[1] * a[2] - a[3];
    else
      c = a[1] * a[2] * a[4] - a[3] * a[4] - a[5] * a[2] - a[6] * a[1] + 2 * a[7];
    prod *= c;
    if (prod == 0) break;
  }
  return prod;
}

// Counts the embeddings of plan p rooted at v0 whose depth-1 vertex is one
// of the entries [lo1, hi1) of the depth-1 slice (all: 0, ~0u): a DFS over
// depths 1..tail-1, each leaf of which adds the tail count.  s.cut[d] =
// part << 16 | parts with parts > 1 cuts every depth-d stream of the DFS
// into `parts` equal ranges of its anchor slice, of which only range `part`
// is walked: the parts partition the subtree.
template <int kMode, bool kMerge>
__device__ uint64_t count_root(const DevGraph& g, const QPlan& p, WaveState& s, uint32_t v0,
                               uint32_t lo1, uint32_t hi1, uint64_t& steps,
                               TaskBudget& budget) {
  const uint32_t lane = lane_id();
  const uint32_t n = p.n;
  if (n == 1) return 1;
  s.emb[0] = v0;
  if (p.tail == 1) return count_tail<kMode, kMerge>(g, p, s, steps, budget);
  uint64_t total = 0;
  init_depth(g, p, s, 1);
  s.cur[1] = lo1;
  s.sl[1] = min(s.sl[1], hi1);
  int d = 1;
  while (d >= 1) {
    if (kMode == kModeTrial && steps >= kTrialBudget) break;
    if (kMode == kModeBudget && over_budget(steps, budget)) break;
    const uint32_t dd = (uint32_t)d;
    if (s.mask[dd] == 0ull) {
      const uint32_t cur = s.cur[dd];
      if (cur >= s.sl[dd]) {
        --d;
        continue;
      }
      steps += weighted_steps(kMode) ? window_cost(g, p, s, dd, false) : 1u;
      const uint32_t i = cur + lane;
      const bool live = i < s.sl[dd];
      const bool ok = window_passes<kMerge>(g, p, s, dd, live ? g.ladj[s.sb[dd] + i] : 0u, live);
      s.mask[dd] = wave_ballot(ok);
      s.base[dd] = cur;
      s.cur[dd] = cur + kWave;
      continue;
    }
    const unsigned long long m = s.mask[dd];
    const uint32_t j = (uint32_t)__builtin_ctzll(m);
    s.mask[dd] = m & (m - 1);
    s.emb[dd] = uniform(g.ladj[s.sb[dd] + s.base[dd] + j]);
    if (dd + 1 == p.tail) {
      total += count_tail<kMode, kMerge>(g, p, s, steps, budget);
    } else {
      init_depth(g, p, s, dd + 1);
      const uint32_t cut = uniform(s.cut[dd + 1]);
      if ((cut & 0xFFFFu) > 1u) {
        const uint64_t sl = s.sl[dd + 1];
        const uint32_t part = cut >> 16, parts = cut & 0xFFFFu;
        s.cur[dd + 1] = (uint32_t)(sl * part / parts);
        s.sl[dd + 1] = (uint32_t)(sl * (part + 1) / parts);
      }
      steps += weighted_steps(kMode) ? 2u * __builtin_popcount(p.bn[dd + 1]) : 1u;
      ++d;
    }
  }
  return total;
}

__device__ __forceinline__ uint32_t find_plan(const QPlan* plans, uint32_t nq, uint32_t t) {
  uint32_t lo = 0, hi = nq;  // last plan with task_off <= t
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (plans[mid].task_off <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Depths 2 .. kCutLevels + 1 of a task's DFS can be cut into ranges.
constexpr int kCutLevels = 4;
constexpr uint32_t kNoCut = 1u;  // part 0 of 1: the whole stream

// Budgeted passes of tuned batches: steps (windows + descents) between a
// running task's checks of the queue, parts per stopped task and level, and
// the pass limit.  Passes run by default on short batches only (those the
// refinement deals out whole): on long ones no task stops, and the budgeted
// kernel, at 90 VGPRs (5 waves per SIMD against the count kernel's 6), took
// 1.62 s where one count launch took 1.41 s (four named queries, labeled
// R-MAT scale 22).  On dense 9-10-vertex queries of the config-4 generator a
// quantum of 512 steps measured 51 ms, 4096 steps 105 ms.
constexpr uint64_t kPassBudget = 512;
constexpr uint32_t kPassBudgetOff = 0xFFFFFFFFu;  // gpm_tuning.pass_budget: one launch
constexpr uint32_t kPassSplit = 16;
constexpr int kMaxPasses = 12;
constexpr size_t kMaxPassTasks = size_t(1) << 22;

__host__ __device__ constexpr uint32_t make_cut(uint32_t part, uint32_t parts) {
  return part << 16 | parts;
}

// A part of one heavy root: depth-1 entries [lo, hi) of root `root` (index
// into the plan's root list) of plan `plan`; cut[i] = make_cut(part, parts)
// restricts every depth-(i + 2) stream of the part's DFS to range `part` of
// `parts` (kNoCut: all of it).
struct HeavyTask {
  uint32_t plan, root, lo, hi;
  uint32_t cut[kCutLevels];
};

// Persistent enumeration: every wave pulls tasks from one queue until it is
// drained.  Tasks [0, n_heavy) are parts of heavy roots (heaviest first);
// the rest are chunks of a query's root list, which skip the heavy roots.
// Counted per task: steps = windows + descents.  A plan trial (kTrial)
// weighs each window and descent by the dependent loads it makes, stops a
// task past kTrialBudget (its counts are then partial and unused), and
// adds, per plan, the total to work[3 qid], the largest task's to
// work[3 qid + 1] and the number of stopped tasks to work[3 qid + 2].
// kMerge: the batch has MERGE plans (a launch without them is built without
// the merge walk, which keeps its register count at 6 waves per SIMD).
// kModeBudget: a task whose steps reach its budget (TaskBudget, quantum
// `budget`) is cut short, adds no count and appends its index to
// aborted[1 ..] (aborted[0] counts them; the list has room for every task).
template <int kMode, bool kMerge>
__global__ __launch_bounds__(kBlock) void match_kernel(
    DevGraph g, const QPlan* __restrict__ plans, uint32_t nq, uint32_t n_tasks,
    const uint32_t* __restrict__ roots, const uint32_t* __restrict__ rlen,
    const HeavyTask* __restrict__ heavy, uint32_t n_heavy, uint32_t split_len,
    unsigned int* __restrict__ queue, unsigned long long* __restrict__ counts,
    unsigned long long* __restrict__ work, uint64_t budget, unsigned int* __restrict__ aborted) {
  __shared__ WaveState s_state[kWavesPerBlock];
  WaveState& s = s_state[threadIdx.x / kWave];
  const uint32_t lane = lane_id();
  // Every lane writes the same values (no lane-dependent branch).
  for (int d = 0; d < kMaxQ; ++d) s.cut[d] = kNoCut;
  while (true) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(queue, 1u);
    t = uniform(wave_bcast(t, 0));
    if (t >= n_tasks) break;
    uint64_t c = 0;
    uint32_t qid;
    uint64_t steps = 0;
    TaskBudget tb{budget, budget, queue, n_tasks};
    if (t < n_heavy) {
      const HeavyTask h = heavy[t];
      const QPlan& p = plans[uniform(h.plan)];
      qid = p.qid;
#pragma unroll
      for (int i = 0; i < kCutLevels; ++i) s.cut[i + 2] = uniform(h.cut[i]);
      c = count_root<kMode, kMerge>(g, p, s, uniform(roots[p.root_off + uniform(h.root)]),
                                    uniform(h.lo), uniform(h.hi), steps, tb);
    } else {
      const uint32_t pi = uniform(find_plan(plans, nq, t - n_heavy));
      const QPlan& p = plans[pi];
      qid = p.qid;
      const uint32_t r0 = (t - n_heavy - p.task_off) * p.chunk;
      const uint32_t r1 = min(r0 + p.chunk, p.n_roots);
#pragma unroll
      for (int i = 0; i < kCutLevels; ++i) s.cut[i + 2] = kNoCut;
      for (uint32_t r = r0; r < r1; ++r) {
        if (p.tail > 1 && uniform(rlen[p.root_off + r]) > split_len) continue;  // a heavy task
        c += count_root<kMode, kMerge>(g, p, s, uniform(roots[p.root_off + r]), 0u, 0xFFFFFFFFu,
                                       steps, tb);
        if (kMode == kModeBudget && over_budget(steps, tb)) break;
      }
    }
    // A task that stopped early ends with steps >= its limit (the limit is
    // only raised below the steps taken); one that merely crossed it on its
    // last step is also counted again (exact, only some work repeated).
    // Every lane holds the same step count; readfirstlane makes the test a
    // scalar branch.
    const bool stopped =
        kMode == kModeBudget && __builtin_amdgcn_readfirstlane(steps >= tb.limit ? 1u : 0u) != 0u;
    if (stopped) {
      if (lane == 0) aborted[1 + atomicAdd(aborted, 1u)] = t;
    } else if (lane == 0 && c) {
      atomicAdd(&counts[qid], (unsigned long long)c);
    }
    if (kMode == kModeLog && lane == 0) work[t] = steps;
    if (kMode == kModeTrial && lane == 0 && steps) {
      atomicAdd(&work[3 * qid], (unsigned long long)steps);
      atomicMax(&work[3 * qid + 1], (unsigned long long)steps);
      if (steps >= kTrialBudget) atomicAdd(&work[3 * qid + 2], 1ull);
    }
    // Keeps the task loop one loop with one latch.  Without it the compiler
    // gave some modes (root log, budgeted count) a second latch behind the
    // lane-0 branches above, i.e. a nested loop around the task fetch's
    // broadcast that lane 0 left alone; the other lanes then read task
    // indices from the inactive lane 0 and never drained the queue (the root
    // log's stall).  tests/test_match_isa.py checks every mode's loop shape.
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ void zero_f32_kernel(float* a, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = 0.f;
}

__global__ void zero_kernel(unsigned long long* a, int na, unsigned int* b, int nb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < na) a[i] = 0ull;
  if (i < nb) b[i] = 0u;
}

struct HostQuery {
  uint32_t n;
  std::vector<std::vector<uint32_t>> adj;
  std::vector<uint32_t> label;
};

// Label statistics of the data graph used by the cost estimates: |V(l)| and
// E(a, b) = adjacency entries from label a to label b.
struct LabelStats {
  const gpm_graph* g;
  double nv(uint32_t l) const {
    return std::max(1.0, (double)(g->h_lvoff[l + 1] - g->h_lvoff[l]));
  }
  double epair(uint32_t a, uint32_t b) const {
    if (!g->h_lpair.empty()) return g->h_lpair[(size_t)a * g->n_labels + b];
    return (double)g->m * nv(a) * nv(b) / ((double)g->n * (double)g->n);
  }
  // Expected label-lb-slice length of a vertex of label la.
  double slice(uint32_t la, uint32_t lb) const { return epair(la, lb) / nv(la); }
  // Probability that a given label-lb vertex is adjacent to a label-la one.
  double adj_prob(uint32_t la, uint32_t lb) const {
    return std::min(1.0, epair(la, lb) / (nv(la) * nv(lb)));
  }
};

// One DFS depth under the estimate: the anchor (shortest expected slice) is
// streamed, every entry is probed against the other backward neighbours and
// survives each with the adjacency probability.  Updates nodes (expected DFS
// nodes at this depth per root) and returns the entries streamed+probed.
double estimate_depth(const LabelStats& st, const uint32_t* label_of_depth, uint32_t bn,
                      uint32_t label, double* nodes) {
  double slice = 1e300;
  uint32_t anchor = 0;
  for (uint32_t m = bn; m; m &= m - 1) {
    const uint32_t b = (uint32_t)__builtin_ctz(m);
    const double len = st.slice(label_of_depth[b], label);
    if (len < slice) {
      slice = len;
      anchor = b;
    }
  }
  if (slice >= 1e300) slice = 0.0;
  double pass = 1.0;
  for (uint32_t m = bn & ~(1u << anchor); m; m &= m - 1)
    pass *= st.adj_prob(label_of_depth[(uint32_t)__builtin_ctz(m)], label);
  const double work = *nodes * slice * (double)__builtin_popcount(bn);
  *nodes *= slice * pass;
  return work;
}

// Expected work (label-slice entries streamed, times the membership probes
// each entry takes) of the DFS below one root of plan p.
double estimate_root_cost(const gpm_graph* g, const QPlan& p) {
  const LabelStats st{g};
  double nodes = 1.0, work = 0.0;
  for (uint32_t d = 1; d < p.n; ++d) work += estimate_depth(st, p.label, p.bn[d], p.label[d], &nodes);
  return std::max(1.0, work);
}

// Membership-probe variant of plan p: the caller's, else under the
// autotuner MERGE when some depth with probes streams slices of expected
// length >= kMergeMinSlice (label statistics) and BINARY otherwise.  MERGE
// only pays on long streams (many windows per cursor); on R-MAT with 8
// labels (expected slices of ~4 entries) it measured 15-27 % slower than
// BINARY on every named query and batch, so short expected slices keep
// BINARY.
constexpr double kMergeMinSlice = 256.0;
uint32_t pick_variant(const gpm_graph* g, const QPlan& p, const gpm_tuning& tun) {
  if (tun.set_variant == GPM_MATCH_SETV_BINARY) return kSetBinary;
  if (tun.set_variant == GPM_MATCH_SETV_MERGE) return kSetMerge;
  if (!tun.autotune) return kSetBinary;
  const LabelStats st{g};
  for (uint32_t d = 1; d < p.n; ++d) {
    if (__builtin_popcount(p.bn[d]) < 2) continue;
    for (uint32_t m = p.bn[d]; m; m &= m - 1)
      if (st.slice(p.label[__builtin_ctz(m)], p.label[d]) >= kMergeMinSlice) return kSetMerge;
  }
  return kSetBinary;
}

// Cost-based order from start vertex s0: the order grows greedily by the
// vertex that keeps the expected number of partial embeddings smallest.
// *total = root candidates x expected work per root.  False when the query
// is disconnected (checked by the caller).
bool greedy_order(const gpm_graph* g, const HostQuery& q, uint32_t s0, std::vector<uint32_t>* out,
                  double* total) {
  const uint32_t n = q.n;
  const LabelStats st{g};
  std::vector<uint32_t> ord{s0};
  std::vector<char> in(n, 0);
  in[s0] = 1;
  uint32_t lab_of[kMaxQ];
  std::vector<uint32_t> at(n, 0);
  lab_of[0] = q.label[s0];
  double nodes = 1.0, work = 1.0;
  while (ord.size() < n) {
    int bu = -1;
    double bnodes = 0.0, bwork = 0.0;
    int bconn = -1;
    for (uint32_t u = 0; u < n; ++u) {
      if (in[u]) continue;
      uint32_t bn = 0;
      for (uint32_t w : q.adj[u])
        if (in[w]) bn |= 1u << at[w];
      if (!bn) continue;
      double nd = nodes;
      const double wk = estimate_depth(st, lab_of, bn, q.label[u], &nd);
      const int conn = __builtin_popcount(bn);
      if (bu < 0 || nd < bnodes || (nd == bnodes && conn > bconn)) {
        bu = (int)u;
        bnodes = nd;
        bwork = wk;
        bconn = conn;
      }
    }
    if (bu < 0) return false;
    at[bu] = (uint32_t)ord.size();
    lab_of[ord.size()] = q.label[bu];
    in[bu] = 1;
    ord.push_back((uint32_t)bu);
    nodes = bnodes;
    work += bwork;
  }
  *total = st.nv(q.label[s0]) * work;
  *out = ord;
  return true;
}

// CECI's start rule (least |V(label)| / degree) and the BFS order from it.
std::vector<uint32_t> bfs_order(const gpm_graph* g, const HostQuery& q) {
  const uint32_t n = q.n;
  auto lab_count = [&](uint32_t l) { return (double)(g->h_lvoff[l + 1] - g->h_lvoff[l]); };
  uint32_t start = 0;
  double best = 1e300;
  for (uint32_t u = 0; u < n; ++u) {
    const double deg = std::max<double>(1.0, (double)q.adj[u].size());
    const double score = lab_count(q.label[u]) / deg;
    if (score < best) {
      best = score;
      start = u;
    }
  }
  std::vector<uint32_t> order;
  std::vector<char> placed(n, 0);
  std::queue<uint32_t> bq;
  bq.push(start);
  placed[start] = 1;
  while (!bq.empty()) {
    const uint32_t u = bq.front();
    bq.pop();
    order.push_back(u);
    for (uint32_t w : q.adj[u])
      if (!placed[w]) {
        placed[w] = 1;
        bq.push(w);
      }
  }
  return order;
}

bool labels_known(const gpm_graph* g, const HostQuery& q) {
  for (uint32_t u = 0; u < q.n; ++u)
    if (q.label[u] >= g->n_labels) return false;
  return true;
}

// Turns a matching order into a plan.
void finish_plan(const HostQuery& q, std::vector<uint32_t> order, QPlan* p) {
  const uint32_t n = q.n;
  // Move vertices whose neighbours all come earlier to the end (keeping the
  // root first), so the independent tail -- counted, not enumerated -- is as
  // long as possible.  Every moved vertex keeps all its backward neighbours.
  for (uint32_t i = n - 1; i-- > 1;) {
    const uint32_t u = order[i];
    bool sink = true;
    for (uint32_t j = i + 1; j < n && sink; ++j)
      for (uint32_t w : q.adj[u])
        if (w == order[j]) sink = false;
    if (sink) {
      order.erase(order.begin() + i);
      order.push_back(u);
    }
  }
  std::vector<uint32_t> pos(n);
  for (uint32_t i = 0; i < n; ++i) pos[order[i]] = i;
  std::memset(p, 0, sizeof(QPlan));
  p->n = n;
  for (uint32_t d = 0; d < n; ++d) {
    const uint32_t u = order[d];
    p->label[d] = q.label[u];
    p->qdeg[d] = (uint32_t)q.adj[u].size();
    for (uint32_t w : q.adj[u])
      if (pos[w] < d) p->bn[d] |= 1u << pos[w];
    for (uint32_t j = 0; j < d; ++j)
      if (q.label[order[j]] == q.label[u]) p->same[d] |= 1u << j;
  }
  // Neighbour-label requirements of every depth.
  for (uint32_t d = 0; d < n; ++d) {
    std::vector<std::pair<uint32_t, uint32_t>> need;
    for (uint32_t w : q.adj[order[d]]) {
      bool found = false;
      for (auto& kv : need)
        if (kv.first == q.label[w]) {
          kv.second++;
          found = true;
        }
      if (!found) need.emplace_back(q.label[w], 1u);
    }
    p->nlf_n[d] = (uint32_t)need.size();
    for (size_t i = 0; i < need.size(); ++i) {
      p->nlf_label[d][i] = need[i].first;
      p->nlf_count[d][i] = need[i].second;
    }
  }
  // Independent tail: grow it backwards from the last depth while the next
  // depth has no edge into it and no label group exceeds kMaxGroup.
  uint32_t t = n - 1;
  while (t > 1) {
    bool ok = true;
    uint32_t same_label = 1;
    for (uint32_t d = t; d < n; ++d) {
      if (p->bn[d] & (1u << (t - 1))) ok = false;
      if (p->label[d] == p->label[t - 1]) ++same_label;
    }
    if (!ok || same_label > kMaxGroup) break;
    --t;
  }
  p->tail = n >= 2 ? t : 1;
  p->n_groups = 0;
  for (uint32_t d = p->tail; d < n; ++d) {
    uint32_t gi = 0;
    while (gi < p->n_groups && p->label[p->grp_dep[gi][0]] != p->label[d]) ++gi;
    if (gi == p->n_groups) p->grp_size[p->n_groups++] = 0;
    p->grp_dep[gi][p->grp_size[gi]++] = d;
  }
}

// Builds the plan of one query: BFS order (plain) or the cost-based order
// from the start with the least expected total work (learned).  Returns
// false when it cannot match at all (a label the data graph does not have).
bool make_plan(const gpm_graph* g, const HostQuery& q, int policy, QPlan* p) {
  if (!labels_known(g, q)) return false;
  if (policy != GPM_SCHED_LEARNED) {
    finish_plan(q, bfs_order(g, q), p);
    return true;
  }
  std::vector<uint32_t> order, ord;
  double best = 1e300, total = 0.0;
  for (uint32_t s0 = 0; s0 < q.n; ++s0)
    if (greedy_order(g, q, s0, &ord, &total) && total < best) {
      best = total;
      order = ord;
    }
  if (order.empty()) return false;
  finish_plan(q, order, p);
  return true;
}

constexpr size_t kMaxCandidates = 6;  // plans a trial compares per query

// Distinct plans a trial compares for one query: the BFS plan (the plain
// policy's, kept on near-ties), then the cost-based orders from every start
// vertex in order of estimated work, at most kMaxCandidates.
std::vector<QPlan> plan_candidates(const gpm_graph* g, const HostQuery& q) {
  std::vector<QPlan> out;
  if (!labels_known(g, q)) return out;
  auto add = [&](const std::vector<uint32_t>& order) {
    if (out.size() >= kMaxCandidates) return;
    QPlan p;
    finish_plan(q, order, &p);
    for (const QPlan& o : out)
      if (std::memcmp(&o, &p, sizeof(QPlan)) == 0) return;
    out.push_back(p);
  };
  std::vector<std::pair<double, std::vector<uint32_t>>> starts;
  std::vector<uint32_t> ord;
  double total = 0.0;
  for (uint32_t s0 = 0; s0 < q.n; ++s0)
    if (greedy_order(g, q, s0, &ord, &total)) starts.emplace_back(total, ord);
  std::stable_sort(starts.begin(), starts.end(),
                   [](const auto& x, const auto& y) { return x.first < y.first; });
  if (starts.empty()) return out;
  add(bfs_order(g, q));
  for (size_t i = 0; i < starts.size(); ++i) add(starts[i].second);
  return out;
}

bool connected(const HostQuery& q) {
  if (q.n == 0) return false;
  std::vector<char> seen(q.n, 0);
  std::vector<uint32_t> st{0};
  seen[0] = 1;
  uint32_t cnt = 1;
  while (!st.empty()) {
    const uint32_t u = st.back();
    st.pop_back();
    for (uint32_t w : q.adj[u])
      if (!seen[w]) {
        seen[w] = 1;
        ++cnt;
        st.push_back(w);
      }
  }
  return cnt == q.n;
}

// A prepared batch: plans laid out for the persistent kernel, root
// candidates, counters, and (when asked for) the instantiated hipGraph of
// the counter reset + enumeration launch.  The last batch is kept on the
// graph handle, so calling again with the same queries and tuning replays
// the graph without planning, filtering or tuning again.
struct MatchBatch {
  std::vector<uint32_t> key;
  void* mem = nullptr;
  QPlan* d_plans = nullptr;
  uint32_t* d_roots = nullptr;
  uint32_t* d_rlen = nullptr;
  unsigned long long* d_counts = nullptr;
  unsigned long long* d_work = nullptr;  // wave steps per counter (plan trials, root logs)
  int mode = kModeCount;
  bool merge = false;  // some plan probes with MERGE
  unsigned int* d_rootcnt = nullptr;
  unsigned int* d_queue = nullptr;
  HeavyTask* d_heavy = nullptr;  // own allocation
  unsigned long long* d_log = nullptr;  // own allocation (root logs)
  // Budgeted passes (tuned batches): host copies of the layout, the pass
  // budget and the buffers of the later passes' tasks and abort lists.
  bool passes = false;
  uint64_t pass_budget = 0;
  // Automatic passes (pass_budget 0): the batch's best time in one launch
  // [0] and in passes [1], negative until measured; a replayed batch runs
  // the faster.
  bool pass_auto = false;
  double pass_ms[2] = {-1.0, -1.0};
  std::vector<QPlan> h_plans;         // as laid out (task_off order)
  std::vector<HeavyTask> h_heavy;     // tasks [0, n_heavy) of the first pass
  std::vector<uint32_t> h_rlen;       // depth-1 slice length per root
  std::vector<uint32_t> h_rlen_dev;   // what d_rlen holds (lengths or marks)
  HeavyTask* d_pass = nullptr;        // own allocation
  size_t pass_cap = 0;
  unsigned int* d_abort = nullptr;    // own allocation: count, then indices
  size_t abort_cap = 0;
  uint32_t n_heavy = 0, split_len = 0;
  uint32_t nq = 0, n_tasks = 0;
  int n_queries = 0;
  unsigned grid = 0;
  DevGraph dg{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

void free_batch(void* p) {
  MatchBatch* b = static_cast<MatchBatch*>(p);
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->mem);
  (void)hipFree(b->d_heavy);
  (void)hipFree(b->d_log);
  (void)hipFree(b->d_pass);
  (void)hipFree(b->d_abort);
  delete b;
}

// Full key material of a batch request: tuning fields, then per query its
// size, edge list and labels, as one word sequence.  A cached batch is
// replayed only when its stored sequence equals the new request's.
std::vector<uint32_t> batch_key(const gpm_query* queries, int n_queries, const gpm_tuning& t) {
  std::vector<uint32_t> k;
  k.push_back((uint32_t)n_queries);
  k.push_back((uint32_t)t.autotune);
  k.push_back((uint32_t)t.policy);
  k.push_back(t.chunk);
  k.push_back((uint32_t)t.use_graph);
  k.push_back((uint32_t)t.set_variant);
  k.push_back(t.pass_budget);
  if (t.root_model) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(t.root_model);
    k.insert(k.end(), w, w + kMlpWeights);
  } else {
    k.push_back(0xFFFFFFFFu);
  }
  for (int i = 0; i < n_queries; ++i) {
    k.push_back(queries[i].n);
    k.push_back(queries[i].n_edges);
    k.insert(k.end(), queries[i].edges, queries[i].edges + 2 * (size_t)queries[i].n_edges);
    k.insert(k.end(), queries[i].labels, queries[i].labels + queries[i].n);
  }
  return k;
}

// Enqueues one enumeration launch over n_tasks tasks, the first n_heavy of
// them heavy parts at d_heavy and the rest b's light chunks.  zero_counts:
// clear the counters first (the first launch of a call).  budget > 0: a
// budgeted count whose stopped tasks are listed in d_abort.
int launch_tasks(gpm_graph* g, MatchBatch* b, const HeavyTask* d_heavy, uint32_t n_heavy,
                 uint32_t n_tasks, bool zero_counts, uint64_t budget, unsigned int* d_abort) {
  hipLaunchKernelGGL(zero_kernel, dim3((b->n_queries + 64) / 64 + 1), dim3(64), 0, g->stream,
                     b->d_counts, zero_counts ? b->n_queries : 0, b->d_queue, 4);
  if (budget)
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, g->stream, nullptr, 0, d_abort, 1);
  auto* kernel = budget ? (b->merge ? &match_kernel<kModeBudget, true>
                                    : &match_kernel<kModeBudget, false>)
                 : b->mode == kModeTrial ? &match_kernel<kModeTrial, false>
                 : b->mode == kModeLog   ? &match_kernel<kModeLog, false>
                 : b->merge              ? &match_kernel<kModeCount, true>
                                         : &match_kernel<kModeCount, false>;
  hipLaunchKernelGGL(kernel, dim3(b->grid), dim3(kBlock), 0, g->stream, b->dg, b->d_plans, b->nq,
                     n_tasks, b->d_roots, b->d_rlen, d_heavy, n_heavy, b->split_len, b->d_queue,
                     b->d_counts, b->d_work, budget, d_abort);
  GPM_HIP(hipGetLastError());
  return GPM_OK;
}

// Enqueues counter reset + enumeration of a prepared batch in one launch.
int launch_batch(gpm_graph* g, MatchBatch* b) {
  if (b->exec) {
    GPM_HIP(hipGraphLaunch(b->exec, g->stream));
    return GPM_OK;
  }
  return launch_tasks(g, b, b->d_heavy, b->n_heavy, b->n_tasks, true, 0, nullptr);
}

// Last plan with task_off <= t (the host side of find_plan).
uint32_t host_find_plan(const std::vector<QPlan>& plans, uint32_t t) {
  uint32_t lo = 0, hi = (uint32_t)plans.size();
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (plans[mid].task_off <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// The parts of a stopped task one level finer, into out: a range of several
// depth-1 entries is cut into kPassSplit ranges; a single entry gets the
// first uncut DFS depth d (2 <= d < tail, d <= kCutLevels + 1) cut into
// kPassSplit ranges.  A task with nothing left to cut goes to `whole`, to
// be counted without a budget.
void split_heavy(const MatchBatch& b, const HeavyTask& h, std::vector<HeavyTask>* out,
                 std::vector<HeavyTask>* whole) {
  const QPlan& p = b.h_plans[h.plan];
  if (p.tail <= 1) {
    whole->push_back(h);
    return;
  }
  const uint32_t hi = std::min(h.hi, b.h_rlen[p.root_off + h.root]);
  if (hi > h.lo + 1) {
    const uint32_t step = (hi - h.lo + kPassSplit - 1) / kPassSplit;
    for (uint32_t a = h.lo; a < hi; a += step) {
      HeavyTask t = h;
      t.lo = a;
      t.hi = std::min(hi, a + step);
      out->push_back(t);
    }
    return;
  }
  for (int i = 0; i < kCutLevels; ++i) {
    if ((uint32_t)i + 2 >= p.tail) break;  // depth i + 2 is not a DFS depth
    if ((h.cut[i] & 0xFFFFu) != 1u) continue;
    for (uint32_t k = 0; k < kPassSplit; ++k) {
      HeavyTask t = h;
      t.cut[i] = make_cut(k, kPassSplit);
      out->push_back(t);
    }
    return;
  }
  whole->push_back(h);
}

// A stopped light chunk (index lt among the light tasks): per root it
// counts, kPassSplit ranges of its depth-1 entries (the whole root when its
// plan has no DFS).
void split_light(const MatchBatch& b, uint32_t lt, std::vector<HeavyTask>* out) {
  const uint32_t pi = host_find_plan(b.h_plans, lt);
  const QPlan& p = b.h_plans[pi];
  const uint32_t r0 = (lt - p.task_off) * p.chunk;
  const uint32_t r1 = std::min(r0 + p.chunk, p.n_roots);
  for (uint32_t r = r0; r < r1; ++r) {
    if (p.tail > 1 && b.h_rlen_dev[p.root_off + r] > b.split_len) continue;  // a heavy task
    if (p.tail <= 1) {
      out->push_back({pi, r, 0u, 0xFFFFFFFFu, {kNoCut, kNoCut, kNoCut, kNoCut}});
      continue;
    }
    const uint32_t len = b.h_rlen[p.root_off + r];
    const uint32_t step = std::max(1u, (len + kPassSplit - 1) / kPassSplit);
    for (uint32_t a = 0; a < len; a += step)
      out->push_back({pi, r, a, std::min(len, a + step), {kNoCut, kNoCut, kNoCut, kNoCut}});
  }
}

int ensure_pass_buffers(MatchBatch* b, size_t tasks) {
  if (tasks > b->pass_cap) {
    (void)hipFree(b->d_pass);
    b->d_pass = nullptr;
    b->pass_cap = 0;
    if (hipMalloc(&b->d_pass, tasks * sizeof(HeavyTask)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(GPM_ENOMEM, "matching pass tasks (%zu)", tasks);
    }
    b->pass_cap = tasks;
  }
  if (tasks + 1 > b->abort_cap) {
    (void)hipFree(b->d_abort);
    b->d_abort = nullptr;
    b->abort_cap = 0;
    if (hipMalloc(&b->d_abort, (tasks + 1) * sizeof(unsigned int)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(GPM_ENOMEM, "matching abort list (%zu)", tasks + 1);
    }
    b->abort_cap = tasks + 1;
  }
  return GPM_OK;
}

// Budgeted passes of a tuned batch.  A search tree can hold most of its
// nodes below a few depth-2 entries, and one wave walks such a subtree
// alone, one dependent load chain per node: the launch then lasts as long
// as that subtree.  So every task checks the task queue every
// b->pass_budget steps; once the queue is drained (idle waves), it stops,
// adds nothing and is counted again in the next pass, cut one level finer
// (split_heavy, split_light), until no task stops.  The last pass, and
// tasks with nothing left to cut, run without a budget.  Every task is
// counted exactly once, so the counts are those of one launch.
int run_passes(gpm_graph* g, MatchBatch* b) {
  std::vector<HeavyTask> cur, next, whole;
  const HeavyTask* d_tasks = b->d_heavy;
  uint32_t n_heavy = b->n_heavy, n_tasks = b->n_tasks;
  int st;
  for (int pass = 0; n_tasks; ++pass) {
    const bool last = pass + 1 >= kMaxPasses;
    const uint64_t budget = last ? 0 : b->pass_budget;
    if ((st = ensure_pass_buffers(b, std::max<size_t>(n_tasks, 1))) != GPM_OK) return st;
    if ((st = launch_tasks(g, b, d_tasks, n_heavy, n_tasks, pass == 0, budget, b->d_abort)) !=
        GPM_OK)
      return st;
    if (!budget) break;
    unsigned int n_ab = 0;
    GPM_HIP(hipMemcpyAsync(&n_ab, b->d_abort, sizeof(n_ab), hipMemcpyDeviceToHost, g->stream));
    GPM_HIP(hipStreamSynchronize(g->stream));
    if (n_ab == 0) break;
    if (n_ab > n_tasks) return fail(GPM_EHIP, "abort list of %u tasks past %u", n_ab, n_tasks);
    std::vector<unsigned int> ab(n_ab);
    GPM_HIP(hipMemcpyAsync(ab.data(), b->d_abort + 1, n_ab * sizeof(unsigned int),
                           hipMemcpyDeviceToHost, g->stream));
    GPM_HIP(hipStreamSynchronize(g->stream));
    next.clear();
    // Splitting stops (the stopped tasks are counted whole) where the next
    // pass would hold more than kMaxPassTasks tasks.
    const bool split = (size_t)n_ab * kPassSplit <= kMaxPassTasks;
    for (unsigned int t : ab) {
      if (pass == 0 && t >= n_heavy)
        split_light(*b, t - n_heavy, split ? &next : &whole);
      else if (split)
        split_heavy(*b, pass == 0 ? b->h_heavy[t] : cur[t], &next, &whole);
      else
        whole.push_back(pass == 0 ? b->h_heavy[t] : cur[t]);
    }
    if (next.size() >= (1ull << 31)) return fail(GPM_EINVAL, "too many matching pass tasks");
    cur.swap(next);
    n_heavy = n_tasks = (uint32_t)cur.size();
    if (n_tasks) {
      if ((st = ensure_pass_buffers(b, n_tasks)) != GPM_OK) return st;
      if ((st = h2d(g, b->d_pass, cur.data(), cur.size() * sizeof(HeavyTask))) != GPM_OK)
        return st;
    }
    d_tasks = b->d_pass;
  }
  if (!whole.empty()) {
    // The last pass may still read d_pass, which is about to be refilled
    // (or reallocated).
    GPM_HIP(hipStreamSynchronize(g->stream));
    const uint32_t n = (uint32_t)whole.size();
    if ((st = ensure_pass_buffers(b, n)) != GPM_OK) return st;
    if ((st = h2d(g, b->d_pass, whole.data(), whole.size() * sizeof(HeavyTask))) != GPM_OK)
      return st;
    if ((st = launch_tasks(g, b, b->d_pass, n, n, false, 0, nullptr)) != GPM_OK) return st;
  }
  return GPM_OK;
}

// Builds the hipGraph of launch_batch explicitly: two kernel nodes, the
// enumeration depending on the reset (no stream capture, no memset node).
int build_graph(MatchBatch* b) {
  GPM_HIP(hipGraphCreate(&b->graph, 0));
  int n_queries = b->n_queries, four = 4;
  void* zargs[] = {&b->d_counts, &n_queries, &b->d_queue, &four};
  hipKernelNodeParams zp{};
  zp.func = reinterpret_cast<void*>(&zero_kernel);
  zp.gridDim = dim3((b->n_queries + 64) / 64 + 1);
  zp.blockDim = dim3(64);
  zp.sharedMemBytes = 0;
  zp.kernelParams = zargs;
  zp.extra = nullptr;
  hipGraphNode_t zn = nullptr, mn = nullptr;
  GPM_HIP(hipGraphAddKernelNode(&zn, b->graph, nullptr, 0, &zp));
  uint64_t no_budget = 0;
  unsigned int* no_abort = nullptr;
  void* margs[] = {&b->dg,      &b->d_plans,   &b->nq,      &b->n_tasks,
                   &b->d_roots, &b->d_rlen,    &b->d_heavy, &b->n_heavy,
                   &b->split_len, &b->d_queue, &b->d_counts, &b->d_work,
                   &no_budget,  &no_abort};
  hipKernelNodeParams mp{};
  mp.func = b->merge ? reinterpret_cast<void*>(&match_kernel<kModeCount, true>)
                     : reinterpret_cast<void*>(&match_kernel<kModeCount, false>);
  mp.gridDim = dim3(b->grid);
  mp.blockDim = dim3(kBlock);
  mp.sharedMemBytes = 0;
  mp.kernelParams = margs;
  mp.extra = nullptr;
  GPM_HIP(hipGraphAddKernelNode(&mn, b->graph, &zn, 1, &mp));
  GPM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return GPM_OK;
}

HostQuery host_query(const gpm_query& gq) {
  HostQuery hq;
  hq.n = gq.n;
  hq.adj.assign(gq.n, {});
  hq.label.assign(gq.labels, gq.labels + gq.n);
  for (uint32_t e = 0; e < gq.n_edges; ++e) {
    const uint32_t a = gq.edges[2 * e], c = gq.edges[2 * e + 1];
    hq.adj[a].push_back(c);
    hq.adj[c].push_back(a);
  }
  for (auto& l : hq.adj) {
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  return hq;
}

// Depth-1 entries per part of a heavy root.  Small parts keep one hub's
// subtree from serialising on a few waves: on labeled R-MAT s22 the batch of
// four named queries took 9.2 s with parts of 128 entries, 3.0 s with 32,
// 2.2 s with 8 and 2.0 s with 4; below 4 the extra tasks start to cost on
// light queries.
constexpr uint32_t kSplitLen = 4;
constexpr uint32_t kTrialUnits = 128;  // tasks a plan trial samples per plan (about)
constexpr double kTrialMargin = 0.9;   // a trial pick must be 10 % cheaper to replace

uint32_t root_hash(uint32_t v) {
  v ^= v >> 16;
  v *= 0x7feb352dU;
  v ^= v >> 15;
  v *= 0x846ca68bU;
  return v ^ (v >> 16);
}

// Is the enumeration unit (root v, depth-1 entries from lo) in a trial
// sample of one unit in `stride`?  Whole roots use lo = 0.
bool in_sample(uint32_t v, uint32_t lo, uint32_t stride) {
  return stride <= 1 || root_hash(v ^ root_hash(lo + 0x9e3779b9U)) % stride == 0;
}

// Predicted weighted DFS steps of every root candidate of the plans on the
// device (d_plans with root_off set, roots and per-plan counts already
// filled): features, then the packed MLP at `model` (log1p domain, clamped
// and mapped back).  cost[root_off + r] for r < rootcnt[plan].
int predict_root_costs(gpm_graph* g, const DevGraph& dg, const QPlan* d_plans, uint32_t nq,
                       const uint32_t* d_roots, const unsigned int* d_rootcnt,
                       uint64_t root_total, uint32_t max_roots, const float* model,
                       std::vector<float>* cost) {
  cost->assign(root_total, 0.f);
  if (!root_total || !nq) return GPM_OK;
  float *d_w = nullptr, *d_feat = nullptr, *d_pred = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(d_w);
    (void)hipFree(d_feat);
    (void)hipFree(d_pred);
  };
  if (hipMalloc(&d_w, kMlpWeights * 4) != hipSuccess ||
      hipMalloc(&d_feat, root_total * 16) != hipSuccess ||
      hipMalloc(&d_pred, root_total * 4) != hipSuccess) {
    cleanup();
    return fail(GPM_ENOMEM, "root-cost prediction buffers (%llu roots)",
                (unsigned long long)root_total);
  }
  int st = h2d(g, d_w, model, kMlpWeights * 4);
  if (st == GPM_OK) {
    hipLaunchKernelGGL(zero_f32_kernel, dim3((unsigned)std::min<uint64_t>((4 * root_total + 255) / 256,
                                                                          65535)),
                       dim3(256), 0, g->stream, d_feat, 4 * root_total);
    const unsigned gx = (unsigned)std::max<uint64_t>(
        1, std::min<uint64_t>(((uint64_t)max_roots * kWave + kBlock - 1) / kBlock, 1024));
    hipLaunchKernelGGL(root_features_kernel, dim3(gx, nq), dim3(kBlock), 0, g->stream, dg,
                       d_plans, d_roots, d_rootcnt, d_feat);
    if (hipGetLastError() != hipSuccess) st = fail(GPM_EHIP, "root features launch");
  }
  if (st == GPM_OK) st = mlp_forward(d_feat, (uint32_t)root_total, d_w, d_pred, device_cus(g),
                                     g->stream);
  if (st == GPM_OK &&
      (hipMemcpyAsync(cost->data(), d_pred, root_total * 4, hipMemcpyDeviceToHost, g->stream) !=
           hipSuccess ||
       hipStreamSynchronize(g->stream) != hipSuccess))
    st = fail(GPM_EHIP, "root-cost prediction");
  cleanup();
  if (st != GPM_OK) return st;
  for (float& c : *cost) c = std::expm1(std::min(std::max(c, 0.f), 60.f));
  return GPM_OK;
}

// Lays out a batch of `plans` (plan.qid = index of its counter; n_counters
// counters): device buffers, root candidates, chunks (roots per task) and
// the parts of heavy roots.
//
// With scale (a plan trial), plan i (counter i) runs a sample of its
// enumeration units -- light roots, and the kSplitLen-entry parts of heavy
// roots -- about kTrialUnits of them, picked by a hash of (vertex, part),
// plus the first part of its root with the longest depth-1 slice.  The
// batch then carries d_work, where the kernel adds each plan's weighted
// steps (work[3i]: total, work[3i+1]: largest task, work[3i+2]: tasks
// stopped at the trial budget), and (*scale)[i] is units / sampled units of
// plan i.  Sampling parts rather than roots keeps
// hubs represented without running them whole.
int build_batch(gpm_graph* g, std::vector<QPlan> plans, int n_counters, const gpm_tuning& tun,
                MatchBatch* b, std::vector<double>* scale, const float* model = nullptr,
                bool log_roots = false) {
  const bool trial = scale != nullptr;
  const uint32_t nq0 = (uint32_t)plans.size();
  uint64_t root_total = 0;
  for (auto& p : plans) {
    p.root_off = (uint32_t)root_total;
    root_total += g->h_lvoff[p.label[0] + 1] - g->h_lvoff[p.label[0]];
  }
  if (root_total >= (1ull << 32)) return fail(GPM_EINVAL, "root candidate total too large");
  const uint32_t plan_cap = nq0;
  const int n_slots = trial ? (int)nq0 : n_counters;
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t plan_bytes = align(std::max<size_t>(plan_cap, 1) * sizeof(QPlan));
  const size_t root_bytes = align(std::max<uint64_t>(root_total, 1) * 4);
  const size_t count_bytes = align((size_t)(n_slots + 1) * 8);
  const size_t work_bytes = trial ? align((size_t)(3 * nq0 + 1) * 8) : 0;
  const size_t small_bytes = align((size_t)(nq0 + 1 + 4) * 4);
  const size_t lvoff_bytes = align(g->h_lvoff.size() * 4);
  const size_t total =
      plan_bytes + 2 * root_bytes + count_bytes + work_bytes + small_bytes + lvoff_bytes;
  if (hipMalloc(&b->mem, total) != hipSuccess) return fail(GPM_ENOMEM, "batch buffers (%zu B)", total);
  char* mp = static_cast<char*>(b->mem);
  b->d_plans = reinterpret_cast<QPlan*>(mp);
  mp += plan_bytes;
  b->d_roots = reinterpret_cast<uint32_t*>(mp);
  mp += root_bytes;
  b->d_rlen = reinterpret_cast<uint32_t*>(mp);
  mp += root_bytes;
  b->d_counts = reinterpret_cast<unsigned long long*>(mp);
  mp += count_bytes;
  if (trial) {
    b->d_work = reinterpret_cast<unsigned long long*>(mp);
    b->mode = kModeTrial;
    mp += work_bytes;
  }
  b->d_rootcnt = reinterpret_cast<unsigned int*>(mp);
  b->d_queue = b->d_rootcnt + nq0 + 1;
  mp += small_bytes;
  uint32_t* d_lvoff = reinterpret_cast<uint32_t*>(mp);
  b->n_queries = n_slots;
  b->nq = nq0;
  b->dg = DevGraph{g->off, g->ladj, g->loff, g->n_labels, g->sig};
  b->grid = (unsigned)device_cus(g) * 8;
  if (nq0 == 0) return GPM_OK;
  int st;
  if ((st = h2d(g, d_lvoff, g->h_lvoff.data(), g->h_lvoff.size() * 4)) != GPM_OK) return st;
  if ((st = h2d(g, b->d_plans, plans.data(), (size_t)nq0 * sizeof(QPlan))) != GPM_OK) return st;
  hipLaunchKernelGGL(zero_kernel, dim3((nq0 + 64) / 64 + 1), dim3(64), 0, g->stream, b->d_counts,
                     0, b->d_rootcnt, (int)nq0);
  uint32_t maxlab = 0;
  for (auto& q : plans)
    maxlab = std::max(maxlab, g->h_lvoff[q.label[0] + 1] - g->h_lvoff[q.label[0]]);
  dim3 rgrid((unsigned)std::max<uint32_t>(1, std::min<uint32_t>((maxlab + kBlock - 1) / kBlock, 256)),
             nq0);
  hipLaunchKernelGGL(roots_kernel, rgrid, dim3(kBlock), 0, g->stream, b->dg, g->lverts, d_lvoff,
                     b->d_plans, b->d_roots, b->d_rlen, b->d_rootcnt);
  GPM_HIP(hipGetLastError());
  std::vector<unsigned int> rootcnt(nq0);
  GPM_HIP(hipMemcpyAsync(rootcnt.data(), b->d_rootcnt, nq0 * 4, hipMemcpyDeviceToHost, g->stream));
  GPM_HIP(hipStreamSynchronize(g->stream));
  for (uint32_t i = 0; i < nq0; ++i) plans[i].n_roots = rootcnt[i];
  std::vector<uint32_t> rlen(root_total);
  if (root_total) {
    GPM_HIP(hipMemcpyAsync(rlen.data(), b->d_rlen, root_total * 4, hipMemcpyDeviceToHost,
                           g->stream));
    GPM_HIP(hipStreamSynchronize(g->stream));
  }
  std::vector<uint32_t> stride(nq0, 1);  // trial: one sampled unit in stride[i]
  std::vector<uint32_t> vmax(nq0, 0);    // trial: root with the longest depth-1 slice
  std::vector<uint32_t> troots;          // trial: the sampled roots (host copy)
  auto sampled = [&](uint32_t i, uint32_t v, uint32_t lo) {
    return in_sample(v, lo, stride[i]) || (v == vmax[i] && lo == 0);
  };
  if (trial) {
    // Root lists come from atomics in arbitrary order; the sample depends
    // only on vertex ids and slice lengths, not on that order.
    std::vector<uint32_t> roots(root_total);
    if (root_total) {
      GPM_HIP(hipMemcpyAsync(roots.data(), b->d_roots, root_total * 4, hipMemcpyDeviceToHost,
                             g->stream));
      GPM_HIP(hipStreamSynchronize(g->stream));
    }
    std::vector<uint32_t>& sroots = troots;
    std::vector<uint32_t> srlen;
    scale->assign(nq0, 0.0);
    for (uint32_t i = 0; i < nq0; ++i) {
      QPlan& p = plans[i];
      p.qid = i;
      const bool split = p.tail > 1;
      auto parts = [&](uint32_t len) {
        return split && len > kSplitLen ? (len + kSplitLen - 1) / kSplitLen : 1u;
      };
      uint64_t units = 0;
      uint32_t lmax = 0;
      for (uint32_t r = 0; r < p.n_roots; ++r) {
        const uint32_t len = rlen[p.root_off + r], v = roots[p.root_off + r];
        units += parts(len);
        if (r == 0 || len > lmax || (len == lmax && v < vmax[i])) {
          lmax = len;
          vmax[i] = v;
        }
      }
      stride[i] = (uint32_t)std::max<uint64_t>(1, (units + kTrialUnits - 1) / kTrialUnits);
      uint64_t picked = 0;
      const uint32_t off0 = (uint32_t)sroots.size();
      for (int pass = 0; pass < 2 && picked == 0 && units; ++pass) {
        if (pass == 1) stride[i] = 1;  // nothing hashed in: take every unit
        for (uint32_t r = 0; r < p.n_roots; ++r) {
          const uint32_t v = roots[p.root_off + r], len = rlen[p.root_off + r];
          uint32_t k = 0;
          for (uint32_t u = 0; u < parts(len); ++u) k += sampled(i, v, u * kSplitLen);
          if (k) {
            sroots.push_back(v);
            srlen.push_back(len);
            picked += k;
          }
        }
      }
      p.root_off = off0;
      p.n_roots = (uint32_t)sroots.size() - off0;
      (*scale)[i] = picked ? (double)units / (double)picked : 0.0;
    }
    if (!sroots.empty()) {  // a subset of the root lists: fits their buffers
      if ((st = h2d(g, b->d_roots, sroots.data(), sroots.size() * 4)) != GPM_OK) return st;
      if ((st = h2d(g, b->d_rlen, srlen.data(), srlen.size() * 4)) != GPM_OK) return st;
    }
    rlen.swap(srlen);
  }
  const uint32_t nq = (uint32_t)plans.size();
  b->nq = nq;
  // Autotuner: estimated work per root of every query.
  std::vector<double> cost_per_root(nq, 1.0);
  if (tun.autotune && !trial)
    for (uint32_t i = 0; i < nq; ++i) cost_per_root[i] = estimate_root_cost(g, plans[i]);
  // Learned root-cost model: predicted weighted steps of every root; they
  // size chunks, order queries and decide which roots are split, and how.
  std::vector<float> rcost;
  const bool learned = model && !trial && !log_roots;
  if (learned) {
    uint32_t max_roots = 0;
    for (const QPlan& q : plans) max_roots = std::max(max_roots, q.n_roots);
    if ((st = predict_root_costs(g, b->dg, b->d_plans, nq, b->d_roots, b->d_rootcnt, root_total,
                                 max_roots, model, &rcost)) != GPM_OK)
      return st;
    for (uint32_t i = 0; i < nq; ++i) {
      double sum = 0.0;
      for (uint32_t r = 0; r < plans[i].n_roots; ++r) sum += rcost[plans[i].root_off + r];
      cost_per_root[i] = std::max(1.0, plans[i].n_roots ? sum / plans[i].n_roots : 1.0);
    }
  }
  // Chunks (roots per task) and layout, heaviest queries first when tuned.
  // A tuned chunk gives a task about kTaskWork estimated entries, but no
  // more roots than keep kTasksPerWave tasks per resident wave in the batch:
  // light roots with large chunks would leave most waves idle.
  const double kTaskWork = 4096.0;  // estimated entries streamed per task
  constexpr uint64_t kTasksPerWave = 2;
  uint64_t batch_roots = 0;
  for (const auto& q : plans) batch_roots += q.n_roots;
  const uint64_t waves = (uint64_t)b->grid * kWavesPerBlock;
  const double chunk_cap =
      (double)std::max<uint64_t>(1, batch_roots / (kTasksPerWave * std::max<uint64_t>(waves, 1)));
  std::vector<uint32_t> perm(nq);
  std::iota(perm.begin(), perm.end(), 0u);
  if (!trial && (tun.autotune || tun.policy == GPM_SCHED_LEARNED))
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t c) {
      return cost_per_root[a] * plans[a].n_roots > cost_per_root[c] * plans[c].n_roots;
    });
  std::vector<QPlan> laid;
  laid.reserve(nq);
  uint32_t t = 0;
  // Learned: a task's target is the batch's predicted work over four tasks
  // per resident wave; light roots are chunked to about that much.
  double target = 0.0;
  if (learned) {
    double total = 0.0;
    for (float c : rcost) total += c;
    target = std::max(1.0, total / (4.0 * (double)std::max<uint64_t>(waves, 1)));
  }
  for (uint32_t i : perm) {
    QPlan q = plans[i];
    q.setv = trial || log_roots ? kSetBinary : pick_variant(g, q, tun);
    b->merge = b->merge || q.setv == kSetMerge;
    if (trial || log_roots)
      q.chunk = 1;
    else if (learned && !tun.chunk)
      q.chunk = (uint32_t)std::min(std::min(256.0, chunk_cap),
                                   std::max(1.0, target / cost_per_root[i]));
    else if (tun.chunk)
      q.chunk = tun.chunk;
    else if (tun.autotune)
      q.chunk = (uint32_t)std::min(std::min(256.0, chunk_cap),
                                   std::max(1.0, kTaskWork / cost_per_root[i]));
    else
      q.chunk = 8;
    q.task_off = t;
    t += q.n_roots ? (q.n_roots + q.chunk - 1) / q.chunk : 0;
    laid.push_back(q);
  }
  // Heavy roots (depth-1 slice longer than kSplitLen) are dealt out as parts
  // of kSplitLen depth-1 entries each, so one hub's subtree is spread over
  // many waves; parts of the longest slices go first.  Queries whose tail
  // starts at depth 1 are counted per root without a DFS and are not split.
  b->split_len = log_roots ? 0xFFFFFFFFu : kSplitLen;
  std::vector<HeavyTask> heavy;
  std::vector<double> hlen;  // sort key: slice length, or predicted part cost
  std::vector<uint32_t> marks;  // learned: UINT32_MAX for a root dealt out as parts
  if (learned) marks.assign(std::max<uint64_t>(root_total, 1), 0u);
  for (uint32_t li = 0; li < nq && !log_roots; ++li) {
    const QPlan& q = laid[li];
    if (q.tail <= 1) continue;
    for (uint32_t r = 0; r < q.n_roots; ++r) {
      const uint32_t len = rlen[q.root_off + r];
      if (learned) {
        // Parts of about `target` predicted work each, at most one per
        // depth-1 entry; roots predicted below the target stay whole.
        const double c = rcost[q.root_off + r];
        if (len <= 1 || c <= target) continue;
        const uint32_t np = (uint32_t)std::min<double>(len, std::ceil(c / target));
        const uint32_t step = (len + np - 1) / np;
        marks[q.root_off + r] = 0xFFFFFFFFu;
        for (uint32_t lo = 0; lo < len; lo += step) {
          heavy.push_back({li, r, lo, std::min(len, lo + step), {kNoCut, kNoCut, kNoCut, kNoCut}});
          hlen.push_back(c * (double)(std::min(len, lo + step) - lo) / (double)len);
        }
        continue;
      }
      if (len <= kSplitLen) continue;
      for (uint32_t lo = 0; lo < len; lo += kSplitLen) {
        if (trial && !sampled(q.qid, troots[q.root_off + r], lo)) continue;
        heavy.push_back(
            {li, r, lo, std::min(len, lo + kSplitLen), {kNoCut, kNoCut, kNoCut, kNoCut}});
        hlen.push_back(len);
      }
    }
  }
  // Short batches (the autotuner, no model): when the tasks so far are
  // fewer than kRefineBelow per resident wave, every plan with a DFS (tail
  // >= 2) is dealt out whole -- every root as tasks of at most kSplitLen
  // depth-1 entries and, where the DFS goes below depth 2, each depth-2
  // stream of a task cut into nparts2 ranges -- so a few dense subtrees
  // (small graphs, dense queries) spread over the chip.  About
  // kRefineTasksPerWave tasks per resident wave in all, more where every
  // root has more than kSplitLen depth-1 entries.  (With 2 tasks per wave as
  // the bar, pyramid-6 on R-MAT scale 16 ran refined under one plan, 2.7 s,
  // and unrefined under another, 7.2 s.)
  constexpr uint64_t kRefineTasksPerWave = 4;
  constexpr uint64_t kRefineBelow = 8;
  const uint64_t have_tasks = (uint64_t)t + heavy.size();
  const bool short_batch = have_tasks < kRefineBelow * waves;
  bool refined = false;
  if (tun.autotune && !learned && !trial && !log_roots && !tun.chunk &&
      have_tasks < kRefineBelow * waves) {
    uint64_t e1 = 0, e1_deep = 0;  // depth-1 entries of the refined roots
    for (uint32_t li = 0; li < nq; ++li)
      if (laid[li].tail >= 2)
        for (uint32_t r = 0; r < laid[li].n_roots; ++r) {
          e1 += rlen[laid[li].root_off + r];
          if (laid[li].tail > 2) e1_deep += rlen[laid[li].root_off + r];
        }
    const uint64_t target_tasks = kRefineTasksPerWave * waves;
    if (e1 > 0) {
      refined = true;
      const uint32_t part1 = (uint32_t)std::min<uint64_t>(
          kSplitLen, std::max<uint64_t>(1, (e1 + target_tasks - 1) / target_tasks));
      // Depth-2 ranges multiply the deep plans' tasks up to the target.
      const uint32_t deep_parts =
          part1 > 1 || e1_deep == 0
              ? 1u
              : (uint32_t)std::min<uint64_t>(
                    16, std::max<uint64_t>(1, (target_tasks - (e1 - e1_deep) + e1_deep - 1) /
                                                  e1_deep));
      marks.assign(std::max<uint64_t>(root_total, 1), 0u);
      heavy.clear();
      hlen.clear();
      for (uint32_t li = 0; li < nq; ++li) {
        const QPlan& q = laid[li];
        if (q.tail < 2) continue;
        const uint32_t nparts2 = q.tail > 2 ? deep_parts : 1u;
        for (uint32_t r = 0; r < q.n_roots; ++r) {
          const uint32_t len = rlen[q.root_off + r];
          marks[q.root_off + r] = 0xFFFFFFFFu;  // no embeddings when len == 0
          for (uint32_t lo = 0; lo < len; lo += part1)
            for (uint32_t k = 0; k < nparts2; ++k) {
              heavy.push_back({li, r, lo, std::min(len, lo + part1),
                               {make_cut(k, nparts2), kNoCut, kNoCut, kNoCut}});
              hlen.push_back(len);
            }
        }
      }
    }
  }
  if (learned || refined) {
    // The kernel skips roots whose rlen exceeds split_len in the light
    // chunks: mark exactly the roots dealt out as parts.
    b->split_len = 0;
    if (root_total && (st = h2d(g, b->d_rlen, marks.data(), root_total * 4)) != GPM_OK) return st;
  }
  {
    std::vector<uint32_t> order(heavy.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t c) { return hlen[a] > hlen[c]; });
    std::vector<HeavyTask> sorted(heavy.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = heavy[order[i]];
    heavy.swap(sorted);
  }
  if (heavy.size() >= (1ull << 31)) return fail(GPM_EINVAL, "too many heavy-root parts");
  b->n_heavy = (uint32_t)heavy.size();
  if (b->n_heavy) {
    if (hipMalloc(&b->d_heavy, heavy.size() * sizeof(HeavyTask)) != hipSuccess)
      return fail(GPM_ENOMEM, "heavy-root tasks");
    if ((st = h2d(g, b->d_heavy, heavy.data(), heavy.size() * sizeof(HeavyTask))) != GPM_OK)
      return st;
  }
  b->n_tasks = b->n_heavy + t;
  if ((st = h2d(g, b->d_plans, laid.data(), (size_t)nq * sizeof(QPlan))) != GPM_OK) return st;
  if (log_roots) {
    // One counter per task (= per root: chunk 1, no parts).
    if (hipMalloc(&b->d_log, std::max<uint64_t>(b->n_tasks, 1) * 8) != hipSuccess)
      return fail(GPM_ENOMEM, "root work log");
    b->d_work = b->d_log;
    b->mode = kModeLog;
  }
  if (tun.autotune && !tun.use_graph && !trial && !log_roots &&
      tun.pass_budget != kPassBudgetOff && (short_batch || tun.pass_budget != 0)) {
    b->passes = true;
    b->pass_auto = tun.pass_budget == 0;
    b->pass_budget = tun.pass_budget ? tun.pass_budget : kPassBudget;
    b->h_plans = laid;
    b->h_heavy = heavy;
    b->h_rlen = rlen;
    if (learned || refined)
      b->h_rlen_dev.assign(marks.begin(), marks.begin() + root_total);
    else
      b->h_rlen_dev = rlen;
  }
  if (tun.use_graph && !trial && !log_roots) return build_graph(b);
  return GPM_OK;
}

// Predicted total work of every plan (sum over its root candidates of the
// model's predicted weighted steps): roots are filtered on the device as for
// a batch, then featurised and scored by the packed MLP at `model`.
int plan_costs_by_model(gpm_graph* g, std::vector<QPlan> plans, const float* model,
                        std::vector<double>* total) {
  const uint32_t nq = (uint32_t)plans.size();
  total->assign(nq, 0.0);
  if (!nq) return GPM_OK;
  uint64_t root_total = 0;
  uint32_t max_class = 0;
  for (QPlan& p : plans) {
    p.root_off = (uint32_t)root_total;
    const uint32_t cls = g->h_lvoff[p.label[0] + 1] - g->h_lvoff[p.label[0]];
    root_total += cls;
    max_class = std::max(max_class, cls);
  }
  if (root_total >= (1ull << 32)) return fail(GPM_EINVAL, "root candidate total too large");
  QPlan* d_plans = nullptr;
  uint32_t *d_roots = nullptr, *d_rlen = nullptr, *d_lvoff = nullptr;
  unsigned int* d_cnt = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(d_plans);
    (void)hipFree(d_roots);
    (void)hipFree(d_rlen);
    (void)hipFree(d_lvoff);
    (void)hipFree(d_cnt);
  };
  if (hipMalloc(&d_plans, nq * sizeof(QPlan)) != hipSuccess ||
      hipMalloc(&d_roots, std::max<uint64_t>(root_total, 1) * 4) != hipSuccess ||
      hipMalloc(&d_rlen, std::max<uint64_t>(root_total, 1) * 4) != hipSuccess ||
      hipMalloc(&d_lvoff, g->h_lvoff.size() * 4) != hipSuccess ||
      hipMalloc(&d_cnt, nq * 4) != hipSuccess) {
    cleanup();
    return fail(GPM_ENOMEM, "plan scoring buffers");
  }
  const DevGraph dg{g->off, g->ladj, g->loff, g->n_labels, g->sig};
  int st = h2d(g, d_plans, plans.data(), nq * sizeof(QPlan));
  if (st == GPM_OK) st = h2d(g, d_lvoff, g->h_lvoff.data(), g->h_lvoff.size() * 4);
  std::vector<unsigned int> cnt(nq, 0);
  std::vector<float> cost;
  if (st == GPM_OK) {
    hipLaunchKernelGGL(zero_kernel, dim3((nq + 64) / 64 + 1), dim3(64), 0, g->stream, nullptr, 0,
                       d_cnt, (int)nq);
    dim3 rgrid((unsigned)std::max<uint32_t>(1, std::min<uint32_t>((max_class + kBlock - 1) / kBlock,
                                                                   256)),
               nq);
    hipLaunchKernelGGL(roots_kernel, rgrid, dim3(kBlock), 0, g->stream, dg, g->lverts, d_lvoff,
                       d_plans, d_roots, d_rlen, d_cnt);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(cnt.data(), d_cnt, nq * 4, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess)
      st = fail(GPM_EHIP, "plan scoring roots");
  }
  if (st == GPM_OK) {
    uint32_t max_roots = 0;
    for (unsigned int c : cnt) max_roots = std::max(max_roots, (uint32_t)c);
    st = predict_root_costs(g, dg, d_plans, nq, d_roots, d_cnt, root_total, max_roots, model,
                            &cost);
  }
  cleanup();
  if (st != GPM_OK) return st;
  for (uint32_t i = 0; i < nq; ++i)
    for (uint32_t r = 0; r < cnt[i]; ++r) (*total)[i] += cost[plans[i].root_off + r];
  return GPM_OK;
}

// Plans every query.  Under the learned policy with the autotuner on, a
// query with more than one candidate plan (plan_candidates: the cost-based
// plan, the BFS plan, the cost-based orders from other starts) gets them all
// tried on a sample of their tasks -- one trial launch for the whole batch,
// counting wave steps -- and keeps the one with the fewest estimated steps.
// The label-statistics estimate assumes uniform adjacency and cannot see
// hubs: on power-law graphs the order it prefers can cost several times
// another one, and which start is best changes with the graph.
int plan_queries(gpm_graph* g, const gpm_query* queries, int n_queries, const gpm_tuning& tun,
                 std::vector<QPlan>* out) {
  std::vector<QPlan> trial_plans;
  struct Tried {
    size_t at, first, count;  // index in *out; candidates in trial_plans
  };
  std::vector<Tried> tried;
  for (int qi = 0; qi < n_queries; ++qi) {
    const HostQuery hq = host_query(queries[qi]);
    QPlan p;
    if (!make_plan(g, hq, tun.policy, &p)) continue;  // a label the graph lacks: count 0
    p.qid = (uint32_t)qi;
    if (tun.policy == GPM_SCHED_LEARNED && (tun.autotune || tun.root_model)) {
      std::vector<QPlan> cand = plan_candidates(g, hq);
      if (cand.size() > 1) {
        tried.push_back({out->size(), trial_plans.size(), cand.size()});
        for (QPlan& c : cand) {
          c.qid = p.qid;
          trial_plans.push_back(c);
        }
      }
    }
    out->push_back(p);
  }
  if (trial_plans.empty()) return GPM_OK;
  if (tun.root_model) {
    // Learned QueryingOrder: the candidate with the least predicted work; a
    // later candidate must be cheaper by more than kTrialMargin to replace an
    // earlier (better estimated) one.
    std::vector<double> total;
    int st = plan_costs_by_model(g, trial_plans, tun.root_model, &total);
    if (st != GPM_OK) return st;
    for (const Tried& tr : tried) {
      size_t best = tr.first;
      for (size_t i = tr.first + 1; i < tr.first + tr.count; ++i)
        if (total[i] < kTrialMargin * total[best]) best = i;
      (*out)[tr.at] = trial_plans[best];
    }
    return GPM_OK;
  }
  MatchBatch* t = new MatchBatch();
  std::vector<double> scale;
  int st = build_batch(g, trial_plans, 0, tun, t, &scale);
  std::vector<unsigned long long> work(3 * (size_t)t->n_queries, 0ull);
  if (st == GPM_OK && t->n_tasks) {
    hipLaunchKernelGGL(zero_kernel, dim3((3 * t->n_queries + 64) / 64 + 1), dim3(64), 0,
                       g->stream, t->d_work, 3 * t->n_queries, nullptr, 0);
    st = launch_batch(g, t);
    if (st == GPM_OK && (hipMemcpyAsync(work.data(), t->d_work, work.size() * 8,
                                        hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
                         hipStreamSynchronize(g->stream) != hipSuccess))
      st = fail(GPM_EHIP, "plan trial");
  }
  // Estimated cost of a plan: its total weighted steps (sampled x scale).
  // Heavy roots are dealt out in parts and short batches are cut further at
  // depth 2, so a plan's largest sampled task is not its critical path.  A
  // task stopped at the budget makes its plan's figure a lower bound, which
  // still rules that plan out when it exceeds another plan's figure; ranking
  // plans with fewer stopped tasks first instead picked, on pyramid-6 at
  // R-MAT scale 16, a plan 2.7x slower than the BFS plan.  A later candidate
  // replaces the current pick only when its estimate is lower by more than
  // the sampling noise (kTrialMargin), so near-ties keep the earlier plan:
  // the BFS plan, then the cost-based ones.
  for (const Tried& tr : tried) {
    size_t best = tr.first;
    double best_est = (double)work[3 * tr.first] * scale[tr.first];
    for (size_t i = tr.first + 1; i < tr.first + tr.count; ++i) {
      const double est = (double)work[3 * i] * scale[i];
      if (est < kTrialMargin * best_est) {
        best_est = est;
        best = i;
      }
    }
    (*out)[tr.at] = trial_plans[best];
  }
  return GPM_OK;
}

// Plans, filters roots, tunes and lays out a batch.
int prepare_batch(gpm_graph* g, const gpm_query* queries, int n_queries, const gpm_tuning& tun,
                  MatchBatch* b) {
  std::vector<QPlan> plans;
  int st = plan_queries(g, queries, n_queries, tun, &plans);
  if (st != GPM_OK) return st;
  return build_batch(g, std::move(plans), n_queries, tun, b, nullptr,
                     tun.policy == GPM_SCHED_LEARNED ? tun.root_model : nullptr);
}

}  // namespace

namespace gpm {
void free_match_cache(gpm_graph* g) {
  free_batch(g->match_cache);
  g->match_cache = nullptr;
}
}  // namespace gpm

namespace {

// Validation shared by the entry points: query shape and connectivity.
int check_query(const gpm_query& gq, int qi) {
  GPM_CHECK_ARG(gq.n >= 1 && gq.n <= (uint32_t)kMaxQ, "query %d: 1 <= n <= 16 (got %u)", qi, gq.n);
  GPM_CHECK_ARG(gq.labels && (gq.n_edges == 0 || gq.edges), "query %d: null arrays", qi);
  HostQuery hq;
  hq.n = gq.n;
  hq.adj.assign(gq.n, {});
  for (uint32_t e = 0; e < gq.n_edges; ++e) {
    const uint32_t a = gq.edges[2 * e], c = gq.edges[2 * e + 1];
    GPM_CHECK_ARG(a < gq.n && c < gq.n && a != c, "query %d: bad edge %u (%u, %u)", qi, e, a, c);
    hq.adj[a].push_back(c);
    hq.adj[c].push_back(a);
  }
  GPM_CHECK_ARG(connected(hq), "query %d is not connected", qi);
  return GPM_OK;
}

// One query's plan (policy's order, no trial) laid out one root per task;
// with `log`, the batch records every task's weighted steps.  *ok = false
// when a query label is absent from the graph (no roots).
int single_query_batch(gpm_graph* g, const gpm_query* query, int policy, bool log, MatchBatch* b,
                       bool* ok) {
  GPM_CHECK_ARG(g && query, "null pointer");
  GPM_CHECK_ARG(g->flags & GPM_GRAPH_SYMMETRIC, "matching needs a SYMMETRIC graph");
  GPM_CHECK_ARG(g->labels && g->ladj, "matching needs a labeled graph");
  GPM_CHECK_ARG(policy == GPM_SCHED_DEGREE || policy == GPM_SCHED_LEARNED, "unknown policy %d",
                policy);
  int st;
  if ((st = check_query(*query, 0)) != GPM_OK) return st;
  QPlan p;
  *ok = make_plan(g, host_query(*query), policy, &p);
  if (!*ok) return GPM_OK;
  p.qid = 0;
  gpm_tuning tun{};
  tun.chunk = 1;
  tun.policy = policy;
  return build_batch(g, {p}, 1, tun, b, nullptr, nullptr, log);
}

}  // namespace

// The root log (gpm_match_root_work, gpm_match_root_features): its stall
// came from the split task loop described in match_kernel.
constexpr bool kRootLogEnabled = true;

extern "C" int gpm_match_root_work(gpm_graph* g, const gpm_query* query, int policy,
                                   uint64_t* work) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    if (!kRootLogEnabled) return fail(GPM_EUNSUPPORTED, "the matching root log is switched off");
    GPM_CHECK_ARG(work, "null pointer");
    MatchBatch b;
    bool ok = false;
    int st = single_query_batch(g, query, policy, true, &b, &ok);
    if (st == GPM_OK) std::fill(work, work + g->n, 0ull);
    if (st == GPM_OK && ok && b.n_tasks) {
      st = launch_batch(g, &b);
      std::vector<unsigned long long> w(b.n_tasks);
      std::vector<uint32_t> roots(b.n_tasks);
      if (st == GPM_OK &&
          (hipMemcpyAsync(w.data(), b.d_work, b.n_tasks * 8, hipMemcpyDeviceToHost, g->stream) !=
               hipSuccess ||
           hipMemcpyAsync(roots.data(), b.d_roots, b.n_tasks * 4, hipMemcpyDeviceToHost,
                          g->stream) != hipSuccess ||
           hipStreamSynchronize(g->stream) != hipSuccess))
        st = fail(GPM_EHIP, "root work log");
      if (st == GPM_OK)
        for (uint32_t t = 0; t < b.n_tasks; ++t) work[roots[t]] = w[t];
    }
    MatchBatch* owned = new MatchBatch(std::move(b));
    free_batch(owned);
    return st;
  });
}

extern "C" int gpm_match_root_features(gpm_graph* g, const gpm_query* query, int policy,
                                       float* feat) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    if (!kRootLogEnabled) return fail(GPM_EUNSUPPORTED, "the matching root log is switched off");
    GPM_CHECK_ARG(feat, "null pointer");
    MatchBatch b;
    bool ok = false;
    int st = single_query_batch(g, query, policy, true, &b, &ok);  // one root per task
    if (st == GPM_OK) std::fill(feat, feat + 4 * (size_t)g->n, 0.f);
    const uint32_t nr = ok ? b.n_tasks : 0;
    if (st == GPM_OK && nr) {
      float* d_feat = nullptr;
      std::vector<float> f(4 * (size_t)nr);
      std::vector<uint32_t> roots(nr);
      if (hipMalloc(&d_feat, (size_t)nr * 16) != hipSuccess) {
        st = fail(GPM_ENOMEM, "root features");
      } else {
        const unsigned gx = (unsigned)std::max<uint64_t>(
            1, std::min<uint64_t>(((uint64_t)nr * kWave + kBlock - 1) / kBlock, 1024));
        hipLaunchKernelGGL(root_features_kernel, dim3(gx, 1), dim3(kBlock), 0, g->stream, b.dg,
                           b.d_plans, b.d_roots, b.d_rootcnt, d_feat);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(f.data(), d_feat, f.size() * 4, hipMemcpyDeviceToHost, g->stream) !=
                hipSuccess ||
            hipMemcpyAsync(roots.data(), b.d_roots, nr * 4, hipMemcpyDeviceToHost, g->stream) !=
                hipSuccess ||
            hipStreamSynchronize(g->stream) != hipSuccess)
          st = fail(GPM_EHIP, "root features");
        (void)hipFree(d_feat);
      }
      if (st == GPM_OK)
        for (uint32_t r = 0; r < nr; ++r)
          for (int k = 0; k < 4; ++k) feat[4 * (size_t)roots[r] + k] = f[4 * (size_t)r + k];
    }
    MatchBatch* owned = new MatchBatch(std::move(b));
    free_batch(owned);
    return st;
  });
}

extern "C" int gpm_match_batch(gpm_graph* g, const gpm_query* queries, int n_queries,
                               const gpm_tuning* tuning, uint64_t* per_query) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g && per_query && (queries || n_queries == 0), "null pointer");
    GPM_CHECK_ARG(n_queries >= 0, "negative query count");
    GPM_CHECK_ARG(g->flags & GPM_GRAPH_SYMMETRIC, "matching needs a SYMMETRIC graph");
    GPM_CHECK_ARG(g->labels && g->ladj, "matching needs a labeled graph");
    gpm_tuning tun{};
    tun.autotune = 1;
    tun.use_graph = 0;
    if (tuning) tun = *tuning;
    GPM_CHECK_ARG(tun.policy == GPM_SCHED_DEGREE || tun.policy == GPM_SCHED_LEARNED,
                  "unknown policy %d", tun.policy);
    GPM_CHECK_ARG(tun.set_variant >= GPM_MATCH_SETV_AUTO && tun.set_variant <= GPM_MATCH_SETV_MERGE,
                  "unknown set variant %d", tun.set_variant);
    // Autotuned batches are not captured into a hipGraph: that combination
    // stopped responding on the device with the earlier probe-based tuner and
    // stays disabled until its cause is known.  Tuned batches run as one
    // persistent launch; graph replay is available with fixed chunks.
    if (tun.autotune && tun.use_graph)
      return fail(GPM_EUNSUPPORTED, "autotune with use_graph is not supported; "
                                    "set autotune = 0 (fixed chunk) or use_graph = 0");
    for (int qi = 0; qi < n_queries; ++qi) {
      const gpm_query& gq = queries[qi];
      GPM_CHECK_ARG(gq.n >= 1 && gq.n <= (uint32_t)kMaxQ, "query %d: 1 <= n <= 16 (got %u)", qi,
                    gq.n);
      GPM_CHECK_ARG(gq.labels && (gq.n_edges == 0 || gq.edges), "query %d: null arrays", qi);
      HostQuery hq;
      hq.n = gq.n;
      hq.adj.assign(gq.n, {});
      for (uint32_t e = 0; e < gq.n_edges; ++e) {
        const uint32_t a = gq.edges[2 * e], c = gq.edges[2 * e + 1];
        GPM_CHECK_ARG(a < gq.n && c < gq.n && a != c, "query %d: bad edge %u (%u, %u)", qi, e, a, c);
        hq.adj[a].push_back(c);
        hq.adj[c].push_back(a);
      }
      GPM_CHECK_ARG(connected(hq), "query %d is not connected", qi);
    }
    for (int i = 0; i < n_queries; ++i) per_query[i] = 0;
    if (n_queries == 0) return GPM_OK;
    std::vector<uint32_t> key = batch_key(queries, n_queries, tun);
    MatchBatch* b = static_cast<MatchBatch*>(g->match_cache);
    if (!b || b->key != key) {
      free_match_cache(g);
      b = new MatchBatch();
      int st = prepare_batch(g, queries, n_queries, tun, b);
      if (st != GPM_OK) {
        free_batch(b);
        return st;
      }
      b->key.swap(key);
      g->match_cache = b;
    }
    if (b->nq == 0) return GPM_OK;
    int st;
    // Automatic passes: the first call of a batch runs passes, the second
    // one launch, later calls the faster so far (a short batch of light
    // queries, such as the 200 HPRD queries, loses more to the passes'
    // launches and host round trips than its tail costs: 6.7 against
    // 2.3 ms).
    bool passes = b->passes;
    if (b->pass_auto)
      passes = b->pass_ms[1] < 0.0 || (b->pass_ms[0] >= 0.0 && b->pass_ms[1] <= b->pass_ms[0]);
    if ((st = timing_begin(g)) != GPM_OK) return st;
    if ((st = passes ? run_passes(g, b) : launch_batch(g, b)) != GPM_OK) return st;
    timing_mark(g, "match_batch");
    if ((st = timing_end(g)) != GPM_OK) return st;
    if (b->pass_auto) {
      double& best = b->pass_ms[passes ? 1 : 0];
      best = best < 0.0 ? g->last_ms : std::min(best, g->last_ms);
    }
    std::vector<unsigned long long> cnt(n_queries);
    GPM_HIP(hipMemcpyAsync(cnt.data(), b->d_counts, (size_t)n_queries * 8, hipMemcpyDeviceToHost,
                           g->stream));
    GPM_HIP(hipStreamSynchronize(g->stream));
    for (int i = 0; i < n_queries; ++i) per_query[i] = cnt[i];
    return GPM_OK;
  });
}
