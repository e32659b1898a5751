// Root tiers: the host-side split of root vertices into size classes that
// decides which kernel serves each root (one wave, one block with LDS rows,
// one block with rows in HBM).  The plain degree scheduler orders each tier
// by decreasing out-degree; the learned scheduler orders it by predicted cost
// (heaviest first), may move a root to a wider tier when its predicted cost
// says so, and splits predicted-heavy roots into parts.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

struct gpm_graph;

namespace gpm {

struct RootTiers {
  uint32_t* d_wave = nullptr;   // roots served by one wave
  uint32_t n_wave = 0;
  uint32_t* d_block = nullptr;  // roots served by one block, rows in LDS
  uint32_t n_block = 0;
  uint32_t max_block_deg = 0;
  uint32_t* d_block2 = nullptr; // roots served by one block, wide LDS rows
  uint32_t n_block2 = 0;
  uint32_t max_block2_deg = 0;
  uint32_t* d_big = nullptr;    // roots served by one block, rows in HBM
  uint32_t n_big = 0;
  uint32_t max_big_deg = 0;
  // Learned policy: a heavy root may appear several times in a block tier,
  // once per part; parts[r] = part | (n_parts << 16) says which share of the
  // root's rows task r counts.  nullptr = every task is a whole root.
  uint32_t* d_block_parts = nullptr;
  uint32_t* d_block2_parts = nullptr;
  uint32_t* d_big_parts = nullptr;
  uint32_t n_split = 0;            // roots split into more than one part
  unsigned int* queues = nullptr;  // work-queue counters of the block tiers (4)
  uint32_t wave_chunk = 1;         // roots per chunk of the wave tier's interleave
};

// What a cached RootTiers was built for (see key_of in tiers.hip).
struct TierKey {
  std::string params;
  std::vector<float> cost;
  std::vector<uint32_t> roots;
};

struct TierRequest {
  uint32_t min_deg = 1;       // roots below this out-degree do nothing
  uint32_t wave_cap = 64;     // max out-degree of a wave root
  uint32_t block_cap = 512;   // max out-degree of an LDS block root
  uint32_t block2_cap = 512;  // max out-degree of a wide-LDS block root
  uint32_t root_begin = 0, root_end = 0;
  const uint32_t* root_list = nullptr;  // explicit roots (overrides the range)
  uint32_t n_root_list = 0;
  int policy = 0;             // gpm_sched_policy
  const float* cost = nullptr;  // per-vertex predicted cost (learned)
  uint32_t chunk = 0;         // 0 = automatic
  float heavy_cost = 0.f;     // learned: wave roots above this go to blocks
  // Learned: blocks of each block tier resident at once on the device; a
  // root predicted to cost more than a quarter of a block's fair share of
  // its tier is split into parts (0 = never split).
  uint32_t resident_block = 0, resident_block2 = 0, resident_big = 0;
  // When set, the uploaded root lists hold rank[v] instead of v (kernels
  // that read the topologically relabelled DAG).
  const uint32_t* rank = nullptr;
};

// Most parts a root is split into, and fewest rows a part is given.
constexpr uint32_t kMaxParts = 64;
constexpr uint32_t kMinRowsPerPart = 8;

// Returns cached tiers for (graph, request); builds and uploads on a miss.
int get_tiers(gpm_graph* g, const TierRequest& req, const RootTiers** out);
// Tiers for vertex-centric triangle counting (min_deg 2, wave_cap as given,
// everything above goes to the block kernel).
int get_tc_tiers(gpm_graph* g, uint32_t wave_cap, const RootTiers** out);
void free_tiers(gpm_graph* g);

}  // namespace gpm
