// Root tier construction and caching (see tiers.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "graph_dev.hpp"

namespace gpm {

namespace {

// Full key material of a tier request: the scalar fields as text, plus
// copies of the cost vector (learned policy) and the explicit root list.
// Compared whole on every lookup, so two requests share tiers only when
// they are identical.
TierKey key_of(const TierRequest& r, uint32_t n) {
  TierKey k;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%u/%u/%u/%u/%u/%u/%d/%u/%a/%u/%u/%u/%d", r.min_deg,
                r.wave_cap, r.block_cap, r.block2_cap, r.root_begin, r.root_end, r.policy, r.chunk,
                (double)r.heavy_cost, r.resident_block, r.resident_block2, r.resident_big,
                r.rank != nullptr);
  k.params = buf;
  if (r.policy == GPM_SCHED_LEARNED && r.cost) k.cost.assign(r.cost, r.cost + n);
  if (r.n_root_list) k.roots.assign(r.root_list, r.root_list + r.n_root_list);
  return k;
}

bool same_key(const TierKey& a, const TierKey& b) {
  return a.params == b.params && a.roots == b.roots && a.cost.size() == b.cost.size() &&
         (a.cost.empty() ||
          std::memcmp(a.cost.data(), b.cost.data(), a.cost.size() * sizeof(float)) == 0);
}

int upload(gpm_graph* g, const std::vector<uint32_t>& v, uint32_t** d) {
  if (v.empty()) return GPM_OK;
  if (hipMalloc(d, v.size() * 4) != hipSuccess) {
    *d = nullptr;
    return fail(GPM_ENOMEM, "tier list");
  }
  return h2d(g, *d, v.data(), v.size() * 4);
}

void release(RootTiers& t) {
  for (uint32_t* p : {t.d_wave, t.d_block, t.d_block2, t.d_big, t.d_block_parts, t.d_block2_parts,
                      t.d_big_parts})
    (void)hipFree(p);
  (void)hipFree(t.queues);
}

}  // namespace

int get_tiers(gpm_graph* g, const TierRequest& req, const RootTiers** out) {
  TierKey key = key_of(req, g->n);
  for (auto& kv : g->tier_cache)
    if (same_key(kv.first, key)) {
      *out = &kv.second;
      return GPM_OK;
    }
  const uint32_t rb = req.root_begin;
  const uint32_t re = (req.root_end == 0 || req.root_end > g->n) ? g->n : req.root_end;
  if (rb > re) return fail(GPM_EINVAL, "root range [%u, %u) is empty or reversed", rb, re);
  std::vector<uint32_t> candidates;
  if (req.n_root_list) {
    std::vector<char> seen(g->n, 0);
    for (uint32_t i = 0; i < req.n_root_list; ++i) {
      const uint32_t v = req.root_list[i];
      if (v >= g->n) return fail(GPM_EINVAL, "root_list[%u] = %u >= n", i, v);
      if (seen[v]) return fail(GPM_EINVAL, "root %u listed twice", v);
      seen[v] = 1;
      candidates.push_back(v);
    }
  } else {
    for (uint32_t v = rb; v < re; ++v) candidates.push_back(v);
  }
  std::vector<uint32_t> wave, block, block2, big;
  uint32_t maxb = 0, maxb2 = 0, maxg = 0;
  const bool learned = req.policy == GPM_SCHED_LEARNED && req.cost;
  for (const uint32_t v : candidates) {
    const uint32_t d = g->h_off[v + 1] - g->h_off[v];
    if (d < req.min_deg) continue;
    bool heavy = learned && req.heavy_cost > 0.f && req.cost[v] > req.heavy_cost;
    if (d <= req.wave_cap && !heavy) {
      wave.push_back(v);
    } else if (d <= req.block_cap) {
      block.push_back(v);
      maxb = std::max(maxb, d);
    } else if (d <= req.block2_cap) {
      block2.push_back(v);
      maxb2 = std::max(maxb2, d);
    } else {
      big.push_back(v);
      maxg = std::max(maxg, d);
    }
  }
  std::vector<uint32_t> block_parts, block2_parts, big_parts;
  uint32_t n_split = 0;
  if (!learned) {
    // Plain degree scheduler: each tier in decreasing out-degree, so the
    // largest roots start first (the reference's degree-sorted order).
    auto by_deg = [&](uint32_t a, uint32_t b) {
      const uint32_t da = g->h_off[a + 1] - g->h_off[a], db = g->h_off[b + 1] - g->h_off[b];
      return da > db || (da == db && a < b);
    };
    for (auto* tier : {&wave, &block, &block2, &big}) std::stable_sort(tier->begin(), tier->end(), by_deg);
  }
  if (learned) {
    // Heaviest predicted roots first (longest-processing-time order), so the
    // expensive subtrees start early and the tail of the launch is short.
    auto by_cost = [&](uint32_t a, uint32_t b) {
      return req.cost[a] > req.cost[b] || (req.cost[a] == req.cost[b] && a < b);
    };
    std::stable_sort(wave.begin(), wave.end(), by_cost);
    // Block tiers: a root predicted to cost more than a quarter of one
    // resident block's fair share of its tier is dealt out as up to
    // kMaxParts tasks (each part counts every n-th row of the root's local
    // DAG; the rows themselves are rebuilt per part).  Tasks are then ordered
    // by predicted cost per part.
    auto split = [&](std::vector<uint32_t>& tier, std::vector<uint32_t>& parts,
                     uint32_t resident) {
      double total = 0.0;
      for (const uint32_t v : tier) total += std::max(0.f, req.cost[v]);
      const double target = resident ? total / (4.0 * resident) : 0.0;
      struct Task {
        float cost;
        uint32_t v, part;
      };
      std::vector<Task> tasks;
      tasks.reserve(tier.size());
      bool any = false;
      for (const uint32_t v : tier) {
        const uint32_t d = g->h_off[v + 1] - g->h_off[v];
        uint32_t np = 1;
        const double c = std::max(0.f, req.cost[v]);
        if (target > 0.0 && c > target)
          np = (uint32_t)std::min<double>(kMaxParts, std::ceil(c / target));
        np = std::max(1u, std::min(np, d / kMinRowsPerPart));
        if (np > 1) {
          any = true;
          ++n_split;
        }
        for (uint32_t p = 0; p < np; ++p) tasks.push_back({(float)(c / np), v, p | (np << 16)});
      }
      std::stable_sort(tasks.begin(), tasks.end(), [](const Task& a, const Task& b) {
        return a.cost > b.cost || (a.cost == b.cost && (a.v < b.v || (a.v == b.v && a.part < b.part)));
      });
      tier.clear();
      for (const Task& t : tasks) {
        tier.push_back(t.v);
        if (any) parts.push_back(t.part);
      }
    };
    split(block, block_parts, req.resident_block);
    split(block2, block2_parts, req.resident_block2);
    split(big, big_parts, req.resident_big);
  }
  RootTiers t;
  t.n_wave = (uint32_t)wave.size();
  t.n_block = (uint32_t)block.size();
  t.n_big = (uint32_t)big.size();
  t.max_block_deg = maxb;
  t.n_block2 = (uint32_t)block2.size();
  t.max_block2_deg = maxb2;
  t.max_big_deg = maxg;
  // One root per chunk: on R-MAT s20 the 5-clique wave tier took 1.20 ms
  // interleaved by single roots, 1.47 ms by 4, 1.91 ms by 8.
  t.wave_chunk = req.chunk ? req.chunk : 1;
  if (req.rank)
    for (auto* tier : {&wave, &block, &block2, &big})
      for (uint32_t& v : *tier) v = req.rank[v];
  int st;
  if ((st = upload(g, wave, &t.d_wave)) != GPM_OK ||
      (st = upload(g, block, &t.d_block)) != GPM_OK ||
      (st = upload(g, block2, &t.d_block2)) != GPM_OK ||
      (st = upload(g, big, &t.d_big)) != GPM_OK ||
      (st = upload(g, block_parts, &t.d_block_parts)) != GPM_OK ||
      (st = upload(g, block2_parts, &t.d_block2_parts)) != GPM_OK ||
      (st = upload(g, big_parts, &t.d_big_parts)) != GPM_OK) {
    release(t);  // frees whatever the failed upload sequence allocated
    return st;
  }
  t.n_split = n_split;
  if (hipMalloc(&t.queues, 4 * sizeof(unsigned int)) != hipSuccess) {
    t.queues = nullptr;
    release(t);
    return fail(GPM_ENOMEM, "work queues");
  }
  g->tier_cache.emplace_back(std::move(key), t);
  *out = &g->tier_cache.back().second;
  // Keep the cache small: learned requests with changing costs would grow it.
  if (g->tier_cache.size() > 16) {
    release(g->tier_cache.front().second);
    g->tier_cache.erase(g->tier_cache.begin());
    *out = &g->tier_cache.back().second;
  }
  return GPM_OK;
}

int get_tc_tiers(gpm_graph* g, uint32_t wave_cap, const RootTiers** out) {
  TierRequest r;
  r.min_deg = 2;
  r.wave_cap = wave_cap;
  r.block_cap = 0xFFFFFFFFu;
  return get_tiers(g, r, out);
}

void free_tiers(gpm_graph* g) {
  for (auto& kv : g->tier_cache) release(kv.second);
  g->tier_cache.clear();
}

}  // namespace gpm
