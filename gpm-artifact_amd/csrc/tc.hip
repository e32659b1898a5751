// Triangle counting on a degree-oriented DAG (replaces src/TCSolver.cuh).
//
//  GPM_TC_EDGE   (e_centric, TCSolver.cuh:27-59): sub-wave groups of G lanes
//                per oriented edge (u, v); each lane takes elements of the
//                shorter list and binary-searches the longer one.  G follows
//                the mean out-degree so that low-degree graphs keep all 64
//                lanes busy.
//  GPM_TC_VERTEX (v_centric, TCSolver.cuh:64-113, tc_hi_warp_vertex.cuh):
//                one wave per root u with |N+(u)| <= 256, one 256-thread
//                block per larger root.  N+(u) goes into an LDS hash table;
//                the lists N+(v), v in N+(u), are walked as one flattened
//                range so that every lane reads one entry per step
//                (coalesced, no idle lanes on short lists).  Roots above the
//                LDS capacity probe N+(u) by binary search instead.
//  Count = sum over oriented edges (u, v) of |N+(u) & N+(v)|.
#include <algorithm>
#include <vector>

#include "graph_dev.hpp"
#include "tiers.hpp"

using namespace gpm;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kTcWaveCap = 256;     // roots served by one wave
constexpr uint32_t kTcWaveTable = 512;   // LDS keys per wave
constexpr uint32_t kTcBlockTableMax = 8192;  // 32 KiB of keys per block

template <int G>
__global__ __launch_bounds__(kBlock) void tc_edge_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ src, uint64_t m, unsigned long long* out) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t sub = threadIdx.x & (G - 1);
  const uint64_t ngrp = (uint64_t)gridDim.x * blockDim.x / G;
  uint64_t cnt = 0;
  for (uint64_t e = tid / G; e < m; e += ngrp) {
    const uint32_t u = src[e], v = adj[e];
    uint32_t ab = off[u], na = off[u + 1] - ab;
    uint32_t bb = off[v], nb = off[v + 1] - bb;
    if (na > nb) {
      uint32_t t = ab;
      ab = bb;
      bb = t;
      t = na;
      na = nb;
      nb = t;
    }
    const uint32_t* small = adj + ab;
    const uint32_t* large = adj + bb;
    for (uint32_t i = sub; i < na; i += G) cnt += sorted_contains(large, nb, small[i]);
  }
  cnt = wave_sum_u64(cnt);
  if (lane_id() == 0 && cnt) atomicAdd(out, (unsigned long long)cnt);
}

// Wave per root: roots[] lists vertices with 2 <= |N+| <= kTcWaveCap.
__global__ __launch_bounds__(kBlock) void tc_vertex_wave_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ roots, uint32_t n_roots, uint32_t chunk,
    unsigned int* __restrict__ queue, unsigned long long* out) {
  constexpr int kWavesPerBlock = kBlock / kWave;
  __shared__ uint32_t s_keys[kWavesPerBlock][kTcWaveTable];
  __shared__ uint32_t s_pre[kWavesPerBlock][kWave];
  __shared__ uint32_t s_beg[kWavesPerBlock][kWave];
  const uint32_t lane = lane_id();
  const uint32_t w = threadIdx.x / kWave;
  uint32_t* keys = s_keys[w];
  uint32_t* pre = s_pre[w];
  uint32_t* beg = s_beg[w];
  constexpr uint32_t mask = kTcWaveTable - 1;
  uint64_t cnt = 0;
  while (true) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(queue, chunk);
    base = wave_bcast(base, 0);
    if (base >= n_roots) break;
    const uint32_t stop = min(base + chunk, n_roots);
    for (uint32_t r = base; r < stop; ++r) {
      const uint32_t u = uniform(roots[r]);
      const uint32_t sb = uniform(off[u]), d = uniform(off[u + 1]) - sb;
      for (uint32_t i = lane; i < kTcWaveTable; i += kWave) keys[i] = kEmpty;
      __builtin_amdgcn_wave_barrier();
      for (uint32_t i = lane; i < d; i += kWave) {
        const uint32_t key = adj[sb + i];
        uint32_t h = hslot(key, mask);
        while (atomicCAS(&keys[h], kEmpty, key) != kEmpty) h = (h + 1) & mask;
      }
      __builtin_amdgcn_wave_barrier();
      for (uint32_t c = 0; c < d; c += kWave) {
        uint32_t len = 0, b = 0;
        if (c + lane < d) {
          const uint32_t v = adj[sb + c + lane];
          b = off[v];
          len = off[v + 1] - b;
        }
        const uint32_t incl = wave_incl_scan(len);
        pre[lane] = incl;
        beg[lane] = b;
        const uint32_t total = wave_bcast(incl, kWave - 1);
        __builtin_amdgcn_wave_barrier();
        for (uint32_t t = 0; t < total; t += kWave) {
          const uint32_t f = t + lane;
          if (f < total) {
            const uint32_t j = upper_index(pre, kWave, f);
            const uint32_t excl = j ? pre[j - 1] : 0u;
            const uint32_t x = adj[beg[j] + (f - excl)];
            cnt += lds_hash_has(keys, mask, x) ? 1u : 0u;
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  cnt = wave_sum_u64(cnt);
  if (lane == 0 && cnt) atomicAdd(out, (unsigned long long)cnt);
}

// Block-wide inclusive scan of one value per thread (kBlock threads).
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t* s_wsum) {
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  uint32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) s_wsum[w] = incl;
  __syncthreads();
  uint32_t add = 0;
  for (uint32_t i = 0; i < w; ++i) add += s_wsum[i];
  __syncthreads();
  return incl + add;
}

// Block per root: roots[] lists vertices with |N+| > kTcWaveCap.  LDS table
// of `table` keys (power of two, >= 2|N+|) or, when table == 0, binary
// search of N+(u) in global memory.
__global__ __launch_bounds__(kBlock) void tc_vertex_block_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ roots, uint32_t n_roots, uint32_t table,
    unsigned int* __restrict__ queue, unsigned long long* out) {
  extern __shared__ uint32_t s_dyn[];
  __shared__ uint32_t s_pre[kBlock];
  __shared__ uint32_t s_beg[kBlock];
  __shared__ uint32_t s_wsum[kBlock / kWave];
  __shared__ uint32_t s_root;
  uint32_t* keys = s_dyn;
  const uint32_t mask = table - 1;
  uint64_t cnt = 0;
  while (true) {
    if (threadIdx.x == 0) s_root = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t r = uniform(s_root);
    __syncthreads();
    if (r >= n_roots) break;
    const uint32_t u = uniform(roots[r]);
    const uint32_t sb = uniform(off[u]), d = uniform(off[u + 1]) - sb;
    const uint32_t* S = adj + sb;
    if (table) {
      for (uint32_t i = threadIdx.x; i < table; i += kBlock) keys[i] = kEmpty;
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < d; i += kBlock) {
        const uint32_t key = S[i];
        uint32_t h = hslot(key, mask);
        while (atomicCAS(&keys[h], kEmpty, key) != kEmpty) h = (h + 1) & mask;
      }
      __syncthreads();
    }
    for (uint32_t c = 0; c < d; c += kBlock) {
      uint32_t len = 0, b = 0;
      if (c + threadIdx.x < d) {
        const uint32_t v = S[c + threadIdx.x];
        b = off[v];
        len = off[v + 1] - b;
      }
      const uint32_t incl = block_incl_scan(len, s_wsum);
      s_pre[threadIdx.x] = incl;
      s_beg[threadIdx.x] = b;
      __syncthreads();
      const uint32_t total = uniform(s_pre[kBlock - 1]);
      for (uint32_t f = threadIdx.x; f < total; f += kBlock) {
        const uint32_t j = upper_index(s_pre, kBlock, f);
        const uint32_t excl = j ? s_pre[j - 1] : 0u;
        const uint32_t x = adj[s_beg[j] + (f - excl)];
        if (table)
          cnt += lds_hash_has(keys, mask, x) ? 1u : 0u;
        else
          cnt += sorted_contains(S, d, x) ? 1u : 0u;
      }
      __syncthreads();
    }
  }
  cnt = wave_sum_u64(cnt);
  if (lane_id() == 0 && cnt) atomicAdd(out, (unsigned long long)cnt);
}

__global__ void zero_u32_kernel(unsigned int* p, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

int edge_group_size(const gpm_graph* g) {
  const double avg = g->n ? (double)g->m / g->n : 0.0;
  if (avg <= 6) return 4;
  if (avg <= 12) return 8;
  if (avg <= 24) return 16;
  if (avg <= 48) return 32;
  return 64;
}

// Enqueues one triangle count (counter reset + kernels) on g's stream; the
// count lands in g->counters[0].  mark: record per-kernel timing marks.
int enqueue_tc(gpm_graph* g, int variant, bool mark) {
  int st;
  const int cus = device_cus(g);
  if (variant == GPM_TC_EDGE) {
    if ((st = zero_counters(g, 1)) != GPM_OK) return st;
    const int G = edge_group_size(g);
    const uint64_t groups_wanted = std::max<uint64_t>(g->m, 1);
    const uint64_t blocks = std::min<uint64_t>((groups_wanted * G + kBlock - 1) / kBlock,
                                               (uint64_t)cus * 32);
    dim3 grid((unsigned)std::max<uint64_t>(blocks, 1)), block(kBlock);
    switch (G) {
      case 4: hipLaunchKernelGGL(tc_edge_kernel<4>, grid, block, 0, g->stream, g->off, g->adj, g->src, g->m, g->counters); break;
      case 8: hipLaunchKernelGGL(tc_edge_kernel<8>, grid, block, 0, g->stream, g->off, g->adj, g->src, g->m, g->counters); break;
      case 16: hipLaunchKernelGGL(tc_edge_kernel<16>, grid, block, 0, g->stream, g->off, g->adj, g->src, g->m, g->counters); break;
      case 32: hipLaunchKernelGGL(tc_edge_kernel<32>, grid, block, 0, g->stream, g->off, g->adj, g->src, g->m, g->counters); break;
      default: hipLaunchKernelGGL(tc_edge_kernel<64>, grid, block, 0, g->stream, g->off, g->adj, g->src, g->m, g->counters); break;
    }
    GPM_HIP(hipGetLastError());
    if (mark) timing_mark(g, "tc_edge");
    return GPM_OK;
  }
  const RootTiers* t = nullptr;
  if ((st = get_tc_tiers(g, kTcWaveCap, &t)) != GPM_OK) return st;
  if ((st = zero_counters(g, 1)) != GPM_OK) return st;
  unsigned int* q = t->queues;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, g->stream, q, 4);
  if (t->n_wave) {
    const uint32_t chunk = 8;
    const uint64_t waves = std::min<uint64_t>((t->n_wave + chunk - 1) / chunk, (uint64_t)cus * 32);
    const unsigned blocks = (unsigned)std::max<uint64_t>(1, (waves * kWave + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(tc_vertex_wave_kernel, dim3(blocks), dim3(kBlock), 0, g->stream, g->off,
                       g->adj, t->d_wave, t->n_wave, chunk, q, g->counters);
    GPM_HIP(hipGetLastError());
    if (mark) timing_mark(g, "tc_vertex_wave");
  }
  if (t->n_block) {
    uint32_t table = 0;
    if (2ull * t->max_block_deg <= kTcBlockTableMax) {
      table = 1;
      while (table < 2 * t->max_block_deg) table <<= 1;
    }
    const unsigned blocks = (unsigned)std::min<uint64_t>(t->n_block, (uint64_t)cus * 4);
    hipLaunchKernelGGL(tc_vertex_block_kernel, dim3(blocks), dim3(kBlock),
                       table * sizeof(uint32_t), g->stream, g->off, g->adj, t->d_block,
                       t->n_block, table, q + 1, g->counters);
    GPM_HIP(hipGetLastError());
    if (mark) timing_mark(g, "tc_vertex_block");
  }
  return GPM_OK;
}

int tc_run(gpm_graph* g, int variant, int reps, uint64_t* out) {
  GPM_CHECK_ARG(g && out, "null pointer");
  GPM_CHECK_ARG(g->flags & GPM_GRAPH_ORIENTED, "triangle counting needs an ORIENTED graph");
  GPM_CHECK_ARG(variant == GPM_TC_EDGE || variant == GPM_TC_VERTEX, "unknown TC variant %d",
                variant);
  GPM_CHECK_ARG(reps >= 1, "reps must be >= 1 (got %d)", reps);
  int st;
  if (variant == GPM_TC_EDGE && (st = ensure_src(g)) != GPM_OK) return st;
  if (variant == GPM_TC_VERTEX) {
    const RootTiers* t = nullptr;  // build (and cache) the tiers before timing
    if ((st = get_tc_tiers(g, kTcWaveCap, &t)) != GPM_OK) return st;
  }
  if ((st = timing_begin(g)) != GPM_OK) return st;
  for (int r = 0; r < reps; ++r)
    if ((st = enqueue_tc(g, variant, r == 0)) != GPM_OK) return st;
  if ((st = timing_end(g)) != GPM_OK) return st;
  g->last_ms /= reps;
  unsigned long long h = 0;
  GPM_HIP(hipMemcpyAsync(&h, g->counters, sizeof(h), hipMemcpyDeviceToHost, g->stream));
  GPM_HIP(hipStreamSynchronize(g->stream));
  *out = h;
  return GPM_OK;
}

}  // namespace

extern "C" {

int gpm_tc(gpm_graph* g, int variant, uint64_t* out) {
  clear_error();
  return tc_run(g, variant, 1, out);
}

int gpm_tc_repeat(gpm_graph* g, int variant, int reps, uint64_t* out) {
  clear_error();
  return tc_run(g, variant, reps, out);
}

}  // extern "C"
