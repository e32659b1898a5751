// Triangle counting on a degree-oriented DAG (replaces src/TCSolver.cuh).
//
//  GPM_TC_EDGE   (e_centric, TCSolver.cuh:27-59): sub-wave groups of G lanes
//                per oriented edge (u, v); each lane takes elements of the
//                shorter list and binary-searches the longer one, or, when
//                the longer list has at most 2G entries, the group compares
//                the two lists' chunks pairwise by XOR lane exchanges
//                (ds_swizzle).  G follows the mean out-degree so that
//                low-degree graphs keep all 64 lanes busy.  Blocks store
//                partial sums that a one-block kernel adds up.
//  GPM_TC_VERTEX (v_centric, TCSolver.cuh:64-113, tc_hi_warp_vertex.cuh):
//                one wave per root u with |N+(u)| <= 256 (roots interleaved
//                statically over the waves), one 256-thread block per larger
//                root.  N+(u) goes into an LDS hash table;
//                the lists N+(v), v in N+(u), are walked as one flattened
//                range so that every lane reads one entry per step
//                (coalesced, no idle lanes on short lists).  Roots above the
//                LDS capacity probe N+(u) by binary search instead.
//  Count = sum over oriented edges (u, v) of |N+(u) & N+(v)|.
#include <algorithm>
#include <mutex>
#include <utility>
#include <vector>

#include "graph_dev.hpp"
#include "tiers.hpp"

using namespace gpm;

// Scratch of gpm_tc_edge_noalloc: device memory the library never
// allocates -- each GPU holds its own copy of this array from the moment the
// code object is loaded until the process ends.  Layout: the edge kernel's
// per-block partial sums (uint64, up to kNoAllocMaxBlocks) from byte 0, then
// the COO source ids (uint32) of up to kNoAllocMaxEdges oriented edges from
// byte kNoAllocSrcOffset.  Host code must not use the array's name as a
// device address: it resolves the per-device address with
// hipGetSymbolAddress once per device (noalloc_scratch below).
constexpr uint64_t kNoAllocMaxEdges = 16ull << 20;  // 16,777,216 oriented edges
constexpr uint64_t kNoAllocMaxBlocks = 8192;        // edge-kernel blocks (partial sums)
constexpr size_t kNoAllocSrcOffset = kNoAllocMaxBlocks * sizeof(unsigned long long);
constexpr size_t kNoAllocScratchBytes = kNoAllocSrcOffset + kNoAllocMaxEdges * sizeof(uint32_t);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kTcWaveCap = 256;     // roots served by one wave
constexpr uint32_t kTcWaveTable = 512;   // LDS keys per wave
constexpr uint32_t kTcBlockTableMax = 8192;  // 32 KiB of keys per block
constexpr uint64_t kTcFusedEdgesPerGroup = 4;  // edges per group, fused repeat
constexpr uint64_t kTcEdgeBlocksPerCu = 32;    // edge-kernel grid cap per CU
constexpr int kTcGraphReps = 32;   // counts per launch of the repeat graph
constexpr int kTcGraphChains = 8;  // independent chains the counts form

// v from the lane whose group index is (this lane's) ^ X, for X < 32: a
// ds_swizzle in bit-mask mode (and 0x1F, or 0, xor X) needs no lane-address
// registers, unlike ds_bpermute.
template <int X>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (X << 10));
}

template <int... X>
__device__ __forceinline__ bool group_has_xor(uint32_t a, uint32_t b,
                                              std::integer_sequence<int, X...>) {
  return (a == b) | ((a == lane_xor<X + 1>(b)) | ...);
}

// True when a equals the b of any lane of this lane's G-lane group: lane s
// meets lane s ^ r for r = 0..G-1, i.e. every lane of the group once.  A
// full wave (G = 64) first swaps its two 32-lane halves once, then runs the
// 32-lane exchange on both copies.
template <int G>
__device__ __forceinline__ bool group_has(uint32_t a, uint32_t b) {
  static_assert(G >= 2 && G <= kWave && (G & (G - 1)) == 0, "group size");
  if constexpr (G <= 32) {
    return group_has_xor(a, b, std::make_integer_sequence<int, G - 1>{});
  } else {
    const uint32_t bh = (uint32_t)__shfl_xor((int)b, 32, kWave);
    return group_has_xor(a, b, std::make_integer_sequence<int, 31>{}) |
           group_has_xor(a, bh, std::make_integer_sequence<int, 31>{});
  }
}

// Edge-parallel count; block (x, y) stores the sum of its lanes' counts in
// partial[y * gridDim.x + x] (a plain store, every block, so the array needs
// no clearing), and tc_sum_kernel adds them up.  A launch with gridDim.y = R
// runs R independent counts (the fused repeat of gpm_tc_repeat_fused).
// One atomic per triangle-holding wave into one counter measured slower:
// atomics to a single address serialise (about 8 ns each here), which made
// the count on Erdos-Renyi n = 16384, M = 131072 (666 triangles) grow with
// its triangles rather than its edges.
template <int G>
__global__ __launch_bounds__(kBlock) void tc_edge_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ src, uint64_t m, unsigned long long* __restrict__ partial) {
  __shared__ unsigned long long s_wave[kBlock / kWave];
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
    if (nb <= 2 * G) {
      // Short lists: one load per lane per list chunk, then every pair of
      // the two chunks is compared by exchanging the large chunk across the
      // group -- no dependent search loads.  Entries are distinct, so a
      // lane's entry matches at most once per chunk: the compares are OR-ed
      // (lane masks, scalar ops) and counted once.
      uint32_t c32 = 0;
      for (uint32_t i0 = 0; i0 < na; i0 += G) {
        const uint32_t a = i0 + sub < na ? small[i0 + sub] : 0xFFFFFFFFu;
        for (uint32_t j0 = 0; j0 < nb; j0 += G) {
          const uint32_t b = j0 + sub < nb ? large[j0 + sub] : 0xFFFFFFFEu;
          c32 += group_has<G>(a, b) ? 1u : 0u;
        }
      }
      cnt += c32;
    } else {
      for (uint32_t i = sub; i < na; i += G) cnt += sorted_contains(large, nb, small[i]);
    }
  }
  // Sparse graphs leave most waves without a triangle: those skip the
  // cross-lane sum.
  if (__ballot(cnt != 0) != 0) cnt = wave_sum_u64(cnt);
  if (lane_id() == 0) s_wave[threadIdx.x / kWave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = 0;
    for (int w = 0; w < kBlock / kWave; ++w) b += s_wave[w];
    partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = b;
  }
}

// out[y] = sum of partial[y * stride, y * stride + nb): one block per count,
// a plain store (the counter needs no clearing).
__global__ __launch_bounds__(kBlock) void tc_sum_kernel(
    const unsigned long long* __restrict__ partial, uint32_t nb, uint32_t stride,
    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_wave[kBlock / kWave];
  const unsigned long long* p = partial + (size_t)blockIdx.x * stride;
  uint64_t v = 0;
  for (uint32_t i = threadIdx.x; i < nb; i += kBlock) v += p[i];
  v = wave_sum_u64(v);
  if (lane_id() == 0) s_wave[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / kWave; ++w) t += s_wave[w];
    out[blockIdx.x] = t;
  }
}

// Wave per root: roots[] lists vertices with 2 <= |N+| <= kTcWaveCap, in
// chunks of `chunk` roots; wave w takes chunks w, w + waves, ... (the list is
// ordered by degree, so the interleave evens the waves' loads out).  No task
// queue and one atomic per block: atomics to one address serialise (about
// 10 ns each), and with one queue fetch per chunk and one sum per wave they
// were most of the kernel's time on Erdos-Renyi n = 16384.
__global__ __launch_bounds__(kBlock) void tc_vertex_wave_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ roots, uint32_t n_roots, uint32_t chunk,
    unsigned long long* out) {
  constexpr int kWavesPerBlock = kBlock / kWave;
  __shared__ uint32_t s_keys[kWavesPerBlock][kTcWaveTable];
  __shared__ uint32_t s_pre[kWavesPerBlock][kWave];
  __shared__ uint32_t s_beg[kWavesPerBlock][kWave];
  __shared__ unsigned long long s_sum[kWavesPerBlock];
  const uint32_t lane = lane_id();
  const uint32_t w = threadIdx.x / kWave;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / kWave;
  uint32_t* keys = s_keys[w];
  uint32_t* pre = s_pre[w];
  uint32_t* beg = s_beg[w];
  uint64_t cnt = 0;
  for (uint64_t base = wave * chunk; base < n_roots; base += nwaves * chunk) {
    const uint32_t stop = (uint32_t)min<uint64_t>(base + chunk, n_roots);
    for (uint32_t r = (uint32_t)base; r < stop; ++r) {
      const uint32_t u = uniform(roots[r]);
      const uint32_t sb = uniform(off[u]), d = uniform(off[u + 1]) - sb;
      // Open-addressing table of at least 2d slots (d <= kTcWaveCap).
      const uint32_t size = d <= 32 ? 64u : d <= 64 ? 128u : d <= 128 ? 256u : kTcWaveTable;
      const uint32_t mask = size - 1;
      for (uint32_t i = lane; i < size; i += kWave) keys[i] = kEmpty;
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
  if (lane == 0) s_sum[w] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long total = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) total += s_sum[i];
    if (total) atomicAdd(out, total);
  }
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

// src[e] = u for every e in [off[u], off[u+1]) with e < m; one wave per
// vertex, grid-stride.  Entries past m are never written, whatever the
// offsets hold.
__global__ void noalloc_src_kernel(const uint32_t* __restrict__ off, uint32_t n, uint64_t m,
                                   uint32_t* __restrict__ src) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const uint32_t lane = lane_id();
  for (uint32_t u = wave; u < n; u += nwaves) {
    const uint64_t b = off[u], e = min<uint64_t>(off[u + 1], m);
    for (uint64_t i = b + lane; i < e; i += kWave) src[i] = u;
  }
}

__global__ void zero_u32_kernel(unsigned int* p, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

int edge_group_size(uint32_t n, uint64_t m) {
  const double avg = n ? (double)m / n : 0.0;
  if (avg <= 6) return 4;
  if (avg <= 12) return 8;
  if (avg <= 24) return 16;
  if (avg <= 48) return 32;
  return 64;
}

// Grid of the edge kernel: one G-lane group per edges_per_group oriented
// edges, at most 32 blocks per CU (the groups then stride over the edges).
uint64_t edge_blocks(uint64_t m, int G, int cus, uint64_t edges_per_group, uint64_t max_blocks) {
  const uint64_t groups = (std::max<uint64_t>(m, 1) + edges_per_group - 1) / edges_per_group;
  const uint64_t cap = std::min<uint64_t>((uint64_t)cus * kTcEdgeBlocksPerCu, max_blocks);
  return std::max<uint64_t>(1, std::min<uint64_t>((groups * G + kBlock - 1) / kBlock, cap));
}

const void* edge_kernel(int G) {
  switch (G) {
    case 4: return reinterpret_cast<const void*>(&tc_edge_kernel<4>);
    case 8: return reinterpret_cast<const void*>(&tc_edge_kernel<8>);
    case 16: return reinterpret_cast<const void*>(&tc_edge_kernel<16>);
    case 32: return reinterpret_cast<const void*>(&tc_edge_kernel<32>);
    default: return reinterpret_cast<const void*>(&tc_edge_kernel<64>);
  }
}

hipError_t launch_edge(int G, dim3 grid, hipStream_t stream, const uint32_t* off,
                       const uint32_t* adj, const uint32_t* src, uint64_t m,
                       unsigned long long* partial) {
  void* args[] = {&off, &adj, &src, &m, &partial};
  return hipLaunchKernel(edge_kernel(G), grid, dim3(kBlock), args, 0, stream);
}

// Per-block partial sums of the edge kernel: kTcGraphReps counts of up to
// kTcEdgeBlocksPerCu blocks per CU each, kept for the handle's lifetime (the
// repeat graph's nodes hold the address).
int ensure_tc_partial(gpm_graph* g) {
  if (g->tc_partial) return GPM_OK;
  const size_t n = (size_t)device_cus(g) * kTcEdgeBlocksPerCu * kTcGraphReps;
  if (hipMalloc(&g->tc_partial, n * sizeof(unsigned long long)) != hipSuccess) {
    g->tc_partial = nullptr;
    (void)hipGetLastError();
    return fail(GPM_ENOMEM, "triangle-count partial sums (%zu entries)", n);
  }
  return GPM_OK;
}

// Enqueues one triangle count (counter reset + kernels) on g's stream; the
// count lands in g->counters[0].  mark: record per-kernel timing marks.
int enqueue_tc(gpm_graph* g, int variant, bool mark) {
  int st;
  const int cus = device_cus(g);
  if (variant == GPM_TC_EDGE) {
    if ((st = ensure_tc_partial(g)) != GPM_OK) return st;
    const int G = edge_group_size(g->n, g->m);
    const uint64_t blocks = edge_blocks(g->m, G, cus, 1, ~0ull);
    GPM_HIP(launch_edge(G, dim3((unsigned)blocks), g->stream, g->off, g->adj, g->src, g->m,
                        g->tc_partial));
    if (mark) timing_mark(g, "tc_edge");
    hipLaunchKernelGGL(tc_sum_kernel, dim3(1), dim3(kBlock), 0, g->stream, g->tc_partial,
                       (uint32_t)blocks, (uint32_t)blocks, g->counters);
    GPM_HIP(hipGetLastError());
    if (mark) timing_mark(g, "tc_sum");
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
                       g->adj, t->d_wave, t->n_wave, chunk, g->counters);
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

// hipGraph of kTcGraphReps edge-parallel counts: count r's edge kernel
// stores its partial sums in slice r of g->tc_partial, and count r starts
// after count r - kTcGraphChains, so the counts form kTcGraphChains
// independent chains the device may run side by side.  One sum node after
// the chains adds every slice into its counter slot.  A launch measures the
// counting throughput of short launches without per-launch host cost.
int ensure_tc_graph(gpm_graph* g) {
  if (g->tc_exec) return GPM_OK;
  int st;
  if ((st = ensure_tc_partial(g)) != GPM_OK) return st;
  const int G = edge_group_size(g->n, g->m);
  const int cus = device_cus(g);
  const uint64_t blocks = edge_blocks(g->m, G, cus, 1, ~0ull);
  const size_t slice = (size_t)cus * kTcEdgeBlocksPerCu;
  GPM_HIP(hipGraphCreate(&g->tc_graph, 0));
  const uint32_t *off = g->off, *adj = g->adj, *src = g->src;
  uint64_t m = g->m;
  hipGraphNode_t nodes[kTcGraphReps] = {};
  for (int r = 0; r < kTcGraphReps; ++r) {
    unsigned long long* part = g->tc_partial + r * slice;
    void* eargs[] = {&off, &adj, &src, &m, &part};
    hipKernelNodeParams ep{};
    ep.func = const_cast<void*>(edge_kernel(G));
    ep.gridDim = dim3((unsigned)blocks);
    ep.blockDim = dim3(kBlock);
    ep.kernelParams = eargs;
    const hipGraphNode_t* dep = r < kTcGraphChains ? nullptr : &nodes[r - kTcGraphChains];
    GPM_HIP(hipGraphAddKernelNode(&nodes[r], g->tc_graph, dep, dep ? 1 : 0, &ep));
  }
  // The slices are `slice` apart: the sum kernel reads the first `blocks`
  // entries of each (its stride argument is the slice, its length blocks).
  unsigned long long* part = g->tc_partial;
  unsigned long long* counters = g->counters;
  uint32_t nb = (uint32_t)blocks, stride = (uint32_t)slice;
  void* sargs[] = {&part, &nb, &stride, &counters};
  hipKernelNodeParams sp{};
  sp.func = reinterpret_cast<void*>(&tc_sum_kernel);
  sp.gridDim = dim3(kTcGraphReps);
  sp.blockDim = dim3(kBlock);
  sp.kernelParams = sargs;
  hipGraphNode_t sn = nullptr;
  GPM_HIP(hipGraphAddKernelNode(&sn, g->tc_graph, &nodes[kTcGraphReps - kTcGraphChains],
                                kTcGraphChains, &sp));
  GPM_HIP(hipGraphInstantiate(&g->tc_exec, g->tc_graph, nullptr, nullptr, 0));
  return GPM_OK;
}

// reps edge-parallel counts as launches of kTcGraphReps counts each: one
// kernel whose grid y index picks the count (and its counter slot), so the
// counts of a launch overlap their dependent-load chains on the device.
int tc_fused_repeat(gpm_graph* g, int reps, uint64_t* out) {
  GPM_CHECK_ARG(g && out, "null pointer");
  GPM_CHECK_ARG(g->flags & GPM_GRAPH_ORIENTED, "triangle counting needs an ORIENTED graph");
  GPM_CHECK_ARG(reps >= 1, "reps must be >= 1 (got %d)", reps);
  int st;
  if ((st = ensure_src(g)) != GPM_OK) return st;
  if ((st = ensure_tc_partial(g)) != GPM_OK) return st;
  const int G = edge_group_size(g->n, g->m);
  const int cus = device_cus(g);
  // The counts of a launch fill the device together, so each group takes
  // kTcFusedEdgesPerGroup edges: fewer waves, each amortising its set-up
  // and final sum over more edges.
  const uint64_t blocks = edge_blocks(g->m, G, cus, kTcFusedEdgesPerGroup, ~0ull);
  unsigned long long h[kTcGraphReps] = {};
  if ((st = timing_begin(g)) != GPM_OK) return st;
  int done = 0;
  while (done < reps) {
    const int batch = std::min(kTcGraphReps, reps - done);
    // batch * blocks partial sums: within the kTcGraphReps slices.
    GPM_HIP(launch_edge(G, dim3((unsigned)blocks, (unsigned)batch), g->stream, g->off, g->adj,
                        g->src, g->m, g->tc_partial));
    if (done == 0) timing_mark(g, "tc_edge_fused");
    hipLaunchKernelGGL(tc_sum_kernel, dim3(batch), dim3(kBlock), 0, g->stream, g->tc_partial,
                       (uint32_t)blocks, (uint32_t)blocks, g->counters);
    GPM_HIP(hipGetLastError());
    done += batch;
    if (done >= reps) {
      GPM_HIP(hipMemcpyAsync(h, g->counters, batch * sizeof(h[0]), hipMemcpyDeviceToHost,
                             g->stream));
      if ((st = timing_end(g)) != GPM_OK) return st;
      GPM_HIP(hipStreamSynchronize(g->stream));
      for (int r = 1; r < batch; ++r)
        if (h[r] != h[0]) return fail(GPM_EHIP, "repeated triangle counts disagree");
    }
  }
  g->last_ms /= (double)reps;
  *out = h[0];
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
  const bool graph = variant == GPM_TC_EDGE && reps > 1;
  if (graph && (st = ensure_tc_graph(g)) != GPM_OK) return st;
  int launches = reps;
  if ((st = timing_begin(g)) != GPM_OK) return st;
  if (graph) {
    launches = (reps + kTcGraphReps - 1) / kTcGraphReps;
    for (int r = 0; r < launches; ++r) GPM_HIP(hipGraphLaunch(g->tc_exec, g->stream));
  } else {
    for (int r = 0; r < reps; ++r)
      if ((st = enqueue_tc(g, variant, r == 0)) != GPM_OK) return st;
  }
  if ((st = timing_end(g)) != GPM_OK) return st;
  g->last_ms /= graph ? (double)launches * kTcGraphReps : (double)reps;
  unsigned long long h[kTcGraphReps] = {};
  const int nread = graph ? kTcGraphReps : 1;
  GPM_HIP(hipMemcpyAsync(h, g->counters, nread * sizeof(h[0]), hipMemcpyDeviceToHost, g->stream));
  GPM_HIP(hipStreamSynchronize(g->stream));
  for (int r = 1; r < nread; ++r)
    if (h[r] != h[0]) return fail(GPM_EHIP, "repeated triangle counts disagree");
  *out = h[0];
  return GPM_OK;
}

// Device address of `scratch` on the current device, resolved once per
// device with that device current and kept; every later call on the device
// uses the kept address.
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

int noalloc_scratch(void** out) {
  int dev = 0;
  GPM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(GPM_EUNSUPPORTED, "device %d", dev);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GPM_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_noalloc_ptr[dev] = p;
  }
  *out = g_noalloc_ptr[dev];
  return GPM_OK;
}

int tc_edge_noalloc(const uint32_t* off, const uint32_t* adj, uint32_t n, uint64_t m,
                    unsigned long long* count, hipStream_t stream) {
  GPM_CHECK_ARG(off && count, "null pointer");
  GPM_CHECK_ARG(adj || m == 0, "null adjacency with m = %llu", (unsigned long long)m);
  GPM_CHECK_ARG(n < 0xFFFFFFFFu, "n = %u too large", n);
  GPM_CHECK_ARG(m <= kNoAllocMaxEdges,
                "m = %llu oriented edges exceeds the no-alloc limit of %llu (its COO staging "
                "lives in a fixed %zu-byte device array)",
                (unsigned long long)m, (unsigned long long)kNoAllocMaxEdges,
                kNoAllocScratchBytes);
  void* base = nullptr;
  int st;
  if ((st = noalloc_scratch(&base)) != GPM_OK) return st;
  unsigned long long* partial = static_cast<unsigned long long*>(base);
  uint32_t* src = reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(base) + kNoAllocSrcOffset);
  int dev = 0, cus = 256;
  GPM_HIP(hipGetDevice(&dev));
  GPM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // Every launched block stores its partial sum and the sum kernel stores
  // the count, so nothing keeps a previous call's value (and no memset node
  // is needed: the sequence may be captured into a hipGraph).  With m = 0
  // one block of the edge kernel runs no edge and stores 0.
  uint64_t blocks = 1;
  if (m) {
    const uint64_t sblocks = std::max<uint64_t>(
        1, std::min<uint64_t>(((uint64_t)n * kWave + kBlock - 1) / kBlock, (uint64_t)cus * 16));
    hipLaunchKernelGGL(noalloc_src_kernel, dim3((unsigned)sblocks), dim3(kBlock), 0, stream, off,
                       n, m, src);
    GPM_HIP(hipGetLastError());
    const int G = edge_group_size(n, m);
    blocks = edge_blocks(m, G, cus, 1, kNoAllocMaxBlocks);
    GPM_HIP(launch_edge(G, dim3((unsigned)blocks), stream, off, adj, src, m, partial));
  } else {
    GPM_HIP(launch_edge(4, dim3(1), stream, off, adj, src, 0, partial));
  }
  hipLaunchKernelGGL(tc_sum_kernel, dim3(1), dim3(kBlock), 0, stream, partial, (uint32_t)blocks,
                     (uint32_t)blocks, count);
  GPM_HIP(hipGetLastError());
  return GPM_OK;
}

}  // namespace

extern "C" {

int gpm_tc_edge_noalloc(const uint32_t* d_offsets, const uint32_t* d_adj, uint32_t n,
                        uint64_t m, uint64_t* d_count, void* stream) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    return tc_edge_noalloc(d_offsets, d_adj, n, m,
                           reinterpret_cast<unsigned long long*>(d_count),
                           static_cast<hipStream_t>(stream));
  });
}

int gpm_tc(gpm_graph* g, int variant, uint64_t* out) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    return tc_run(g, variant, 1, out);
  });
}

int gpm_tc_repeat_fused(gpm_graph* g, int reps, uint64_t* out) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    return tc_fused_repeat(g, reps, out);
  });
}

int gpm_tc_repeat(gpm_graph* g, int variant, int reps, uint64_t* out) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    return tc_run(g, variant, reps, out);
  });
}

}  // extern "C"
