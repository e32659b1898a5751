// Motif4 by counting identities (the counting formulation of gpm_pattern).
//
// Non-induced counts of the six connected 4-vertex shapes:
//   N_star    = sum_v C(d_v, 3)
//   N_path    = sum_{edges uv} (d_u - 1)(d_v - 1) - 3T
//   N_paw     = sum_v t_v (d_v - 2)            (tailed triangle)
//             = 1/2 sum_{edges uv} c_uv (d_u + d_v - 4)
//   N_cycle   = 4-cycles, counted per max-rank corner (below)
//   N_diamond = sum_{edges e} C(c_e, 2)        (c_e = triangles on e)
//   N_K4      = 4-cliques of the degree-oriented DAG
// The edge sums take every undirected edge once (the DAG's edges); c_e
// comes from listing every triangle once on the DAG (its first edge and apex)
// and adding one to each of its three edges.
// and the induced counts by inclusion-exclusion over how many copies of
// each shape a larger shape holds:
//   I_K4 = N_K4
//   I_diamond = N_diamond - 6 I_K4
//   I_cycle   = N_cycle - I_diamond - 3 I_K4
//   I_paw     = N_paw - 4 I_diamond - 12 I_K4
//   I_path    = N_path - 2 I_paw - 4 I_cycle - 6 I_diamond - 12 I_K4
//   I_star    = N_star - I_paw - 2 I_diamond - 4 I_K4
//
// 4-cycles: with rank = position in (degree, id) order, every 4-cycle has a
// unique top corner u of maximal rank; its wedges u-v-w with v, w below u
// land on the opposite corner w.  The walk runs on the graph relabelled by
// rank with sorted lists, so "below u" is a prefix of each list (one binary
// search, no rank lookups).  For each root u a block counts, for every w,
// the wedges reaching it in a block-private counter array in HBM; each
// increment adds the counter's previous value, which sums to
// sum_w C(cnt_w, 2).  The first increment of a counter also appends w to a
// block-private list, and only the listed counters are cleared afterwards.
#include <algorithm>
#include <numeric>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "graph_dev.hpp"

using namespace gpm;

namespace {

constexpr int kBlock = 256;
// 4-cycle corners: threads per block and blocks per CU (the best of the
// shapes measured at the launch).
constexpr int kCycBlock = 256;
constexpr int kCycBlocksPerCu = 1;
constexpr uint32_t kWedgesInFlight = 8;
// Counters of the wedge endpoints ranked just below the root, [u - kCycBand,
// u), live in LDS rather than in the block's HBM array.  On R-MAT scale 22,
// 32768 ranks below the root hold 43.6 % of the wedges of the top 1 % of
// roots (hubs link to hubs).  The walk took 1.39 s with every counter in
// HBM; with the band (words, wedges in flight): 32768 x 4 1.08 s,
// 32768 x 8 1.07 s, 40000 x 4 1.05 s, 40000 x 8 1.04 s.  40000 words with
// the walk's other arrays use 158 KB of the CU's 160 KB of LDS; the kernel
// runs one block per CU.
constexpr uint32_t kCycBand = 40000;

// Blocks of 256 threads for `threads` threads, at most 32 per CU (the
// kernels stride over the rest).
unsigned grid_of(const gpm_graph* g, uint64_t threads) {
  const uint64_t want = (threads + 255) / 256;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)device_cus(g) * 32));
}

// Triangles per DAG edge (c_e = |N(u) & N(v)| for the undirected edge of e):
// every triangle is found once, from its first DAG edge (u, v) and its apex
// w in N+(u) & N+(v) -- G lanes per edge, each taking entries of the shorter
// list and searching the longer (on R-MAT scale 22, 14 % fewer search steps
// than always taking N+(u)) -- and adds one to each of its three edges:
// (u, w) and (v, w) by their positions in the DAG lists, (u, v) once per
// edge with the group's total.  tri must be zero on entry; m_dag entries.
template <int G>
__global__ __launch_bounds__(kBlock) void edge_triangles_kernel(
    const uint32_t* __restrict__ doff, const uint32_t* __restrict__ dadj,
    const uint32_t* __restrict__ dsrc, uint64_t m, uint32_t* __restrict__ tri) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t sub = threadIdx.x & (G - 1);
  const uint64_t ngrp = (uint64_t)gridDim.x * blockDim.x / G;
  for (uint64_t e = tid / G; e < m; e += ngrp) {
    const uint32_t u = dsrc[e], v = dadj[e];
    const uint32_t ub = doff[u], nu = doff[u + 1] - ub;
    const uint32_t vb = doff[v], nv = doff[v + 1] - vb;
    // a = the list walked, b = the list searched
    const bool swap = nu > nv;
    const uint32_t ab = swap ? vb : ub, na = swap ? nv : nu;
    const uint32_t bb = swap ? ub : vb, nb = swap ? nu : nv;
    uint32_t cnt = 0;
    for (uint32_t i = sub; i < na; i += G) {
      const uint32_t w = dadj[ab + i];
      // position of w in list b (lower bound), if present
      uint32_t lo = 0, len = nb;
      while (len > 1) {
        const uint32_t half = len >> 1;
        lo = dadj[bb + lo + half] <= w ? lo + half : lo;
        len -= half;
      }
      if (nb && dadj[bb + lo] == w) {
        atomicAdd(&tri[ab + i], 1u);
        atomicAdd(&tri[bb + lo], 1u);
        ++cnt;
      }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, kWave);
    if (sub == 0 && cnt) atomicAdd(&tri[e], cnt);
  }
}

// Edge sums over the undirected edges {u, v} (each once: the DAG's edges),
// from the per-edge triangle counts c and the symmetric degrees:
//   out[0] += 2c                      (sum: 6T)
//   out[2] += c (d_u + d_v - 4)       (sum: 2 N_paw)
//   out[3] += c (c - 1)               (sum: 2 N_diamond)
//   out[4] += 2 (d_u - 1)(d_v - 1)    (sum: 2 x the path edge sum)
__global__ __launch_bounds__(kBlock) void edge_stats_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ dsrc,
    const uint32_t* __restrict__ dadj, const uint32_t* __restrict__ tri, uint64_t m,
    unsigned long long* __restrict__ out) {
  uint64_t s6t = 0, spaw = 0, sdia = 0, spath = 0;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = dsrc[e], v = dadj[e];
    const uint64_t c = tri[e], da = off[u + 1] - off[u], db = off[v + 1] - off[v];
    s6t += 2 * c;
    spaw += c ? c * (da + db - 4) : 0;
    sdia += c * (c - (c ? 1 : 0));
    spath += 2 * (da - 1) * (db - 1);
  }
  s6t = wave_sum_u64(s6t);
  spaw = wave_sum_u64(spaw);
  sdia = wave_sum_u64(sdia);
  spath = wave_sum_u64(spath);
  if (lane_id() == 0) {
    atomicAdd(&out[0], (unsigned long long)s6t);
    atomicAdd(&out[2], (unsigned long long)spaw);
    atomicAdd(&out[3], (unsigned long long)sdia);
    atomicAdd(&out[4], (unsigned long long)spath);
  }
}

// out[1] += sum_v C(d_v, 3).
__global__ __launch_bounds__(kBlock) void star_sum_kernel(const uint32_t* __restrict__ off,
                                                          uint32_t n,
                                                          unsigned long long* __restrict__ out) {
  uint64_t acc = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint64_t d = off[v + 1] - off[v];
    acc += d >= 3 ? d * (d - 1) * (d - 2) / 6 : 0;
  }
  acc = wave_sum_u64(acc);
  if (lane_id() == 0 && acc) atomicAdd(&out[1], (unsigned long long)acc);
}

// Entries of the sorted list a[0..n) below key.
__device__ __forceinline__ uint32_t count_below(const uint32_t* __restrict__ a, uint32_t n,
                                                uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// 4-cycle walk of root u over the rank-relabelled graph (ids = rank, lists
// sorted): the wedges u - v - x with v, x < u are the prefixes below u of
// N(u) and of every such N(v).  Atomic increments whose previous values sum
// to sum_x C(cnt_x, 2); every x seen first is appended to `touched`
// (*s_nt entries).
template <int BS>
__device__ __forceinline__ void cycle_walk(const uint32_t* __restrict__ off,
                                           const uint32_t* __restrict__ adj, uint32_t u,
                                           uint32_t* __restrict__ cnt,
                                           uint32_t* __restrict__ touched, uint32_t* s_nt,
                                           uint64_t* acc, uint32_t* s_pre, uint32_t* s_beg,
                                           uint32_t* s_wsum, uint32_t* s_band) {
  const uint32_t band_lo = u > kCycBand ? u - kCycBand : 0u;
  const uint32_t b = uniform(off[u]);
  const uint32_t dl = uniform(count_below(adj + b, uniform(off[u + 1]) - b, u));
  for (uint32_t c0 = 0; c0 < dl; c0 += BS) {
    const uint32_t i = c0 + threadIdx.x;
    uint32_t len = 0, vb = 0;
    if (i < dl) {
      const uint32_t v = adj[b + i];
      vb = off[v];
      len = count_below(adj + vb, off[v + 1] - vb, u);
    }
    // Block scan of the prefix lengths.
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
    const uint32_t incl0 = wave_incl_scan(len);
    if (lane == kWave - 1) s_wsum[w] = incl0;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t k = 0; k < w; ++k) add += s_wsum[k];
    s_pre[threadIdx.x] = incl0 + add;
    s_beg[threadIdx.x] = vb;
    __syncthreads();
    const uint32_t total = uniform(s_pre[BS - 1]);
    // kWedgesInFlight wedges per thread per step: their endpoint loads and
    // counter atomics are independent, so their latencies overlap (the walk
    // runs few waves per CU: one block per CU holds the counters).  On R-MAT
    // scale 22, before the LDS band (device-scope fences per root): 1 in
    // flight 1.96 s, 4 1.85 s, 8 1.88 s; with workgroup-scope atomics and
    // fences 1.39 s.  With the band, 8 in flight beat 4.  Aggregating a
    // root's increments in an LDS hash table first (4.5 wedges per distinct
    // endpoint on the heaviest roots) measured 3.01 s: the table's inserts
    // and the barriers of its flushes cost more than the atomics saved.
    for (uint32_t f0 = threadIdx.x; f0 < total; f0 += BS * kWedgesInFlight) {
      uint32_t x[kWedgesInFlight], old[kWedgesInFlight];
#pragma unroll
      for (uint32_t k = 0; k < kWedgesInFlight; ++k) {
        const uint32_t f = f0 + k * BS;
        x[k] = kEmpty;
        if (f < total) {
          const uint32_t j = upper_index(s_pre, BS, f);
          const uint32_t excl = j ? s_pre[j - 1] : 0u;
          x[k] = adj[s_beg[j] + (f - excl)];
        }
      }
#pragma unroll
      for (uint32_t k = 0; k < kWedgesInFlight; ++k)
        old[k] = x[k] == kEmpty ? 1u
                 : x[k] >= band_lo ? wg_atomic_add(&s_band[x[k] - band_lo], 1u)
                                   : wg_atomic_add(&cnt[x[k]], 1u);
#pragma unroll
      for (uint32_t k = 0; k < kWedgesInFlight; ++k) {
        if (x[k] == kEmpty) continue;
        *acc += old[k];
        if (old[k] == 0u) touched[atomicAdd(s_nt, 1u)] = x[k];
      }
    }
    __syncthreads();
  }
}

// Block per root, highest rank first (dynamic queue).  Each block owns
// 2 n words at counters + blockIdx.x * stride: n counters, then the list of
// the counters the current root touched.
template <int BS>
__global__ __launch_bounds__(BS) void cycles_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj, uint32_t n,
    unsigned int* __restrict__ queue, uint32_t* __restrict__ counters, uint64_t stride,
    unsigned long long* __restrict__ out) {
  __shared__ uint32_t s_pre[BS];
  __shared__ uint32_t s_beg[BS];
  __shared__ uint32_t s_wsum[BS / kWave];
  __shared__ uint32_t s_root;
  __shared__ uint32_t s_nt;
  __shared__ uint32_t s_band[kCycBand];
  uint32_t* cnt = counters + (uint64_t)blockIdx.x * stride;
  uint32_t* touched = cnt + n;
  uint64_t acc = 0;
  if (threadIdx.x == 0) s_nt = 0u;
  for (uint32_t i = threadIdx.x; i < kCycBand; i += BS) s_band[i] = 0u;
  while (true) {
    if (threadIdx.x == 0) s_root = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t q = uniform(s_root);
    __syncthreads();
    if (q >= n) break;
    const uint32_t u = n - 1 - q;
    cycle_walk<BS>(off, adj, u, cnt, touched, &s_nt, &acc, s_pre, s_beg, s_wsum, s_band);
    // Counter updates and list appends of this root must be complete before
    // the listed counters are cleared, and the clears before the next root's
    // increments; the counters are this block's alone, so workgroup-scope
    // fences order them.
    __threadfence_block();
    __syncthreads();
    const uint32_t nt = uniform(s_nt);
    const uint32_t band_lo = u > kCycBand ? u - kCycBand : 0u;
    for (uint32_t i = threadIdx.x; i < nt; i += BS) {
      const uint32_t x = touched[i];
      if (x >= band_lo)
        s_band[x - band_lo] = 0u;
      else
        cnt[x] = 0u;
    }
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x == 0) s_nt = 0u;
  }
  acc = wave_sum_u64(acc);
  if (lane_id() == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

__global__ void zero_u32_kernel(unsigned int* p, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// keys[e] = rank(u) << 32 | rank(x) for every adjacency entry e = (u, x).
__global__ __launch_bounds__(kBlock) void rank_keys_kernel(const uint32_t* __restrict__ off,
                                                           const uint32_t* __restrict__ adj,
                                                           const uint32_t* __restrict__ rank,
                                                           uint32_t n, uint64_t* __restrict__ keys) {
  const uint32_t lane = lane_id();
  const uint32_t nw = gridDim.x * blockDim.x / kWave;
  for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; v < n; v += nw) {
    const uint64_t hi = (uint64_t)rank[v] << 32;
    const uint32_t b = off[v], e = off[v + 1];
    for (uint32_t i = b + lane; i < e; i += kWave) keys[i] = hi | rank[adj[i]];
  }
}

__global__ void low_words_kernel(const uint64_t* __restrict__ keys, uint64_t m,
                                 uint32_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)keys[i];
}

// The graph relabelled by degree rank (position in (degree, id) order),
// every list sorted: rows of (rank(u), rank(x)) keys sorted as 64-bit
// integers, the row offsets from the degrees in rank order.
int ensure_rank_csr(gpm_graph* g) {
  if (g->rk_off) return GPM_OK;
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  std::vector<uint32_t> order(n), rank(n), rk_off(n + 1, 0u);
  std::iota(order.begin(), order.end(), 0u);
  auto deg = [&](uint32_t v) { return g->h_off[v + 1] - g->h_off[v]; };
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return deg(a) < deg(b) || (deg(a) == deg(b) && a < b);
  });
  for (uint32_t i = 0; i < n; ++i) {
    rank[order[i]] = i;
    rk_off[i + 1] = rk_off[i] + deg(order[i]);
  }
  uint32_t* d_rank = nullptr;
  uint64_t* keys = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  unsigned bits = 32;
  while (bits < 64 && (1ull << (bits - 32)) < (uint64_t)std::max(n, 1u)) ++bits;
  int st = GPM_OK;
  auto cleanup = [&]() {
    (void)hipFree(d_rank);
    (void)hipFree(keys);
    (void)hipFree(temp);
  };
  auto fail_all = [&](int code) {  // leaves no half-built graph on the handle
    cleanup();
    (void)hipFree(g->rk_off);
    (void)hipFree(g->rk_adj);
    g->rk_off = g->rk_adj = nullptr;
    return code;
  };
  if (hipMalloc(&g->rk_off, (size_t)(n + 1) * 4) != hipSuccess ||
      hipMalloc(&g->rk_adj, std::max<uint64_t>(m, 1) * 4) != hipSuccess ||
      hipMalloc(&d_rank, (size_t)std::max(n, 1u) * 4) != hipSuccess ||
      hipMalloc(&keys, std::max<uint64_t>(m, 1) * 16) != hipSuccess)
    return fail_all(
        fail(GPM_ENOMEM, "rank-relabelled graph (%llu entries)", (unsigned long long)m));
  if ((st = h2d(g, g->rk_off, rk_off.data(), (size_t)(n + 1) * 4)) != GPM_OK ||
      (st = h2d(g, d_rank, rank.data(), (size_t)n * 4)) != GPM_OK)
    return fail_all(st);
  if (m) {
    uint64_t* sorted = keys + m;
    hipLaunchKernelGGL(rank_keys_kernel, dim3(grid_of(g, (uint64_t)n * kWave)),
                       dim3(kBlock), 0, g->stream, g->off, g->adj, d_rank, n, keys);
    if (rocprim::radix_sort_keys(nullptr, temp_bytes, keys, sorted, m, 0u, bits, g->stream) !=
            hipSuccess ||
        hipMalloc(&temp, std::max<size_t>(temp_bytes, 1)) != hipSuccess ||
        rocprim::radix_sort_keys(temp, temp_bytes, keys, sorted, m, 0u, bits, g->stream) !=
            hipSuccess)
      return fail_all(fail(GPM_EHIP, "sorting the rank-relabelled graph"));
    hipLaunchKernelGGL(low_words_kernel, dim3(grid_of(g, m)), dim3(256), 0, g->stream,
                       sorted, m, g->rk_adj);
  }
  if (hipStreamSynchronize(g->stream) != hipSuccess)
    return fail_all(fail(GPM_EHIP, "building the rank-relabelled graph"));
  cleanup();
  return GPM_OK;
}

}  // namespace

namespace gpm {

int motif4_by_identities(gpm_graph* g, uint64_t out[6]) {
  int st;
  if ((st = ensure_rank_csr(g)) != GPM_OK) return st;
  if (!g->dag && (st = gpm_graph_orient(g, &g->dag)) != GPM_OK) return st;
  if ((st = ensure_src(g->dag)) != GPM_OK) return st;
  const int cus = device_cus(g);
  // Scratch: per block, n 4-cycle counters and a list of up to n touched
  // counters; kCycBlocksPerCu blocks of kCycBlock threads per CU, within
  // 8 GiB (fewer blocks when that much is not free).  The wedge increments
  // are random atomics into these arrays: more or wider blocks keep more of
  // them in flight but spread them over more memory.  4-cycle corners on
  // R-MAT scale 22 (blocks x threads): 256 x 256 1.96 s, 128 x 256 2.62 s,
  // 512 x 256 2.24 s, 1024 x 256 2.91 s, 2048 x 256 2.97 s, 256 x 512 2.47 s,
  // 256 x 1024 3.64 s, 128 x 1024 3.44 s.
  const uint64_t stride = 2 * (uint64_t)std::max<uint32_t>(g->n, 1);
  unsigned blocks_c = (unsigned)std::min<uint64_t>((uint64_t)cus * kCycBlocksPerCu, std::max(g->n, 1u));
  while (blocks_c > 16 && (uint64_t)blocks_c * stride * 4 > (8ull << 30)) blocks_c /= 2;
  // The per-edge triangle counts (one uint32 per DAG edge) use the same
  // scratch before the 4-cycle counters do.
  const uint64_t tri_bytes = std::max<uint64_t>(g->dag->m, 1) * 4;
  while (true) {
    const uint64_t want = std::max((uint64_t)blocks_c * stride * 4, tri_bytes) + 256;
    if ((st = ensure_scratch(g, want)) == GPM_OK) break;
    if (st != GPM_ENOMEM || blocks_c <= 16) return st;
    clear_error();
    blocks_c /= 2;
  }
  const uint64_t cyc_bytes = (uint64_t)blocks_c * stride * 4;
  uint32_t* tri = static_cast<uint32_t*>(g->scratch);
  uint32_t* ccnt = static_cast<uint32_t*>(g->scratch);
  if ((st = timing_begin(g)) != GPM_OK) return st;
  if ((st = zero_counters(g, 16)) != GPM_OK) return st;
  unsigned int* queue = reinterpret_cast<unsigned int*>(g->counters + 8);
  const uint64_t me = g->dag->m;
  if (me) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_of(g, me)), dim3(256), 0, g->stream, tri, me);
    hipLaunchKernelGGL(edge_triangles_kernel<8>, dim3(grid_of(g, me * 8)), dim3(kBlock), 0,
                       g->stream, g->dag->off, g->dag->adj, g->dag->src, me, tri);
    hipLaunchKernelGGL(edge_stats_kernel, dim3(grid_of(g, me)), dim3(kBlock), 0, g->stream,
                       g->off, g->dag->src, g->dag->adj, tri, me, g->counters);
    GPM_HIP(hipGetLastError());
  }
  timing_mark(g, "motif_edge_sums");
  hipLaunchKernelGGL(star_sum_kernel, dim3(grid_of(g, g->n)), dim3(kBlock), 0, g->stream, g->off,
                     g->n, g->counters);
  GPM_HIP(hipGetLastError());
  timing_mark(g, "motif_star_sum");
  // The 4-cycle counters start at zero (the triangle counts lived here).
  const uint64_t zero_n = cyc_bytes / 4;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_of(g, zero_n)), dim3(256), 0, g->stream, ccnt,
                     zero_n);
  hipLaunchKernelGGL(cycles_kernel<kCycBlock>, dim3(blocks_c), dim3(kCycBlock), 0, g->stream,
                     g->rk_off, g->rk_adj, g->n, queue, ccnt, stride, g->counters + 5);
  GPM_HIP(hipGetLastError());
  timing_mark(g, "motif_cycles");
  if ((st = timing_end(g)) != GPM_OK) return st;
  double ms_main = g->last_ms;
  unsigned long long h[6];
  GPM_HIP(hipMemcpyAsync(h, g->counters, sizeof(h), hipMemcpyDeviceToHost, g->stream));
  GPM_HIP(hipStreamSynchronize(g->stream));
  uint64_t k4 = 0;
  if ((st = gpm_clique(g->dag, 4, nullptr, &k4)) != GPM_OK) return st;
  g->last_ms = ms_main + g->dag->last_ms;
  const uint64_t T = h[0] / 6;
  const uint64_t n_star = h[1];
  const uint64_t n_paw = h[2] / 2;
  const uint64_t n_dia = h[3] / 2;
  const uint64_t n_path = h[4] / 2 - 3 * T;
  const uint64_t n_cyc = h[5];
  const uint64_t i_k4 = k4;
  const uint64_t i_dia = n_dia - 6 * i_k4;
  const uint64_t i_cyc = n_cyc - i_dia - 3 * i_k4;
  const uint64_t i_paw = n_paw - 4 * i_dia - 12 * i_k4;
  const uint64_t i_path = n_path - 2 * i_paw - 4 * i_cyc - 6 * i_dia - 12 * i_k4;
  const uint64_t i_star = n_star - i_paw - 2 * i_dia - 4 * i_k4;
  out[0] = i_star;
  out[1] = i_path;
  out[2] = i_paw;
  out[3] = i_cyc;
  out[4] = i_dia;
  out[5] = i_k4;
  return GPM_OK;
}

}  // namespace gpm
