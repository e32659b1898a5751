// Motif4 by counting identities (the counting formulation of gpm_pattern).
//
// Non-induced counts of the six connected 4-vertex shapes:
//   N_star    = sum_v C(d_v, 3)
//   N_path    = sum_{edges uv} (d_u - 1)(d_v - 1) - 3T
//   N_paw     = sum_v t_v (d_v - 2)            (tailed triangle)
//   N_cycle   = 4-cycles, counted per max-rank corner (below)
//   N_diamond = sum_{edges e} C(c_e, 2)        (c_e = triangles on e)
//   N_K4      = 4-cliques of the degree-oriented DAG
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
// land on the opposite corner w.  For each root u a block counts, for every
// w, the wedges reaching it in a block-private counter array in HBM; each
// increment adds the counter's previous value, which sums to
// sum_w C(cnt_w, 2), and a second walk clears the counters.
#include <algorithm>
#include <numeric>
#include <vector>

#include "graph_dev.hpp"

using namespace gpm;

namespace {

constexpr int kBlock = 256;

// c[e] = |N(u) & N(v)| for every directed edge e = (u, v); G lanes per edge.
template <int G>
__global__ __launch_bounds__(kBlock) void edge_triangles_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ src, uint64_t m, uint32_t* __restrict__ c) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t sub = threadIdx.x & (G - 1);
  const uint64_t ngrp = (uint64_t)gridDim.x * blockDim.x / G;
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
    uint32_t cnt = 0;
    for (uint32_t i = sub; i < na; i += G) cnt += sorted_contains(adj + bb, nb, adj[ab + i]);
    // Sum over the G lanes of this group (aligned, power of two).
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, kWave);
    if (sub == 0) c[e] = cnt;
  }
}

// Per-vertex sums (wave per vertex): out[0] += sum_e c_e (6T),
// out[1] += sum_v C(d,3), out[2] += sum_v t_v (d_v - 2),
// out[3] += sum_e C(c_e, 2) (directed: 2 N_diamond),
// out[4] += sum_e (d_u - 1)(d_v - 1) (directed: 2 x the edge sum).
__global__ __launch_bounds__(kBlock) void vertex_sums_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ c, uint32_t n, unsigned long long* __restrict__ out) {
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  uint64_t s6t = 0, sstar = 0, spaw = 0, sdia = 0, spath = 0;
  for (uint32_t v = wave; v < n; v += nwaves) {
    const uint32_t b = off[v], e = off[v + 1];
    const uint64_t dv = e - b;
    uint64_t t2 = 0, dia = 0, path = 0;
    for (uint32_t i = b + lane; i < e; i += kWave) {
      const uint64_t ce = c[i];
      const uint32_t u = adj[i];
      const uint64_t du = off[u + 1] - off[u];
      t2 += ce;
      dia += ce * (ce - (ce ? 1 : 0)) / 2;
      path += (dv - 1) * (du - 1);
    }
    t2 = wave_sum_u64(t2);
    dia = wave_sum_u64(dia);
    path = wave_sum_u64(path);
    if (lane == 0) {
      s6t += t2;
      sstar += dv >= 3 ? dv * (dv - 1) * (dv - 2) / 6 : 0;
      spaw += dv >= 2 ? (t2 / 2) * (dv - 2) : 0;
      sdia += dia;
      spath += path;
    }
  }
  if (lane == 0) {
    atomicAdd(&out[0], (unsigned long long)s6t);
    atomicAdd(&out[1], (unsigned long long)sstar);
    atomicAdd(&out[2], (unsigned long long)spaw);
    atomicAdd(&out[3], (unsigned long long)sdia);
    atomicAdd(&out[4], (unsigned long long)spath);
  }
}

// Block per root u (dynamic queue).  PASS 0 counts (atomic increments whose
// previous values sum to sum_w C(cnt_w, 2)); PASS 1 clears the counters.
__device__ __forceinline__ void cycle_walk(const uint32_t* __restrict__ off,
                                           const uint32_t* __restrict__ adj,
                                           const uint32_t* __restrict__ rank, uint32_t u,
                                           uint32_t* __restrict__ cnt, bool clear,
                                           uint64_t* acc, uint32_t* s_pre, uint32_t* s_beg,
                                           uint32_t* s_wsum) {
  const uint32_t ru = rank[u];
  const uint32_t b = uniform(off[u]), d = uniform(off[u + 1]) - b;
  for (uint32_t c0 = 0; c0 < d; c0 += kBlock) {
    const uint32_t i = c0 + threadIdx.x;
    uint32_t len = 0, vb = 0;
    if (i < d) {
      const uint32_t v = adj[b + i];
      if (rank[v] < ru) {
        vb = off[v];
        len = off[v + 1] - vb;
      }
    }
    // Block scan of the qualifying list lengths.
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
    const uint32_t incl0 = wave_incl_scan(len);
    if (lane == kWave - 1) s_wsum[w] = incl0;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t k = 0; k < w; ++k) add += s_wsum[k];
    s_pre[threadIdx.x] = incl0 + add;
    s_beg[threadIdx.x] = vb;
    __syncthreads();
    const uint32_t total = uniform(s_pre[kBlock - 1]);
    for (uint32_t f = threadIdx.x; f < total; f += kBlock) {
      const uint32_t j = upper_index(s_pre, kBlock, f);
      const uint32_t excl = j ? s_pre[j - 1] : 0u;
      const uint32_t x = adj[s_beg[j] + (f - excl)];
      if (x != u && rank[x] < ru) {
        if (clear)
          cnt[x] = 0u;
        else
          *acc += atomicAdd(&cnt[x], 1u);
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void cycles_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ rank, uint32_t n, unsigned int* __restrict__ queue,
    uint32_t* __restrict__ counters, uint64_t stride, unsigned long long* __restrict__ out) {
  __shared__ uint32_t s_pre[kBlock];
  __shared__ uint32_t s_beg[kBlock];
  __shared__ uint32_t s_wsum[kBlock / kWave];
  __shared__ uint32_t s_root;
  uint32_t* cnt = counters + (uint64_t)blockIdx.x * stride;
  uint64_t acc = 0;
  while (true) {
    if (threadIdx.x == 0) s_root = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t u = uniform(s_root);
    __syncthreads();
    if (u >= n) break;
    cycle_walk(off, adj, rank, u, cnt, false, &acc, s_pre, s_beg, s_wsum);
    // Counter updates of this root must be complete before the clearing
    // walk reads nothing but writes zeros; the fence orders them.
    __threadfence();
    __syncthreads();
    cycle_walk(off, adj, rank, u, cnt, true, &acc, s_pre, s_beg, s_wsum);
    __threadfence();
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if (lane_id() == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

__global__ void zero_u32_kernel(unsigned int* p, uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

int ensure_rank(gpm_graph* g) {
  if (g->rank) return GPM_OK;
  std::vector<uint32_t> order(g->n), rank(g->n);
  std::iota(order.begin(), order.end(), 0u);
  auto deg = [&](uint32_t v) { return g->h_off[v + 1] - g->h_off[v]; };
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return deg(a) < deg(b) || (deg(a) == deg(b) && a < b);
  });
  for (uint32_t i = 0; i < g->n; ++i) rank[order[i]] = i;
  if (hipMalloc(&g->rank, (size_t)std::max(g->n, 1u) * 4) != hipSuccess)
    return fail(GPM_ENOMEM, "rank array");
  int st = h2d(g, g->rank, rank.data(), (size_t)g->n * 4);
  if (st != GPM_OK) return st;
  return GPM_OK;
}

}  // namespace

namespace gpm {

int motif4_by_identities(gpm_graph* g, uint64_t out[6]) {
  int st;
  if ((st = ensure_src(g)) != GPM_OK) return st;
  if ((st = ensure_rank(g)) != GPM_OK) return st;
  if (!g->dag && (st = gpm_graph_orient(g, &g->dag)) != GPM_OK) return st;
  const int cus = device_cus(g);
  // Scratch: per-edge triangle counts, then block-private 4-cycle counters.
  const uint64_t tri_bytes = std::max<uint64_t>(g->m, 1) * 4;
  const uint64_t stride = std::max<uint32_t>(g->n, 1);
  unsigned blocks_c = (unsigned)std::min<uint64_t>((uint64_t)cus * 2, std::max(g->n, 1u));
  while (blocks_c > 16 && (uint64_t)blocks_c * stride * 4 > (4ull << 30)) blocks_c /= 2;
  const uint64_t cyc_bytes = (uint64_t)blocks_c * stride * 4;
  if ((st = ensure_scratch(g, tri_bytes + cyc_bytes + 256)) != GPM_OK) return st;
  uint32_t* tri = static_cast<uint32_t*>(g->scratch);
  uint32_t* ccnt = reinterpret_cast<uint32_t*>(static_cast<char*>(g->scratch) + tri_bytes);
  if ((st = timing_begin(g)) != GPM_OK) return st;
  if ((st = zero_counters(g, 16)) != GPM_OK) return st;
  unsigned int* queue = reinterpret_cast<unsigned int*>(g->counters + 8);
  const uint64_t zero_n = cyc_bytes / 4;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((zero_n + 255) / 256)), dim3(256), 0,
                     g->stream, ccnt, zero_n);
  const unsigned eblocks =
      (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((g->m * 8 + kBlock - 1) / kBlock,
                                                         (uint64_t)cus * 32));
  hipLaunchKernelGGL(edge_triangles_kernel<8>, dim3(eblocks), dim3(kBlock), 0, g->stream, g->off,
                     g->adj, g->src, g->m, tri);
  GPM_HIP(hipGetLastError());
  timing_mark(g, "motif_edge_triangles");
  const unsigned vblocks = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>(((uint64_t)g->n * kWave + kBlock - 1) / kBlock, (uint64_t)cus * 16));
  hipLaunchKernelGGL(vertex_sums_kernel, dim3(vblocks), dim3(kBlock), 0, g->stream, g->off,
                     g->adj, tri, g->n, g->counters);
  GPM_HIP(hipGetLastError());
  timing_mark(g, "motif_vertex_sums");
  hipLaunchKernelGGL(cycles_kernel, dim3(blocks_c), dim3(kBlock), 0, g->stream, g->off, g->adj,
                     g->rank, g->n, queue, ccnt, stride, g->counters + 5);
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
  const uint64_t n_paw = h[2];
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
