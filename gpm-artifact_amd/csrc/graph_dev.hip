// Device graph handle: creation (H2D of the CSR, label index), destruction,
// device-side orientation, timing and scratch management.
#include <algorithm>
#include <numeric>

#include "graph_dev.hpp"

using namespace gpm;

namespace {

__global__ void fill_u64_kernel(unsigned long long* p, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0ull;
}

// src[e] = u for every e in [off[u], off[u+1]); one wave per vertex.
__global__ void build_src_kernel(const uint32_t* __restrict__ off,
                                 uint32_t* __restrict__ src, uint32_t n) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const uint32_t lane = lane_id();
  for (uint32_t u = wave; u < n; u += nwaves) {
    const uint32_t b = off[u], e = off[u + 1];
    for (uint32_t i = b + lane; i < e; i += kWave) src[i] = u;
  }
}

// Orientation pass 1: out-degree under the rule of graph.h:128-129.
__global__ void orient_count_kernel(const uint32_t* __restrict__ off,
                                    const uint32_t* __restrict__ adj,
                                    uint32_t n, uint32_t* __restrict__ cnt) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const uint32_t lane = lane_id();
  for (uint32_t u = wave; u < n; u += nwaves) {
    const uint32_t b = off[u], e = off[u + 1], du = e - b;
    uint32_t c = 0;
    for (uint32_t i = b + lane; i < e; i += kWave) {
      const uint32_t v = adj[i];
      const uint32_t dv = off[v + 1] - off[v];
      c += (dv > du || (dv == du && v > u)) ? 1u : 0u;
    }
    c = wave_sum_u32(c);
    if (lane == 0) cnt[u] = c;
  }
}

// Orientation pass 2: stable compaction of each list (order kept).
__global__ void orient_fill_kernel(const uint32_t* __restrict__ off,
                                   const uint32_t* __restrict__ adj, uint32_t n,
                                   const uint32_t* __restrict__ new_off,
                                   uint32_t* __restrict__ new_adj) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const uint32_t lane = lane_id();
  for (uint32_t u = wave; u < n; u += nwaves) {
    const uint32_t b = off[u], e = off[u + 1], du = e - b;
    uint32_t pos = new_off[u];
    for (uint32_t i0 = b; i0 < e; i0 += kWave) {
      const uint32_t i = i0 + lane;
      bool keep = false;
      uint32_t v = 0;
      if (i < e) {
        v = adj[i];
        const uint32_t dv = off[v + 1] - off[v];
        keep = dv > du || (dv == du && v > u);
      }
      const uint64_t mask = wave_ballot(keep);
      if (keep) new_adj[pos + lanes_below(mask)] = v;
      pos += (uint32_t)__popcll(mask);
    }
  }
}

int grid_for(const gpm_graph* g, uint64_t items, int block) {
  int cus = device_cus(g);
  uint64_t want = (items + block - 1) / block;
  uint64_t cap = (uint64_t)cus * 16;
  return (int)std::max<uint64_t>(1, std::min(want, cap));
}

}  // namespace

namespace gpm {

int device_cus(const gpm_graph* g) {
  static thread_local int cached_dev = -1, cached_cus = 0;
  if (cached_dev != g->device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, g->device) != hipSuccess) return 256;
    cached_dev = g->device;
    cached_cus = prop.multiProcessorCount;
  }
  return cached_cus;
}

int h2d(gpm_graph* g, void* dst, const void* src, size_t bytes) {
  if (!bytes) return GPM_OK;
  GPM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g->stream));
  GPM_HIP(hipStreamSynchronize(g->stream));
  return GPM_OK;
}

int ensure_topo(gpm_graph* g) {
  if (g->toff) return GPM_OK;
  std::vector<uint32_t> h_adj(std::max<uint64_t>(g->m, 1));
  if (g->m) {
    GPM_HIP(hipMemcpyAsync(h_adj.data(), g->adj, g->m * 4, hipMemcpyDeviceToHost, g->stream));
    GPM_HIP(hipStreamSynchronize(g->stream));
  }
  std::vector<uint32_t> rank, noff, nadj;
  if (!topo_relabel(g->n, g->h_off.data(), h_adj.data(), rank, noff, nadj))
    return fail(GPM_EINVAL, "the ORIENTED graph has a cycle: k-clique counting needs an acyclic "
                            "orientation");
  uint32_t *toff = nullptr, *tadj = nullptr;
  if (hipMalloc(&toff, noff.size() * 4) != hipSuccess ||
      hipMalloc(&tadj, std::max<size_t>(nadj.size(), 1) * 4) != hipSuccess) {
    (void)hipFree(toff);
    return fail(GPM_ENOMEM, "topological copy of the DAG");
  }
  int st = h2d(g, toff, noff.data(), noff.size() * 4);
  if (st == GPM_OK) st = h2d(g, tadj, nadj.data(), nadj.size() * 4);
  if (st != GPM_OK) {
    (void)hipFree(toff);
    (void)hipFree(tadj);
    return st;
  }
  g->toff = toff;
  g->tadj = tadj;
  g->h_rank.swap(rank);
  return GPM_OK;
}

int ensure_scratch(gpm_graph* g, size_t bytes) {
  if (g->scratch_bytes >= bytes) return GPM_OK;
  if (g->scratch) {
    GPM_HIP(hipStreamSynchronize(g->stream));
    GPM_HIP(hipFree(g->scratch));
    g->scratch = nullptr;
    g->scratch_bytes = 0;
  }
  size_t want = std::max(bytes, (size_t)1 << 20);
  if (hipMalloc(&g->scratch, want) != hipSuccess) {
    g->scratch = nullptr;
    (void)hipGetLastError();  // clear the sticky error: callers may retry smaller
    return fail(GPM_ENOMEM, "device scratch of %zu bytes", want);
  }
  g->scratch_bytes = want;
  return GPM_OK;
}

int ensure_src(gpm_graph* g) {
  if (g->src || g->m == 0) return GPM_OK;
  if (hipMalloc(&g->src, g->m * sizeof(uint32_t)) != hipSuccess)
    return fail(GPM_ENOMEM, "device COO source array");
  hipLaunchKernelGGL(build_src_kernel, dim3(grid_for(g, (uint64_t)g->n * kWave, 256)),
                     dim3(256), 0, g->stream, g->off, g->src, g->n);
  GPM_HIP(hipGetLastError());
  GPM_HIP(hipStreamSynchronize(g->stream));
  return GPM_OK;
}

int zero_counters(gpm_graph* g, int count) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3((count + 63) / 64), dim3(64), 0,
                     g->stream, g->counters, count);
  GPM_HIP(hipGetLastError());
  return GPM_OK;
}

int timing_begin(gpm_graph* g) {
  g->n_marks = 0;
  GPM_HIP(hipEventRecord(g->ev0, g->stream));
  return GPM_OK;
}

int timing_mark(gpm_graph* g, const char* name) {
  if (g->n_marks >= gpm_graph::kMaxMarks) return GPM_OK;
  hipEvent_t& e = g->mark_ev[g->n_marks];
  if (!e) GPM_HIP(hipEventCreate(&e));
  GPM_HIP(hipEventRecord(e, g->stream));
  g->mark_name[g->n_marks++] = name;
  return GPM_OK;
}

int timing_end(gpm_graph* g) {
  GPM_HIP(hipEventRecord(g->ev1, g->stream));
  GPM_HIP(hipEventSynchronize(g->ev1));
  float ms = 0.f;
  GPM_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  g->last_ms = ms;
  for (int i = 0; i < g->n_marks; ++i) {
    float mi = 0.f;
    GPM_HIP(hipEventElapsedTime(&mi, i ? g->mark_ev[i - 1] : g->ev0, g->mark_ev[i]));
    g->mark_ms[i] = mi;
  }
  return GPM_OK;
}

}  // namespace gpm

namespace {

void release(gpm_graph* g) {
  if (!g) return;
  if (g->stream) hipStreamSynchronize(g->stream);
  free_match_cache(g);
  if (g->tc_exec) (void)hipGraphExecDestroy(g->tc_exec);
  if (g->tc_graph) (void)hipGraphDestroy(g->tc_graph);
  if (g->dag) gpm_graph_destroy(g->dag);
  (void)hipFree(g->rk_off);
  (void)hipFree(g->rk_adj);
  free_tiers(g);
  hipFree(g->off);
  hipFree(g->adj);
  hipFree(g->src);
  hipFree(g->labels);
  hipFree(g->ladj);
  hipFree(g->loff);
  hipFree(g->lverts);
  (void)hipFree(g->sig);
  (void)hipFree(g->toff);
  (void)hipFree(g->tadj);
  (void)hipFree(g->bmask);
  hipFree(g->scratch);
  hipFree(g->counters);
  (void)hipFree(g->tc_partial);
  for (auto& e : g->mark_ev)
    if (e) (void)hipEventDestroy(e);
  if (g->ev0) hipEventDestroy(g->ev0);
  if (g->ev1) hipEventDestroy(g->ev1);
  if (g->stream) hipStreamDestroy(g->stream);
  delete g;
}

// Builds the label-partitioned adjacency on the host (sorted by (label, id))
// and uploads it.
int build_label_index(gpm_graph* g, const uint32_t* offsets, const uint32_t* adj,
                      const uint32_t* labels) {
  std::vector<uint32_t> ladj, loff;
  build_label_csr(g->n, g->n_labels, offsets, adj, labels, ladj, loff);
  if (hipMalloc(&g->ladj, std::max<size_t>(g->m, 1) * 4) != hipSuccess ||
      hipMalloc(&g->loff, loff.size() * 4) != hipSuccess)
    return fail(GPM_ENOMEM, "device label index");
  GPM_HIP(hipMemcpyAsync(g->ladj, ladj.data(), g->m * 4, hipMemcpyHostToDevice, g->stream));
  GPM_HIP(hipMemcpyAsync(g->loff, loff.data(), loff.size() * 4, hipMemcpyHostToDevice,
                         g->stream));
  // Vertices grouped by label (counting sort, ids ascending inside a label).
  const uint32_t L = g->n_labels;
  if (L <= 1024) {
    g->h_lpair.assign((size_t)L * L, 0.0);
    for (uint32_t v = 0; v < g->n; ++v) {
      const uint32_t* lo = loff.data() + (size_t)v * (L + 1);
      double* row = g->h_lpair.data() + (size_t)labels[v] * L;
      for (uint32_t l = 0; l < L; ++l) row[l] += lo[l + 1] - lo[l];
    }
  }
  // Packed filter signature for small label alphabets: degree in the high
  // word, per label the neighbour count saturated at 15 in 4 bits.
  if (L <= kSigLabels) {
    std::vector<unsigned long long> sig(std::max(g->n, 1u), 0ull);
    for (uint32_t v = 0; v < g->n; ++v) {
      const uint32_t* lo = loff.data() + (size_t)v * (L + 1);
      unsigned long long x = (unsigned long long)(offsets[v + 1] - offsets[v]) << 32;
      for (uint32_t l = 0; l < L; ++l)
        x |= (unsigned long long)std::min<uint32_t>(15u, lo[l + 1] - lo[l]) << (4 * l);
      sig[v] = x;
    }
    if (hipMalloc(&g->sig, sig.size() * 8) != hipSuccess) return fail(GPM_ENOMEM, "signatures");
    GPM_HIP(hipMemcpyAsync(g->sig, sig.data(), sig.size() * 8, hipMemcpyHostToDevice, g->stream));
    GPM_HIP(hipStreamSynchronize(g->stream));
  }
  g->h_lvoff.assign(L + 1, 0);
  for (uint32_t v = 0; v < g->n; ++v) g->h_lvoff[labels[v] + 1]++;
  for (uint32_t l = 0; l < L; ++l) g->h_lvoff[l + 1] += g->h_lvoff[l];
  std::vector<uint32_t> lverts(std::max(g->n, 1u)), cur(g->h_lvoff.begin(), g->h_lvoff.end() - 1);
  for (uint32_t v = 0; v < g->n; ++v) lverts[cur[labels[v]]++] = v;
  if (hipMalloc(&g->lverts, lverts.size() * 4) != hipSuccess)
    return fail(GPM_ENOMEM, "device label vertex lists");
  GPM_HIP(hipMemcpyAsync(g->lverts, lverts.data(), (size_t)g->n * 4, hipMemcpyHostToDevice,
                         g->stream));
  GPM_HIP(hipStreamSynchronize(g->stream));
  return GPM_OK;
}

}  // namespace

namespace gpm {

int adopt_device_csr(uint32_t n, uint64_t m, uint32_t* d_off, uint32_t* d_adj,
                     std::vector<uint32_t>&& h_off, const uint32_t* h_adj, const uint32_t* labels,
                     uint32_t n_labels, uint32_t flags, hipStream_t stream, gpm_graph** out) {
  *out = nullptr;
  int dev = 0;
  GPM_HIP(hipGetDevice(&dev));
  gpm_graph* g = new gpm_graph();
  g->device = dev;
  g->n = n;
  g->m = m;
  g->flags = flags;
  g->n_labels = labels ? n_labels : 0;
  g->off = d_off;
  g->adj = d_adj;
  g->stream = stream;
  for (uint32_t v = 0; v < n; ++v) g->max_deg = std::max(g->max_deg, h_off[v + 1] - h_off[v]);
  g->h_off = std::move(h_off);
  if (labels) g->h_labels.assign(labels, labels + n);
  int st = GPM_OK;
  do {
    if (hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess) {
      st = fail(GPM_EHIP, "event creation failed");
      break;
    }
    if (hipMalloc(&g->counters, 64 * sizeof(unsigned long long)) != hipSuccess) {
      st = fail(GPM_ENOMEM, "device counters");
      break;
    }
    if (labels) {
      if (hipMalloc(&g->labels, (size_t)std::max(n, 1u) * 4) != hipSuccess) {
        st = fail(GPM_ENOMEM, "device labels");
        break;
      }
      if ((st = h2d(g, g->labels, labels, (size_t)n * 4)) != GPM_OK) break;
      if ((st = build_label_index(g, g->h_off.data(), h_adj, labels)) != GPM_OK) break;
    }
    if (hipStreamSynchronize(g->stream) != hipSuccess) st = fail(GPM_EHIP, "stream synchronize");
  } while (0);
  if (st != GPM_OK) {
    release(g);  // frees d_off, d_adj and the stream with the rest
    return st;
  }
  *out = g;
  return GPM_OK;
}

}  // namespace gpm

extern "C" {

int gpm_graph_create(const uint32_t* offsets, uint32_t n, const uint32_t* adj,
                     uint64_t m, const uint32_t* labels, uint32_t n_labels,
                     uint32_t flags, gpm_graph** out) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(out && offsets, "null pointer");
    GPM_CHECK_ARG(m == 0 || adj, "null adjacency");
    GPM_CHECK_ARG(n < 0xFFFFFFFFu, "n must be < 2^32 - 1");
    GPM_CHECK_ARG(m < (1ull << 32), "m must be < 2^32");
    GPM_CHECK_ARG(offsets[0] == 0 && offsets[n] == m, "offsets[0] must be 0 and offsets[n] == m");
    GPM_CHECK_ARG(!labels || n_labels >= 1, "labels given with n_labels == 0");
    GPM_CHECK_ARG(n_labels <= 4096, "n_labels must be <= 4096");
    *out = nullptr;
    // Host-side validation so that no kernel can index out of range.
    uint32_t maxd = 0;
    for (uint32_t v = 0; v < n; ++v) {
      if (offsets[v + 1] < offsets[v]) return fail(GPM_EINVAL, "offsets not monotone at %u", v);
      maxd = std::max(maxd, offsets[v + 1] - offsets[v]);
    }
    for (uint64_t i = 0; i < m; ++i)
      if (adj[i] >= n) return fail(GPM_EINVAL, "adjacency entry %llu = %u >= n",
                                   (unsigned long long)i, adj[i]);
    if (labels)
      for (uint32_t v = 0; v < n; ++v)
        if (labels[v] >= n_labels) return fail(GPM_EINVAL, "label of %u >= n_labels", v);
    int dev = 0;
    GPM_HIP(hipGetDevice(&dev));
    gpm_graph* g = new gpm_graph();
    g->device = dev;
    g->n = n;
    g->m = m;
    g->max_deg = maxd;
    g->flags = flags;
    g->n_labels = labels ? n_labels : 0;
    g->h_off.assign(offsets, offsets + n + 1);
    if (labels) g->h_labels.assign(labels, labels + n);
    int st = GPM_OK;
    do {
      if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess) {
        st = fail(GPM_EHIP, "stream/event creation failed");
        break;
      }
      if (hipMalloc(&g->off, (size_t)(n + 1) * 4) != hipSuccess ||
          hipMalloc(&g->adj, std::max<uint64_t>(m, 1) * 4) != hipSuccess ||
          hipMalloc(&g->counters, 64 * sizeof(unsigned long long)) != hipSuccess) {
        st = fail(GPM_ENOMEM, "device CSR allocation (n=%u, m=%llu)", n, (unsigned long long)m);
        break;
      }
      if (hipMemcpyAsync(g->off, offsets, (size_t)(n + 1) * 4, hipMemcpyHostToDevice,
                         g->stream) != hipSuccess ||
          (m && hipMemcpyAsync(g->adj, adj, m * 4, hipMemcpyHostToDevice, g->stream) !=
                    hipSuccess)) {
        st = fail(GPM_EHIP, "H2D copy of the CSR failed");
        break;
      }
      if (labels) {
        if (hipMalloc(&g->labels, (size_t)std::max(n, 1u) * 4) != hipSuccess) {
          st = fail(GPM_ENOMEM, "device labels");
          break;
        }
        if (hipMemcpyAsync(g->labels, labels, (size_t)n * 4, hipMemcpyHostToDevice,
                           g->stream) != hipSuccess) {
          st = fail(GPM_EHIP, "H2D copy of labels failed");
          break;
        }
        st = build_label_index(g, offsets, adj, labels);
        if (st != GPM_OK) break;
      }
      if (hipStreamSynchronize(g->stream) != hipSuccess) {
        st = fail(GPM_EHIP, "stream synchronize failed");
        break;
      }
    } while (0);
    if (st != GPM_OK) {
      release(g);
      return st;
    }
    *out = g;
    return GPM_OK;
  });
}

int gpm_graph_destroy(gpm_graph* g) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    release(g);
    return GPM_OK;
  });
}

int gpm_graph_info(const gpm_graph* g, uint32_t* n, uint64_t* m, uint32_t* max_deg) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g, "null graph");
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (max_deg) *max_deg = g->max_deg;
    return GPM_OK;
  });
}

int gpm_last_kernel_ms(const gpm_graph* g, double* ms) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g && ms, "null pointer");
    *ms = g->last_ms;
    return GPM_OK;
  });
}

int gpm_last_kernel_breakdown(const gpm_graph* g, double* ms, const char** names,
                              int capacity, int* count) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g && count, "null pointer");
    const int n = std::min(capacity, g->n_marks);
    for (int i = 0; i < n; ++i) {
      if (ms) ms[i] = g->mark_ms[i];
      if (names) names[i] = g->mark_name[i];
    }
    *count = g->n_marks;
    return GPM_OK;
  });
}

int gpm_graph_orient(gpm_graph* sym, gpm_graph** dag) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(sym && dag, "null pointer");
    GPM_CHECK_ARG(sym->flags & GPM_GRAPH_SYMMETRIC, "graph is not SYMMETRIC");
    const uint32_t n = sym->n;
    uint32_t* cnt = nullptr;
    GPM_HIP(hipMalloc(&cnt, (size_t)(n + 1) * 4));
    const int grid = grid_for(sym, (uint64_t)n * kWave, 256);
    int st = timing_begin(sym);
    if (st != GPM_OK) {
      hipFree(cnt);
      return st;
    }
    hipLaunchKernelGGL(orient_count_kernel, dim3(grid), dim3(256), 0, sym->stream, sym->off,
                       sym->adj, n, cnt);
    std::vector<uint32_t> h_cnt(n + 1, 0);
    if (hipMemcpyAsync(h_cnt.data(), cnt, (size_t)n * 4, hipMemcpyDeviceToHost, sym->stream) !=
            hipSuccess ||
        hipStreamSynchronize(sym->stream) != hipSuccess) {
      hipFree(cnt);
      return fail(GPM_EHIP, "orientation count pass failed");
    }
    // Exclusive scan on the host (n+1 entries); the list pass is on the GPU.
    std::vector<uint32_t> h_new(n + 1, 0);
    uint32_t maxd = 0;
    for (uint32_t v = 0; v < n; ++v) {
      h_new[v + 1] = h_new[v] + h_cnt[v];
      maxd = std::max(maxd, h_cnt[v]);
    }
    const uint64_t m2 = h_new[n];
    gpm_graph* g = new gpm_graph();
    g->device = sym->device;
    g->n = n;
    g->m = m2;
    g->max_deg = maxd;
    g->flags = GPM_GRAPH_ORIENTED;
    g->h_off = h_new;
    do {
      if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess) {
        st = fail(GPM_EHIP, "stream/event creation failed");
        break;
      }
      if (hipMalloc(&g->off, (size_t)(n + 1) * 4) != hipSuccess ||
          hipMalloc(&g->adj, std::max<uint64_t>(m2, 1) * 4) != hipSuccess ||
          hipMalloc(&g->counters, 64 * sizeof(unsigned long long)) != hipSuccess) {
        st = fail(GPM_ENOMEM, "device CSR allocation");
        break;
      }
      if (hipMemcpyAsync(g->off, h_new.data(), (size_t)(n + 1) * 4, hipMemcpyHostToDevice,
                         sym->stream) != hipSuccess) {
        st = fail(GPM_EHIP, "H2D of offsets failed");
        break;
      }
      hipLaunchKernelGGL(orient_fill_kernel, dim3(grid), dim3(256), 0, sym->stream, sym->off,
                         sym->adj, n, g->off, g->adj);
      if (hipGetLastError() != hipSuccess) {
        st = fail(GPM_EHIP, "orientation fill launch failed");
        break;
      }
      st = timing_end(sym);
    } while (0);
    hipFree(cnt);
    if (st != GPM_OK) {
      release(g);
      return st;
    }
    *dag = g;
    return GPM_OK;
  });
}

}  // extern "C"
