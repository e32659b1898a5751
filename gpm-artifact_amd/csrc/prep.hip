// Graph preparation on the device: CSR build from an edge list (symmetrise,
// drop self loops, sort, dedup) and the degree relabel, replacing the host
// passes of Graph::Init (src/graph/graph.h:254-340, with the symmetrise /
// dedup the configurations' inputs need) and Graph::SortCSRGraph
// (graph.h:206-252).  Orientation already runs on the device
// (gpm_graph_orient).
//
// CSR build: every kept edge becomes one or two records (u, v); sorted mode
// (GPM_CSR_SORT or GPM_CSR_DEDUP) radix-sorts 64-bit keys u << 32 | v, so
// lists come out sorted, and DEDUP keeps the first of equal keys; input-order
// mode radix-sorts the records by u alone (LSD radix sort is stable), so
// every list keeps the order of the edge list, the reverse record of an edge
// right after the forward one -- the order Graph::Init and
// gpm_csr_from_edges produce.  Dropped self loops get the source key n and
// sort past every real record.  Offsets come from the sorted sources in one
// pass (entry i starts the lists of vertices src[i-1]+1 .. src[i]).
#include <algorithm>
#include <numeric>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "graph_dev.hpp"

using namespace gpm;

namespace {

constexpr int kBlock = 256;

unsigned grid_for_items(uint64_t items, int cus) {
  return (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((items + kBlock - 1) / kBlock, (uint64_t)cus * 32));
}

int bits_for(uint64_t max_value) {
  int b = 1;
  while (b < 64 && (max_value >> b)) ++b;
  return b;
}

// Records of edge i: (u, v) and, when symmetrising, (v, u) right after it.
__global__ void emit_keys_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                 uint64_t n_edges, uint32_t n, bool sym, bool drop_self,
                                 unsigned long long* __restrict__ keys) {
  const uint32_t per = sym ? 2u : 1u;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_edges;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = src[i], v = dst[i];
    const bool drop = drop_self && u == v;
    const unsigned long long sentinel = (unsigned long long)n << 32;
    keys[per * i] = drop ? sentinel : ((unsigned long long)u << 32 | v);
    if (sym) keys[per * i + 1] = drop ? sentinel : ((unsigned long long)v << 32 | u);
  }
}

// Input-order mode: source keys (n for dropped records) and target values.
__global__ void emit_pairs_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                  uint64_t n_edges, uint32_t n, bool sym, bool drop_self,
                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t per = sym ? 2u : 1u;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_edges;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = src[i], v = dst[i];
    const bool drop = drop_self && u == v;
    keys[per * i] = drop ? n : u;
    vals[per * i] = v;
    if (sym) {
      keys[per * i + 1] = drop ? n : v;
      vals[per * i + 1] = u;
    }
  }
}

__global__ void split_keys_kernel(const unsigned long long* __restrict__ keys, uint64_t m,
                                  uint32_t* __restrict__ hi, uint32_t* __restrict__ lo) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = keys[i];
    hi[i] = (uint32_t)(k >> 32);
    lo[i] = (uint32_t)k;
  }
}

// off[v] = first record with source >= v, for v in [0, n]; src holds the
// sorted sources of the m kept records.
__global__ void offsets_kernel(const uint32_t* __restrict__ src, uint64_t m, uint32_t n,
                               uint32_t* __restrict__ off) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t first = i == 0 ? 0 : (uint64_t)src[i - 1] + 1;  // vertices starting here
    const uint64_t last = i == m ? n : src[i];
    for (uint64_t v = first; v <= last; ++v) off[v] = (uint32_t)i;
  }
}

// Degree relabel: sort keys (max_deg - deg(v)) with values v, then
// new_id[order[r]] = r.
__global__ void degree_keys_kernel(const uint32_t* __restrict__ off, uint32_t n, uint32_t max_deg,
                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    keys[v] = max_deg - (off[v + 1] - off[v]);
    ids[v] = v;
  }
}

__global__ void scatter_rank_kernel(const uint32_t* __restrict__ order, uint32_t n,
                                    uint32_t* __restrict__ new_id) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    new_id[order[r]] = r;
}

// Keys new_id[u] << 32 | new_id[x] for every entry (u, x); one wave per u.
__global__ void relabel_keys_kernel(const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                    const uint32_t* __restrict__ new_id, uint32_t n,
                                    unsigned long long* __restrict__ keys) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  for (uint32_t u = wave; u < n; u += nwaves) {
    const unsigned long long hi = (unsigned long long)new_id[u] << 32;
    for (uint32_t i = off[u] + lane_id(); i < off[u + 1]; i += kWave) keys[i] = hi | new_id[adj[i]];
  }
}

// Device buffers of one preparation call, freed on every exit path unless
// handed to the new graph.
struct Bufs {
  std::vector<void*> owned;
  template <typename T>
  int alloc(T** p, uint64_t count) {
    *p = nullptr;
    if (hipMalloc(p, std::max<uint64_t>(count, 1) * sizeof(T)) != hipSuccess) {
      *p = nullptr;
      return fail(GPM_ENOMEM, "device buffer of %llu bytes",
                  (unsigned long long)(count * sizeof(T)));
    }
    owned.push_back(*p);
    return GPM_OK;
  }
  void release(void* p) { owned.erase(std::remove(owned.begin(), owned.end(), p), owned.end()); }
  ~Bufs() {
    for (void* p : owned) (void)hipFree(p);
  }
};

// Sorted 64-bit keys (u << 32 | v, the m_kept first valid) -> CSR in new
// buffers *d_off, *d_adj (owned by `b` until released).
int keys_to_csr(Bufs& b, const unsigned long long* keys, uint64_t m, uint32_t n, int cus,
                hipStream_t stream, uint32_t** d_off, uint32_t** d_adj) {
  int st;
  uint32_t* d_src = nullptr;
  if ((st = b.alloc(&d_src, m)) != GPM_OK || (st = b.alloc(d_adj, m)) != GPM_OK ||
      (st = b.alloc(d_off, (uint64_t)n + 1)) != GPM_OK)
    return st;
  if (m)
    hipLaunchKernelGGL(split_keys_kernel, dim3(grid_for_items(m, cus)), dim3(kBlock), 0, stream,
                       keys, m, d_src, *d_adj);
  hipLaunchKernelGGL(offsets_kernel, dim3(grid_for_items(m + 1, cus)), dim3(kBlock), 0, stream,
                     d_src, m, n, *d_off);
  GPM_HIP(hipGetLastError());
  return GPM_OK;
}

int build(uint32_t n, const uint32_t* src, const uint32_t* dst, uint64_t n_edges, uint32_t csr_flags,
          const uint32_t* labels, uint32_t n_labels, uint32_t graph_flags, gpm_graph** out) {
  GPM_CHECK_ARG(out, "null pointer");
  GPM_CHECK_ARG(n_edges == 0 || (src && dst), "null edge arrays");
  GPM_CHECK_ARG(n < 0xFFFFFFFFu, "n must be < 2^32 - 1");
  GPM_CHECK_ARG(!labels || (n_labels >= 1 && n_labels <= 4096), "n_labels must be in [1, 4096]");
  uint64_t self_loops = 0;
  for (uint64_t i = 0; i < n_edges; ++i) {
    if (src[i] >= n || dst[i] >= n) return fail(GPM_EINVAL, "edge endpoint >= n (%u)", n);
    self_loops += src[i] == dst[i];
  }
  if (labels)
    for (uint32_t v = 0; v < n; ++v)
      if (labels[v] >= n_labels) return fail(GPM_EINVAL, "label of %u >= n_labels", v);
  const bool sym = csr_flags & GPM_CSR_SYMMETRIZE;
  const bool dedup = csr_flags & GPM_CSR_DEDUP;
  const bool drop_self = csr_flags & GPM_CSR_DROP_SELF;
  const bool sorted = (csr_flags & GPM_CSR_SORT) || dedup;
  const uint64_t records = n_edges * (sym ? 2 : 1);
  // Dropped self loops become sentinel records that sort past the real ones.
  const uint64_t dropped = drop_self ? self_loops * (sym ? 2 : 1) : 0;
  GPM_CHECK_ARG(records < (1ull << 32), "too many adjacency records (%llu)",
                (unsigned long long)records);
  int dev = 0, cus = 256;
  GPM_HIP(hipGetDevice(&dev));
  GPM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  hipStream_t stream = nullptr;
  GPM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t* s;
    ~StreamGuard() {
      if (*s) {
        (void)hipStreamSynchronize(*s);
        (void)hipStreamDestroy(*s);
      }
    }
  } guard{&stream};
  Bufs b;
  int st;
  uint32_t *d_src = nullptr, *d_dst = nullptr;
  if ((st = b.alloc(&d_src, n_edges)) != GPM_OK || (st = b.alloc(&d_dst, n_edges)) != GPM_OK)
    return st;
  if (n_edges) {
    GPM_HIP(hipMemcpyAsync(d_src, src, n_edges * 4, hipMemcpyHostToDevice, stream));
    GPM_HIP(hipMemcpyAsync(d_dst, dst, n_edges * 4, hipMemcpyHostToDevice, stream));
  }
  const int vbits = bits_for(n);  // source keys go up to n (the drop sentinel)
  uint32_t *d_off = nullptr, *d_adj = nullptr;
  uint64_t m = 0;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  if (sorted) {
    unsigned long long *keys = nullptr, *sorted_keys = nullptr;
    if ((st = b.alloc(&keys, records)) != GPM_OK || (st = b.alloc(&sorted_keys, records)) != GPM_OK)
      return st;
    if (records) {
      hipLaunchKernelGGL(emit_keys_kernel, dim3(grid_for_items(n_edges, cus)), dim3(kBlock), 0,
                         stream, d_src, d_dst, n_edges, n, sym, drop_self, keys);
      GPM_HIP(hipGetLastError());
      const unsigned end_bit = 32u + (unsigned)vbits;
      GPM_HIP(rocprim::radix_sort_keys(nullptr, temp_bytes, keys, sorted_keys, records, 0u,
                                       end_bit, stream));
      if ((st = b.alloc(reinterpret_cast<char**>(&temp), temp_bytes)) != GPM_OK) return st;
      GPM_HIP(rocprim::radix_sort_keys(temp, temp_bytes, keys, sorted_keys, records, 0u, end_bit,
                                       stream));
    }
    unsigned long long* final_keys = sorted_keys;
    uint64_t kept = records;
    if (dedup && records) {
      // Unique keys into `keys` (reused), count on the device.
      unsigned long long* d_count = nullptr;
      if ((st = b.alloc(&d_count, 1)) != GPM_OK) return st;
      size_t ubytes = 0;
      GPM_HIP(rocprim::unique(nullptr, ubytes, sorted_keys, keys, d_count, records,
                              rocprim::equal_to<unsigned long long>(), stream));
      void* utemp = nullptr;
      if ((st = b.alloc(reinterpret_cast<char**>(&utemp), ubytes)) != GPM_OK) return st;
      GPM_HIP(rocprim::unique(utemp, ubytes, sorted_keys, keys, d_count, records,
                              rocprim::equal_to<unsigned long long>(), stream));
      unsigned long long h_count = 0;
      GPM_HIP(hipMemcpyAsync(&h_count, d_count, 8, hipMemcpyDeviceToHost, stream));
      GPM_HIP(hipStreamSynchronize(stream));
      kept = h_count;
      final_keys = keys;
    }
    // Dedup keeps one sentinel record when any was dropped.
    m = dedup ? kept - (dropped ? 1 : 0) : records - dropped;
    if ((st = keys_to_csr(b, final_keys, m, n, cus, stream, &d_off, &d_adj)) != GPM_OK) return st;
  } else {
    uint32_t *keys = nullptr, *vals = nullptr, *skeys = nullptr;
    if ((st = b.alloc(&keys, records)) != GPM_OK || (st = b.alloc(&vals, records)) != GPM_OK ||
        (st = b.alloc(&skeys, records)) != GPM_OK || (st = b.alloc(&d_adj, records)) != GPM_OK)
      return st;
    if (records) {
      hipLaunchKernelGGL(emit_pairs_kernel, dim3(grid_for_items(n_edges, cus)), dim3(kBlock), 0,
                         stream, d_src, d_dst, n_edges, n, sym, drop_self, keys, vals);
      GPM_HIP(hipGetLastError());
      GPM_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys, skeys, vals, d_adj, records, 0u,
                                        (unsigned)vbits, stream));
      if ((st = b.alloc(reinterpret_cast<char**>(&temp), temp_bytes)) != GPM_OK) return st;
      GPM_HIP(rocprim::radix_sort_pairs(temp, temp_bytes, keys, skeys, vals, d_adj, records, 0u,
                                        (unsigned)vbits, stream));
    }
    m = records - dropped;
    if ((st = b.alloc(&d_off, (uint64_t)n + 1)) != GPM_OK) return st;
    hipLaunchKernelGGL(offsets_kernel, dim3(grid_for_items(m + 1, cus)), dim3(kBlock), 0, stream,
                       skeys, m, n, d_off);
    GPM_HIP(hipGetLastError());
  }
  std::vector<uint32_t> h_off((size_t)n + 1);
  GPM_HIP(hipMemcpyAsync(h_off.data(), d_off, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, stream));
  std::vector<uint32_t> h_adj;
  if (labels) {
    h_adj.resize(std::max<uint64_t>(m, 1));
    if (m) GPM_HIP(hipMemcpyAsync(h_adj.data(), d_adj, m * 4, hipMemcpyDeviceToHost, stream));
  }
  GPM_HIP(hipStreamSynchronize(stream));
  b.release(d_off);
  b.release(d_adj);
  hipStream_t s = stream;
  stream = nullptr;  // the handle owns it now
  return adopt_device_csr(n, m, d_off, d_adj, std::move(h_off), h_adj.data(), labels, n_labels,
                          graph_flags, s, out);
}

int relabel(gpm_graph* g, gpm_graph** out, uint32_t** perm) {
  GPM_CHECK_ARG(g && out && perm, "null pointer");
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  int cus = device_cus(g);
  hipStream_t stream = nullptr;
  GPM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t* s;
    ~StreamGuard() {
      if (*s) {
        (void)hipStreamSynchronize(*s);
        (void)hipStreamDestroy(*s);
      }
    }
  } guard{&stream};
  // The source graph's pending work must be done before this stream reads it.
  GPM_HIP(hipStreamSynchronize(g->stream));
  Bufs b;
  int st;
  uint32_t *dkeys = nullptr, *ids = nullptr, *skeys = nullptr, *order = nullptr, *new_id = nullptr;
  if ((st = b.alloc(&dkeys, n)) != GPM_OK || (st = b.alloc(&ids, n)) != GPM_OK ||
      (st = b.alloc(&skeys, n)) != GPM_OK || (st = b.alloc(&order, n)) != GPM_OK ||
      (st = b.alloc(&new_id, n)) != GPM_OK)
    return st;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  if (n) {
    hipLaunchKernelGGL(degree_keys_kernel, dim3(grid_for_items(n, cus)), dim3(kBlock), 0, stream,
                       g->off, n, g->max_deg, dkeys, ids);
    GPM_HIP(hipGetLastError());
    const unsigned end_bit = (unsigned)bits_for(g->max_deg);
    GPM_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, dkeys, skeys, ids, order, n, 0u,
                                      end_bit, stream));
    if ((st = b.alloc(reinterpret_cast<char**>(&temp), temp_bytes)) != GPM_OK) return st;
    GPM_HIP(rocprim::radix_sort_pairs(temp, temp_bytes, dkeys, skeys, ids, order, n, 0u, end_bit,
                                      stream));
    hipLaunchKernelGGL(scatter_rank_kernel, dim3(grid_for_items(n, cus)), dim3(kBlock), 0, stream,
                       order, n, new_id);
    GPM_HIP(hipGetLastError());
  }
  unsigned long long *keys = nullptr, *sorted_keys = nullptr;
  if ((st = b.alloc(&keys, m)) != GPM_OK || (st = b.alloc(&sorted_keys, m)) != GPM_OK) return st;
  if (m) {
    const unsigned blocks = grid_for_items((uint64_t)n * kWave, cus);
    hipLaunchKernelGGL(relabel_keys_kernel, dim3(blocks), dim3(kBlock), 0, stream, g->off, g->adj,
                       new_id, n, keys);
    GPM_HIP(hipGetLastError());
    const unsigned end_bit = 32u + (unsigned)bits_for(n);
    size_t tb = 0;
    GPM_HIP(rocprim::radix_sort_keys(nullptr, tb, keys, sorted_keys, m, 0u, end_bit, stream));
    void* t2 = nullptr;
    if ((st = b.alloc(reinterpret_cast<char**>(&t2), tb)) != GPM_OK) return st;
    GPM_HIP(rocprim::radix_sort_keys(t2, tb, keys, sorted_keys, m, 0u, end_bit, stream));
  }
  uint32_t *d_off = nullptr, *d_adj = nullptr;
  if ((st = keys_to_csr(b, sorted_keys, m, n, cus, stream, &d_off, &d_adj)) != GPM_OK) return st;
  std::vector<uint32_t> h_off((size_t)n + 1), h_new(std::max(n, 1u));
  GPM_HIP(hipMemcpyAsync(h_off.data(), d_off, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, stream));
  if (n) GPM_HIP(hipMemcpyAsync(h_new.data(), new_id, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
  std::vector<uint32_t> h_adj;
  std::vector<uint32_t> nlab;
  if (g->labels) {
    h_adj.resize(std::max<uint64_t>(m, 1));
    if (m) GPM_HIP(hipMemcpyAsync(h_adj.data(), d_adj, m * 4, hipMemcpyDeviceToHost, stream));
    nlab.resize(n);
    for (uint32_t v = 0; v < n; ++v) nlab[h_new[v]] = g->h_labels[v];
  }
  GPM_HIP(hipStreamSynchronize(stream));
  uint32_t* p = xmalloc<uint32_t>(std::max(n, 1u));
  if (!p) return fail(GPM_ENOMEM, "host allocation failed");
  std::copy(h_new.begin(), h_new.begin() + n, p);
  b.release(d_off);
  b.release(d_adj);
  hipStream_t s = stream;
  stream = nullptr;
  st = adopt_device_csr(n, m, d_off, d_adj, std::move(h_off), h_adj.data(),
                        g->labels ? nlab.data() : nullptr, g->n_labels, g->flags, s, out);
  if (st != GPM_OK) {
    std::free(p);
    return st;
  }
  *perm = p;
  return GPM_OK;
}

}  // namespace

extern "C" {

int gpm_graph_build(uint32_t n, const uint32_t* src, const uint32_t* dst, uint64_t n_edges,
                    uint32_t csr_flags, const uint32_t* labels, uint32_t n_labels,
                    uint32_t graph_flags, gpm_graph** out) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    return build(n, src, dst, n_edges, csr_flags, labels, n_labels, graph_flags, out);
  });
}

int gpm_graph_relabel_by_degree(gpm_graph* g, gpm_graph** out, uint32_t** perm) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    return relabel(g, out, perm);
  });
}

int gpm_graph_copy_csr(const gpm_graph* g, uint32_t* offsets, uint32_t* adj) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g && offsets && (adj || g->m == 0), "null pointer");
    GPM_HIP(hipMemcpyAsync(offsets, g->off, ((size_t)g->n + 1) * 4, hipMemcpyDeviceToHost,
                           g->stream));
    if (g->m) GPM_HIP(hipMemcpyAsync(adj, g->adj, g->m * 4, hipMemcpyDeviceToHost, g->stream));
    GPM_HIP(hipStreamSynchronize(g->stream));
    return GPM_OK;
  });
}

}  // extern "C"
