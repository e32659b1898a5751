// Host-side graph preparation of libgpm: CSR construction, degree
// orientation, degree relabel, the matrix-market reader and the seeded
// generators used by the workloads.  OpenMP-parallel where it pays; every
// result is independent of the thread count.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <fstream>
#include <limits>
#include <numeric>
#include <string>
#include <unordered_set>
#include <vector>

#include "common.hpp"

namespace gpm {

static thread_local std::string g_err;

int fail(int status, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return status;
}

void clear_error() { g_err.clear(); }

void build_label_csr(uint32_t n, uint32_t L, const uint32_t* offsets,
                     const uint32_t* adj, const uint32_t* labels,
                     std::vector<uint32_t>& ladj, std::vector<uint32_t>& loff) {
  ladj.assign(offsets[n], 0);
  loff.assign((size_t)n * (L + 1), 0);
#pragma omp parallel
  {
    std::vector<uint32_t> cnt(L + 1);
#pragma omp for schedule(dynamic, 1024)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
      const uint32_t b = offsets[v], e = offsets[v + 1];
      std::fill(cnt.begin(), cnt.end(), 0u);
      for (uint32_t i = b; i < e; ++i) cnt[labels[adj[i]] + 1]++;
      for (uint32_t l = 0; l < L; ++l) cnt[l + 1] += cnt[l];
      for (uint32_t l = 0; l <= L; ++l) loff[(size_t)v * (L + 1) + l] = b + cnt[l];
      // Lists are sorted by id, so a stable bucket pass keeps ids sorted
      // within each label.
      for (uint32_t i = b; i < e; ++i) ladj[b + cnt[labels[adj[i]]]++] = adj[i];
    }
  }
}

namespace {

// Exclusive prefix sum of 32-bit counts into 64-bit offsets (n+1 entries).
uint64_t prefix_sum(const std::vector<uint32_t>& cnt, std::vector<uint64_t>& off) {
  const size_t n = cnt.size();
  off.assign(n + 1, 0);
  for (size_t i = 0; i < n; ++i) off[i + 1] = off[i] + cnt[i];
  return off[n];
}

int export_csr(const std::vector<uint64_t>& off64, std::vector<uint32_t>& adj,
               uint32_t** offsets, uint32_t** adj_out, uint64_t* m) {
  const size_t n = off64.size() - 1;
  const uint64_t total = off64[n];
  if (total >= (1ull << 32))
    return fail(GPM_EINVAL, "CSR has %llu entries; uint32 offsets hold < 2^32",
                (unsigned long long)total);
  uint32_t* o = xmalloc<uint32_t>(n + 1);
  uint32_t* a = xmalloc<uint32_t>(total);
  if (!o || !a) {
    std::free(o);
    std::free(a);
    return fail(GPM_ENOMEM, "host allocation of CSR (%llu entries) failed",
                (unsigned long long)total);
  }
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i <= n; ++i) o[i] = (uint32_t)off64[i];
  if (total) std::memcpy(a, adj.data(), total * sizeof(uint32_t));
  *offsets = o;
  *adj_out = a;
  *m = total;
  return GPM_OK;
}

}  // namespace
}  // namespace gpm

using namespace gpm;

extern "C" {

const char* gpm_last_error(void) { return g_err.c_str(); }
int gpm_abi_version(void) { return 2; }
void gpm_free(void* p) { std::free(p); }

int gpm_csr_from_edges(uint32_t n, const uint32_t* src, const uint32_t* dst,
                       uint64_t n_edges, uint32_t flags, uint32_t** offsets,
                       uint32_t** adj, uint64_t* m) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(offsets && adj && m, "null output pointer");
    GPM_CHECK_ARG(n_edges == 0 || (src && dst), "null edge arrays");
    const bool sym = flags & GPM_CSR_SYMMETRIZE;
    const bool dedup = flags & GPM_CSR_DEDUP;
    const bool noself = flags & GPM_CSR_DROP_SELF;
    const bool sorted = (flags & GPM_CSR_SORT) || dedup;
    // Validate ids first so that no later loop indexes out of range.
    std::atomic<bool> bad{false};
  #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n_edges; ++i)
      if (src[i] >= n || dst[i] >= n) bad.store(true, std::memory_order_relaxed);
    if (bad) return fail(GPM_EINVAL, "edge endpoint >= n (%u)", n);

    std::vector<uint32_t> deg(n, 0);
    auto keep = [&](uint32_t u, uint32_t v) { return !(noself && u == v); };
    if (sorted) {
  #pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < (int64_t)n_edges; ++i) {
        uint32_t u = src[i], v = dst[i];
        if (!keep(u, v)) continue;
        __atomic_fetch_add(&deg[u], 1u, __ATOMIC_RELAXED);
        if (sym) __atomic_fetch_add(&deg[v], 1u, __ATOMIC_RELAXED);
      }
    } else {
      for (uint64_t i = 0; i < n_edges; ++i) {
        uint32_t u = src[i], v = dst[i];
        if (!keep(u, v)) continue;
        deg[u]++;
        if (sym) deg[v]++;
      }
    }
    std::vector<uint64_t> off;
    uint64_t total = prefix_sum(deg, off);
    std::vector<uint32_t> buf(total);
    if (sorted) {
      // Order inside a list does not matter before the per-list sort.
      std::vector<uint64_t> cur(off.begin(), off.end() - 1);
  #pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < (int64_t)n_edges; ++i) {
        uint32_t u = src[i], v = dst[i];
        if (!keep(u, v)) continue;
        buf[__atomic_fetch_add(&cur[u], 1ull, __ATOMIC_RELAXED)] = v;
        if (sym) buf[__atomic_fetch_add(&cur[v], 1ull, __ATOMIC_RELAXED)] = u;
      }
  #pragma omp parallel for schedule(dynamic, 256)
      for (int64_t u = 0; u < (int64_t)n; ++u)
        std::sort(buf.begin() + off[u], buf.begin() + off[u + 1]);
    } else {
      // Input order, as Graph::Init (src/graph/graph.h:281-287).
      std::vector<uint64_t> cur(off.begin(), off.end() - 1);
      for (uint64_t i = 0; i < n_edges; ++i) {
        uint32_t u = src[i], v = dst[i];
        if (!keep(u, v)) continue;
        buf[cur[u]++] = v;
        if (sym) buf[cur[v]++] = u;
      }
    }
    if (dedup) {
      std::vector<uint32_t> ndeg(n);
  #pragma omp parallel for schedule(dynamic, 256)
      for (int64_t u = 0; u < (int64_t)n; ++u) {
        auto b = buf.begin() + off[u], e = buf.begin() + off[u + 1];
        ndeg[u] = (uint32_t)(std::unique(b, e) - b);
      }
      std::vector<uint64_t> noff;
      uint64_t ntotal = prefix_sum(ndeg, noff);
      std::vector<uint32_t> nbuf(ntotal);
  #pragma omp parallel for schedule(dynamic, 256)
      for (int64_t u = 0; u < (int64_t)n; ++u)
        std::copy(buf.begin() + off[u], buf.begin() + off[u] + ndeg[u],
                  nbuf.begin() + noff[u]);
      return export_csr(noff, nbuf, offsets, adj, m);
    }
    return export_csr(off, buf, offsets, adj, m);
  });
}

int gpm_csr_orient(uint32_t n, const uint32_t* offsets, const uint32_t* adj,
                   uint32_t** out_offsets, uint32_t** out_adj, uint64_t* out_m) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(offsets && out_offsets && out_adj && out_m, "null pointer");
    GPM_CHECK_ARG(offsets[n] == 0 || adj, "null adjacency");
    // deg = list length in the input CSR (odegs_ of graph.h:128-129).
    auto deg = [&](uint32_t v) { return offsets[v + 1] - offsets[v]; };
    auto keep = [&](uint32_t u, uint32_t v) {
      uint32_t du = deg(u), dv = deg(v);
      return dv > du || (dv == du && v > u);
    };
    std::vector<uint32_t> cnt(n, 0);
  #pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t u = 0; u < (int64_t)n; ++u) {
      uint32_t c = 0;
      for (uint32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
        if (adj[e] >= n) continue;  // rejected below
        c += keep((uint32_t)u, adj[e]);
      }
      cnt[u] = c;
    }
    for (uint64_t e = 0; e < offsets[n]; ++e)
      if (adj[e] >= n) return fail(GPM_EINVAL, "adjacency id %u >= n", adj[e]);
    std::vector<uint64_t> off;
    uint64_t total = prefix_sum(cnt, off);
    std::vector<uint32_t> buf(total);
  #pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t u = 0; u < (int64_t)n; ++u) {
      uint64_t p = off[u];
      for (uint32_t e = offsets[u]; e < offsets[u + 1]; ++e)
        if (keep((uint32_t)u, adj[e])) buf[p++] = adj[e];
    }
    return export_csr(off, buf, out_offsets, out_adj, out_m);
  });
}

int gpm_csr_sort_by_degree(uint32_t n, const uint32_t* offsets,
                           const uint32_t* adj, uint32_t** out_offsets,
                           uint32_t** out_adj, uint32_t** perm) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(offsets && out_offsets && out_adj && perm, "null pointer");
    auto deg = [&](uint32_t v) { return offsets[v + 1] - offsets[v]; };
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return deg(a) > deg(b); });
    uint32_t* newid = xmalloc<uint32_t>(n);
    if (!newid) return fail(GPM_ENOMEM, "host allocation failed");
    for (uint32_t i = 0; i < n; ++i) newid[order[i]] = i;
    std::vector<uint32_t> cnt(n);
    for (uint32_t i = 0; i < n; ++i) cnt[i] = deg(order[i]);
    std::vector<uint64_t> off;
    uint64_t total = prefix_sum(cnt, off);
    std::vector<uint32_t> buf(total);
  #pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      uint32_t v = order[i];
      uint64_t p = off[i];
      for (uint32_t e = offsets[v]; e < offsets[v + 1]; ++e) buf[p++] = newid[adj[e]];
      std::sort(buf.begin() + off[i], buf.begin() + off[i + 1]);
    }
    int st = export_csr(off, buf, out_offsets, out_adj, &total);
    if (st != GPM_OK) {
      std::free(newid);
      return st;
    }
    *perm = newid;
    return GPM_OK;
  });
}

int gpm_load_mtx(const char* path, uint32_t* n, uint32_t** src, uint32_t** dst,
                 uint64_t* n_edges) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(path && n && src && dst && n_edges, "null pointer");
    std::ifstream fin(path);
    if (!fin.is_open()) return fail(GPM_EIO, "cannot open file %s", path);
    // Skip lines until one starts with a digit (io.h:91-96).
    while (true) {
      int c = fin.peek();
      if (c == std::char_traits<char>::eof())
        return fail(GPM_EIO, "%s: no data line found", path);
      if (c >= '0' && c <= '9') break;
      fin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    }
    int64_t hn0 = 0, hn1 = 0, hm = 0;
    fin >> hn0 >> hn1 >> hm;  // header, read and ignored (io.h:99-100)
    std::vector<uint32_t> s, d;
    int64_t vmin = std::numeric_limits<int64_t>::max(), vmax = -1;
    // Pair loop of io.h:111-126: a pair is kept only if the stream is not at
    // EOF after skipping the rest of its line.
    while (fin.good()) {
      int64_t v0 = 0, v1 = 0;
      fin >> v0 >> v1;
      fin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
      if (fin.eof()) break;
      if (v0 == v1) continue;
      if (v0 < 0 || v1 < 0 || v0 > 0xFFFFFFFEll || v1 > 0xFFFFFFFEll)
        return fail(GPM_EIO, "%s: vertex id out of range", path);
      vmin = std::min(vmin, std::min(v0, v1));
      vmax = std::max(vmax, std::max(v0, v1));
      s.push_back((uint32_t)v0);
      d.push_back((uint32_t)v1);
    }
    uint64_t ne = s.size();
    uint32_t nv = ne ? (uint32_t)(vmax - vmin + 1) : 0;
    uint32_t* so = xmalloc<uint32_t>(ne);
    uint32_t* dso = xmalloc<uint32_t>(ne);
    if (!so || !dso) {
      std::free(so);
      std::free(dso);
      return fail(GPM_ENOMEM, "host allocation failed");
    }
    for (uint64_t i = 0; i < ne; ++i) {
      so[i] = (uint32_t)(s[i] - vmin);
      dso[i] = (uint32_t)(d[i] - vmin);
    }
    *n = nv;
    *src = so;
    *dst = dso;
    *n_edges = ne;
    return GPM_OK;
  });
}

}  // extern "C"

namespace gpm {

bool topo_relabel(uint32_t n, const uint32_t* offsets, const uint32_t* adj,
                  std::vector<uint32_t>& rank, std::vector<uint32_t>& new_off,
                  std::vector<uint32_t>& new_adj) {
  const uint64_t m = offsets[n];
  std::vector<uint32_t> indeg(n, 0);
  for (uint64_t e = 0; e < m; ++e) indeg[adj[e]]++;
  std::vector<uint32_t> order;
  order.reserve(n);
  for (uint32_t v = 0; v < n; ++v)
    if (!indeg[v]) order.push_back(v);
  for (size_t h = 0; h < order.size(); ++h) {
    const uint32_t v = order[h];
    for (uint32_t e = offsets[v]; e < offsets[v + 1]; ++e)
      if (--indeg[adj[e]] == 0) order.push_back(adj[e]);
  }
  if (order.size() != n) return false;
  rank.assign(n, 0);
  for (uint32_t r = 0; r < n; ++r) rank[order[r]] = r;
  new_off.assign((size_t)n + 1, 0);
  for (uint32_t r = 0; r < n; ++r) new_off[r + 1] = new_off[r] + (offsets[order[r] + 1] - offsets[order[r]]);
  new_adj.resize(m);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < (int64_t)n; ++r) {
    const uint32_t v = order[r];
    uint32_t* out = new_adj.data() + new_off[r];
    const uint32_t len = offsets[v + 1] - offsets[v];
    for (uint32_t i = 0; i < len; ++i) out[i] = rank[adj[offsets[v] + i]];
    std::sort(out, out + len);
  }
  return true;
}

}  // namespace gpm

extern "C" {

int gpm_load_labeled(const char* path, uint32_t* n, uint32_t** offsets, uint32_t** adj,
                     uint64_t* m, uint32_t** labels, uint32_t* n_labels) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(path && n && offsets && adj && m && labels && n_labels, "null pointer");
    std::ifstream fin(path);
    if (!fin.is_open()) return fail(GPM_EIO, "cannot open file %s", path);
    // Header "t |V| |E|" (sub_ceci/graph/graph.cpp:94-95).
    char type = 0;
    int64_t nv = -1, ne = -1;
    if (!(fin >> type >> nv >> ne) || type != 't' || nv < 0 || ne < 0 || nv > 0xFFFFFFFEll ||
        2 * ne >= (1ll << 32))
      return fail(GPM_EIO, "%s: bad header (want \"t |V| |E|\")", path);
    std::vector<uint32_t> off(nv + 1, 0), lab(nv, 0), fill(nv, 0);
    std::vector<uint32_t> nb((size_t)2 * ne);
    int64_t next_v = 0, n_e = 0;
    uint32_t max_label = 0;
    // "v id label degree" lines in id order, then "e u v" lines, each
    // undirected edge once (graph.cpp:106-144).  The declared degrees size the
    // lists, as in the reference; they must add up and be met exactly.
    while (fin >> type) {
      if (type == 'v') {
        int64_t id = -1, l = -1, deg = -1;
        if (!(fin >> id >> l >> deg) || id != next_v || l < 0 || l > 0xFFFFFFFEll || deg < 0)
          return fail(GPM_EIO, "%s: bad vertex line for vertex %lld", path, (long long)next_v);
        if ((int64_t)off[id] + deg > 2 * ne)
          return fail(GPM_EIO, "%s: declared degrees exceed 2|E|", path);
        lab[id] = (uint32_t)l;
        max_label = std::max(max_label, (uint32_t)l);
        off[id + 1] = off[id] + (uint32_t)deg;
        ++next_v;
      } else if (type == 'e') {
        int64_t a = -1, b = -1;
        if (!(fin >> a >> b) || a < 0 || b < 0 || a >= next_v || b >= next_v || a == b)
          return fail(GPM_EIO, "%s: bad edge line %lld", path, (long long)n_e);
        if (off[a] + fill[a] >= off[a + 1] || off[b] + fill[b] >= off[b + 1])
          return fail(GPM_EIO, "%s: edge %lld exceeds a declared degree", path, (long long)n_e);
        nb[off[a] + fill[a]++] = (uint32_t)b;
        nb[off[b] + fill[b]++] = (uint32_t)a;
        ++n_e;
      } else {
        return fail(GPM_EIO, "%s: unknown line type '%c'", path, type);
      }
    }
    if (next_v != nv || n_e != ne || off[nv] != 2 * ne)
      return fail(GPM_EIO, "%s: header says %lld vertices / %lld edges, file has %lld / %lld",
                  path, (long long)nv, (long long)ne, (long long)next_v, (long long)n_e);
    for (int64_t v = 0; v < nv; ++v) {
      if (fill[v] != off[v + 1] - off[v])
        return fail(GPM_EIO, "%s: vertex %lld has fewer edges than declared", path, (long long)v);
      std::sort(nb.begin() + off[v], nb.begin() + off[v + 1]);  // graph.cpp:156-158
    }
    uint32_t* o = xmalloc<uint32_t>((size_t)nv + 1);
    uint32_t* a = xmalloc<uint32_t>(std::max<size_t>(nb.size(), 1));
    uint32_t* l = xmalloc<uint32_t>(std::max<int64_t>(nv, 1));
    if (!o || !a || !l) {
      std::free(o);
      std::free(a);
      std::free(l);
      return fail(GPM_ENOMEM, "host allocation failed");
    }
    std::copy(off.begin(), off.end(), o);
    std::copy(nb.begin(), nb.end(), a);
    std::copy(lab.begin(), lab.end(), l);
    *n = (uint32_t)nv;
    *offsets = o;
    *adj = a;
    *m = nb.size();
    *labels = l;
    *n_labels = nv ? max_label + 1 : 0;
    return GPM_OK;
  });
}

int gpm_gen_er(uint32_t n, uint64_t m_undirected, uint64_t seed,
               uint32_t** src, uint32_t** dst, uint64_t* n_edges) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(src && dst && n_edges, "null pointer");
    GPM_CHECK_ARG(n >= 2, "n must be >= 2");
    const uint64_t max_m = (uint64_t)n * (n - 1) / 2;
    GPM_CHECK_ARG(m_undirected <= max_m, "M=%llu exceeds n(n-1)/2 = %llu",
                  (unsigned long long)m_undirected, (unsigned long long)max_m);
    uint32_t* s = xmalloc<uint32_t>(m_undirected);
    uint32_t* d = xmalloc<uint32_t>(m_undirected);
    if (!s || !d) {
      std::free(s);
      std::free(d);
      return fail(GPM_ENOMEM, "host allocation failed");
    }
    // Sequential rejection sampling over one splitmix64 stream: draw (u, v),
    // reject self loops and pairs already drawn; emit (min, max) in draw order.
    std::unordered_set<uint64_t> seen;
    seen.reserve(m_undirected * 2);
    SplitMix rng(seed);
    uint64_t k = 0;
    while (k < m_undirected) {
      uint64_t r = rng.next();
      uint32_t u = (uint32_t)(((r & 0xFFFFFFFFull) * n) >> 32);
      uint32_t v = (uint32_t)(((r >> 32) * n) >> 32);
      if (u == v) continue;
      uint32_t a = std::min(u, v), b = std::max(u, v);
      uint64_t key = ((uint64_t)a << 32) | b;
      if (!seen.insert(key).second) continue;
      s[k] = a;
      d[k] = b;
      ++k;
    }
    *src = s;
    *dst = d;
    *n_edges = m_undirected;
    return GPM_OK;
  });
}

int gpm_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                 uint64_t seed, uint32_t** src, uint32_t** dst,
                 uint64_t* n_edges) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(src && dst && n_edges, "null pointer");
    GPM_CHECK_ARG(scale >= 1 && scale <= 31, "scale must be in [1, 31]");
    GPM_CHECK_ARG(edge_factor >= 1, "edge_factor must be >= 1");
    GPM_CHECK_ARG(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0,
                  "bad quadrant probabilities");
    const uint64_t ne = (uint64_t)edge_factor << scale;
    uint32_t* s = xmalloc<uint32_t>(ne);
    uint32_t* d = xmalloc<uint32_t>(ne);
    if (!s || !d) {
      std::free(s);
      std::free(d);
      return fail(GPM_ENOMEM, "host allocation failed");
    }
    const double ab = a + b, abc = a + b + c;
    const uint64_t base = mix64(seed);
    // Sample i, level l draws value mix64(base + i*scale + l): counter based,
    // so the output does not depend on the thread count.
  #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)ne; ++i) {
      uint32_t u = 0, v = 0;
      const uint64_t ctr = base + (uint64_t)i * (uint64_t)scale;
      for (int l = 0; l < scale; ++l) {
        double r = u01(mix64(ctr + (uint64_t)l));
        uint32_t bit = 1u << (scale - 1 - l);
        if (r < a) {
        } else if (r < ab) {
          v |= bit;
        } else if (r < abc) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      s[i] = u;
      d[i] = v;
    }
    *src = s;
    *dst = d;
    *n_edges = ne;
    return GPM_OK;
  });
}

int gpm_gen_labels(uint32_t n, uint32_t n_labels, uint64_t seed,
                   uint32_t** labels) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(labels, "null pointer");
    GPM_CHECK_ARG(n_labels >= 1, "n_labels must be >= 1");
    uint32_t* l = xmalloc<uint32_t>(n);
    if (!l) return fail(GPM_ENOMEM, "host allocation failed");
    const uint64_t base = mix64(seed ^ 0x6c6162656c73ull);
  #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
      l[i] = (uint32_t)(((mix64(base + (uint64_t)i) >> 32) * n_labels) >> 32);
    *labels = l;
    return GPM_OK;
  });
}

}  // extern "C"
