// Binary on-disk graph formats of the reference, read and written on the
// host:
//   * sub_ceci compressed graph (sub_ceci/graph/graph.cpp:184-304 reader,
//     :306-366 writer): a degree file -- int32 4 (the int size), uint32 |V|,
//     uint32 |E| (undirected edges), uint32 degree[|V|] -- an adjacency file
//     holding the neighbour lists of vertices 0..|V|-1 back to back (2|E|
//     uint32 ids), and a label file of uint32 label[|V|].  The reader sorts
//     every list, as graph.cpp:238-239 does.
//   * GraphMiner binary directory (scripts/gbin_reader/greader.h:129-272,
//     the converter's output): graph.vertex.bin = int64 row pointer[|V|+1],
//     graph.edge.bin = int32 column ids[nnz], graph.meta.txt = "|V|\n nnz\n
//     4\n max_degree\n".  Lists are kept in file order.
// Unlike the reference readers, these check sizes against the headers and
// reject inconsistent files with GPM_EIO instead of reading past them.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "common.hpp"

using namespace gpm;

namespace {

bool read_exact(std::ifstream& f, void* dst, size_t bytes) {
  f.read(static_cast<char*>(dst), (std::streamsize)bytes);
  return (size_t)f.gcount() == bytes;
}

bool at_eof(std::ifstream& f) { return f.peek() == std::ifstream::traits_type::eof(); }

// Hands host vectors out as malloc'ed arrays (released with gpm_free).
template <typename T>
int hand_out(const std::vector<T>& v, T** out) {
  T* p = xmalloc<T>(std::max<size_t>(v.size(), 1));
  if (!p) return fail(GPM_ENOMEM, "host allocation failed");
  std::copy(v.begin(), v.end(), p);
  *out = p;
  return GPM_OK;
}

}  // namespace

extern "C" {

int gpm_load_ceci_bin(const char* degree_path, const char* adj_path, const char* label_path,
                      uint32_t* n, uint32_t** offsets, uint32_t** adj, uint64_t* m,
                      uint32_t** labels, uint32_t* n_labels) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(degree_path && adj_path && n && offsets && adj && m, "null pointer");
    GPM_CHECK_ARG(!label_path || (labels && n_labels), "null label outputs");
    std::ifstream df(degree_path, std::ios::binary);
    if (!df.is_open()) return fail(GPM_EIO, "cannot open degree file %s", degree_path);
    int32_t int_size = 0;
    uint32_t nv = 0, ne = 0;
    if (!read_exact(df, &int_size, 4) || !read_exact(df, &nv, 4) || !read_exact(df, &ne, 4))
      return fail(GPM_EIO, "%s: truncated header", degree_path);
    if (int_size != 4) return fail(GPM_EIO, "%s: int size %d (want 4)", degree_path, int_size);
    if (nv == 0xFFFFFFFFu || 2ull * ne >= (1ull << 32))
      return fail(GPM_EIO, "%s: %u vertices / %u edges exceed 32-bit ids", degree_path, nv, ne);
    std::vector<uint32_t> deg(nv);
    if (nv && !read_exact(df, deg.data(), (size_t)nv * 4))
      return fail(GPM_EIO, "%s: fewer than %u degrees", degree_path, nv);
    if (!at_eof(df)) return fail(GPM_EIO, "%s: trailing bytes after the degrees", degree_path);
    std::vector<uint32_t> off((size_t)nv + 1, 0);
    for (uint32_t v = 0; v < nv; ++v) {
      off[v + 1] = off[v] + deg[v];
      if ((uint64_t)off[v] + deg[v] > 2ull * ne)
        return fail(GPM_EIO, "%s: degrees add up to more than 2|E| = %llu", degree_path,
                    2ull * ne);
    }
    if (off[nv] != 2ull * ne)
      return fail(GPM_EIO, "%s: degrees add up to %u, header says 2|E| = %llu", degree_path,
                  off[nv], 2ull * ne);
    std::ifstream af(adj_path, std::ios::binary);
    if (!af.is_open()) return fail(GPM_EIO, "cannot open adjacency file %s", adj_path);
    std::vector<uint32_t> nb((size_t)2 * ne);
    if (!nb.empty() && !read_exact(af, nb.data(), nb.size() * 4))
      return fail(GPM_EIO, "%s: fewer than 2|E| = %zu neighbour ids", adj_path, nb.size());
    if (!at_eof(af)) return fail(GPM_EIO, "%s: trailing bytes after 2|E| ids", adj_path);
    for (uint32_t x : nb)
      if (x >= nv) return fail(GPM_EIO, "%s: neighbour id %u >= |V| = %u", adj_path, x, nv);
    for (uint32_t v = 0; v < nv; ++v) std::sort(nb.begin() + off[v], nb.begin() + off[v + 1]);
    std::vector<uint32_t> lab;
    uint32_t max_label = 0;
    if (label_path) {
      std::ifstream lf(label_path, std::ios::binary);
      if (!lf.is_open()) return fail(GPM_EIO, "cannot open label file %s", label_path);
      lab.resize(nv);
      if (nv && !read_exact(lf, lab.data(), (size_t)nv * 4))
        return fail(GPM_EIO, "%s: fewer than %u labels", label_path, nv);
      if (!at_eof(lf)) return fail(GPM_EIO, "%s: trailing bytes after the labels", label_path);
      for (uint32_t l : lab) max_label = std::max(max_label, l);
      if (max_label == 0xFFFFFFFFu) return fail(GPM_EIO, "%s: label 2^32-1", label_path);
    }
    uint32_t *o = nullptr, *a = nullptr, *l = nullptr;
    int st;
    if ((st = hand_out(off, &o)) != GPM_OK) return st;
    if ((st = hand_out(nb, &a)) != GPM_OK) {
      std::free(o);
      return st;
    }
    if (label_path && (st = hand_out(lab, &l)) != GPM_OK) {
      std::free(o);
      std::free(a);
      return st;
    }
    *n = nv;
    *offsets = o;
    *adj = a;
    *m = nb.size();
    if (label_path) {
      *labels = l;
      *n_labels = nv ? max_label + 1 : 0;
    }
    return GPM_OK;
  });
}

int gpm_store_ceci_bin(const char* degree_path, const char* adj_path, const char* label_path,
                       uint32_t n, const uint32_t* offsets, const uint32_t* adj,
                       const uint32_t* labels) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(degree_path && adj_path && offsets, "null pointer");
    GPM_CHECK_ARG(!label_path || labels, "label path without labels");
    const uint64_t m = offsets[n];
    GPM_CHECK_ARG(m % 2 == 0 && m / 2 < (1ull << 32), "adjacency entries %llu: not 2|E|",
                  (unsigned long long)m);
    GPM_CHECK_ARG(adj || m == 0, "null adjacency");
    std::ofstream df(degree_path, std::ios::binary);
    if (!df.is_open()) return fail(GPM_EIO, "cannot create %s", degree_path);
    const int32_t int_size = 4;
    const uint32_t ne = (uint32_t)(m / 2);
    df.write(reinterpret_cast<const char*>(&int_size), 4);
    df.write(reinterpret_cast<const char*>(&n), 4);
    df.write(reinterpret_cast<const char*>(&ne), 4);
    for (uint32_t v = 0; v < n; ++v) {
      GPM_CHECK_ARG(offsets[v] <= offsets[v + 1], "offsets decrease at %u", v);
      const uint32_t d = offsets[v + 1] - offsets[v];
      df.write(reinterpret_cast<const char*>(&d), 4);
    }
    if (!df.good()) return fail(GPM_EIO, "write failed: %s", degree_path);
    std::ofstream af(adj_path, std::ios::binary);
    if (!af.is_open()) return fail(GPM_EIO, "cannot create %s", adj_path);
    if (m) af.write(reinterpret_cast<const char*>(adj), (std::streamsize)(m * 4));
    if (!af.good()) return fail(GPM_EIO, "write failed: %s", adj_path);
    if (label_path) {
      std::ofstream lf(label_path, std::ios::binary);
      if (!lf.is_open()) return fail(GPM_EIO, "cannot create %s", label_path);
      if (n) lf.write(reinterpret_cast<const char*>(labels), (std::streamsize)n * 4);
      if (!lf.good()) return fail(GPM_EIO, "write failed: %s", label_path);
    }
    return GPM_OK;
  });
}

int gpm_load_gbin(const char* dir, uint32_t* n, uint32_t** offsets, uint32_t** adj,
                  uint64_t* m, uint32_t* max_degree) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(dir && n && offsets && adj && m && max_degree, "null pointer");
    const std::string base(dir);
    std::ifstream meta(base + "/graph.meta.txt");
    if (!meta.is_open()) return fail(GPM_EIO, "cannot open %s/graph.meta.txt", dir);
    int64_t nv = -1, nnz = -1, bytes = -1, maxd = -1;
    if (!(meta >> nv >> nnz >> bytes >> maxd) || nv < 0 || nnz < 0 || maxd < 0)
      return fail(GPM_EIO, "%s/graph.meta.txt: want |V|, nnz, 4, max degree", dir);
    if (bytes != 4) return fail(GPM_EIO, "%s/graph.meta.txt: id size %lld (want 4)", dir,
                                (long long)bytes);
    if (nv >= 0xFFFFFFFFll || nnz >= (1ll << 32))
      return fail(GPM_EIO, "%s: %lld vertices / %lld entries exceed 32-bit ids", dir,
                  (long long)nv, (long long)nnz);
    std::ifstream vf(base + "/graph.vertex.bin", std::ios::binary);
    if (!vf.is_open()) return fail(GPM_EIO, "cannot open %s/graph.vertex.bin", dir);
    std::vector<int64_t> rp((size_t)nv + 1);
    if (!read_exact(vf, rp.data(), rp.size() * 8) || !at_eof(vf))
      return fail(GPM_EIO, "%s/graph.vertex.bin: want %lld int64 row pointers", dir,
                  (long long)nv + 1);
    std::ifstream ef(base + "/graph.edge.bin", std::ios::binary);
    if (!ef.is_open()) return fail(GPM_EIO, "cannot open %s/graph.edge.bin", dir);
    std::vector<int32_t> col((size_t)nnz);
    if ((nnz && !read_exact(ef, col.data(), col.size() * 4)) || !at_eof(ef))
      return fail(GPM_EIO, "%s/graph.edge.bin: want %lld int32 ids", dir, (long long)nnz);
    if (rp[0] != 0 || rp[nv] != nnz)
      return fail(GPM_EIO, "%s: row pointers span [%lld, %lld), want [0, %lld)", dir,
                  (long long)rp[0], (long long)rp[nv], (long long)nnz);
    std::vector<uint32_t> off((size_t)nv + 1);
    int64_t seen_max = 0;
    for (int64_t v = 0; v <= nv; ++v) {
      if (v && rp[v] < rp[v - 1]) return fail(GPM_EIO, "%s: row pointer decreases at %lld", dir,
                                              (long long)v);
      off[v] = (uint32_t)rp[v];
      if (v) seen_max = std::max<int64_t>(seen_max, rp[v] - rp[v - 1]);
    }
    if (seen_max != maxd)
      return fail(GPM_EIO, "%s: max degree %lld, meta says %lld", dir, (long long)seen_max,
                  (long long)maxd);
    std::vector<uint32_t> a((size_t)nnz);
    for (int64_t e = 0; e < nnz; ++e) {
      if (col[e] < 0 || col[e] >= nv)
        return fail(GPM_EIO, "%s: id %d out of [0, %lld)", dir, col[e], (long long)nv);
      a[e] = (uint32_t)col[e];
    }
    uint32_t *o = nullptr, *aa = nullptr;
    int st;
    if ((st = hand_out(off, &o)) != GPM_OK) return st;
    if ((st = hand_out(a, &aa)) != GPM_OK) {
      std::free(o);
      return st;
    }
    *n = (uint32_t)nv;
    *offsets = o;
    *adj = aa;
    *m = (uint64_t)nnz;
    *max_degree = (uint32_t)maxd;
    return GPM_OK;
  });
}

int gpm_store_gbin(const char* dir, uint32_t n, const uint32_t* offsets, const uint32_t* adj) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(dir && offsets, "null pointer");
    const uint64_t m = offsets[n];
    GPM_CHECK_ARG(adj || m == 0, "null adjacency");
    GPM_CHECK_ARG(n < 0x7FFFFFFFu, "n = %u exceeds int32 ids", n);
    const std::string base(dir);
    std::vector<int64_t> rp((size_t)n + 1);
    uint32_t maxd = 0;
    for (uint32_t v = 0; v <= n; ++v) {
      rp[v] = offsets[v];
      if (v) {
        GPM_CHECK_ARG(offsets[v] >= offsets[v - 1], "offsets decrease at %u", v);
        maxd = std::max(maxd, offsets[v] - offsets[v - 1]);
      }
    }
    std::ofstream vf(base + "/graph.vertex.bin", std::ios::binary);
    if (!vf.is_open()) return fail(GPM_EIO, "cannot create %s/graph.vertex.bin", dir);
    vf.write(reinterpret_cast<const char*>(rp.data()), (std::streamsize)(rp.size() * 8));
    if (!vf.good()) return fail(GPM_EIO, "write failed: %s/graph.vertex.bin", dir);
    std::ofstream ef(base + "/graph.edge.bin", std::ios::binary);
    if (!ef.is_open()) return fail(GPM_EIO, "cannot create %s/graph.edge.bin", dir);
    if (m) ef.write(reinterpret_cast<const char*>(adj), (std::streamsize)(m * 4));
    if (!ef.good()) return fail(GPM_EIO, "write failed: %s/graph.edge.bin", dir);
    std::ofstream mf(base + "/graph.meta.txt");
    if (!mf.is_open()) return fail(GPM_EIO, "cannot create %s/graph.meta.txt", dir);
    mf << n << "\n" << m << "\n" << 4 << "\n" << maxd << "\n";
    if (!mf.good()) return fail(GPM_EIO, "write failed: %s/graph.meta.txt", dir);
    return GPM_OK;
  });
}

}  // extern "C"
