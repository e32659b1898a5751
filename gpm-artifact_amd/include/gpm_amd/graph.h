// Host graph of the reference API (src/graph/graph.h:88-438), kept as host
// vectors; the device copy is a gpm_graph handle made on first use.
//
// Naming follows the reference, whose members are swapped against their
// usual meaning: row_ptr_ holds the adjacency entries and col_idx_ the
// offsets (graph.h:422-423).
#pragma once

#include <algorithm>
#include <iostream>
#include <memory>
#include <utility>
#include <vector>

#include "utils.h"

namespace project_AntiRF {

template <typename VID, typename VLABEL>
class Graph {
 public:
  Graph() = default;
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;
  ~Graph() { release_device(); }

  // Graph::Init (graph.h:254-340): CSR in input order, COO copy.
  void Init(std::vector<VID> const& vids, std::vector<VLABEL> const& vlabels,
            std::vector<std::pair<VID, VID>> const& edges, bool use_label = false) {
    vsize_ = vids.size();
    esize_ = edges.size();
    if (use_label) vlabels_ = vlabels;
    std::vector<uint32_t> s(esize_), d(esize_);
    for (size_t i = 0; i < esize_; ++i) {
      s[i] = (uint32_t)edges[i].first;
      d[i] = (uint32_t)edges[i].second;
    }
    uint32_t *po = nullptr, *pa = nullptr;
    uint64_t m = 0;
    check_status(gpm_csr_from_edges((uint32_t)vsize_, s.data(), d.data(), esize_, 0u, &po, &pa, &m),
                 "Graph::Init");
    col_idx_.assign(po, po + vsize_ + 1);
    row_ptr_.assign(pa, pa + m);
    gpm_free(po);
    gpm_free(pa);
    src_list_ = std::vector<VID>(s.begin(), s.end());
    dst_list_ = std::vector<VID>(d.begin(), d.end());
    odegs_.resize(vsize_);
    for (size_t v = 0; v < vsize_; ++v) odegs_[v] = col_idx_[v + 1] - col_idx_[v];
    max_degree_ = vsize_ ? *std::max_element(odegs_.begin(), odegs_.end()) : 0;
    oriented_ = false;
    release_device();
  }

  // Graph::orientation (graph.h:116-204): keep u->v iff deg v > deg u or
  // (equal and v > u).  Lists are then sorted and deduplicated, which the
  // engine's set operations need (the reference assumes sorted input).
  void orientation(bool NeedToLoadToDevice = true) {
    uint32_t *po = nullptr, *pa = nullptr;
    uint64_t m = 0;
    sort_dedup();
    check_status(gpm_csr_orient((uint32_t)vsize_, offsets_u32().data(), adj_u32().data(), &po,
                                &pa, &m),
                 "Graph::orientation");
    col_idx_.assign(po, po + vsize_ + 1);
    row_ptr_.assign(pa, pa + m);
    gpm_free(po);
    gpm_free(pa);
    esize_ = m;
    src_list_.resize(m);
    for (size_t v = 0; v < vsize_; ++v)
      for (VID e = col_idx_[v]; e < col_idx_[v + 1]; ++e) src_list_[e] = (VID)v;
    dst_list_ = row_ptr_;
    max_degree_ = 0;
    for (size_t v = 0; v < vsize_; ++v)
      max_degree_ = std::max<VID>(max_degree_, col_idx_[v + 1] - col_idx_[v]);
    oriented_ = true;
    release_device();
    if (NeedToLoadToDevice) DeviceObject();
  }

  // Graph::SortCSRGraph (graph.h:206-252): relabel by out-degree, descending.
  void SortCSRGraph(bool NeedToLoadToDevice = true) {
    uint32_t *po = nullptr, *pa = nullptr, *perm = nullptr;
    check_status(gpm_csr_sort_by_degree((uint32_t)vsize_, offsets_u32().data(), adj_u32().data(),
                                        &po, &pa, &perm),
                 "Graph::SortCSRGraph");
    col_idx_.assign(po, po + vsize_ + 1);
    row_ptr_.assign(pa, pa + esize_);
    for (size_t v = 0; v < vsize_; ++v) odegs_[v] = col_idx_[v + 1] - col_idx_[v];
    gpm_free(po);
    gpm_free(pa);
    gpm_free(perm);
    release_device();
    if (NeedToLoadToDevice) DeviceObject();
  }

  // Graph::DeviceObject (graph.h:342-366): the device copy (a gpm_graph).
  gpm_graph* DeviceObject() {
    if (!dev_) {
      sort_dedup();
      std::vector<uint32_t> lab;
      if (!vlabels_.empty()) lab.assign(vlabels_.begin(), vlabels_.end());
      uint32_t nl = 0;
      for (uint32_t l : lab) nl = std::max(nl, l + 1);
      check_status(gpm_graph_create(offsets_u32().data(), (uint32_t)vsize_, adj_u32().data(),
                                    row_ptr_.size(), lab.empty() ? nullptr : lab.data(), nl,
                                    oriented_ ? GPM_GRAPH_ORIENTED : GPM_GRAPH_SYMMETRIC, &dev_),
                   "Graph::DeviceObject");
    }
    return dev_;
  }

  VID* getNeighbor(VID vid) const {
    return const_cast<VID*>(row_ptr_.data()) + col_idx_[vid];
  }
  VID edge_begin(VID src) { return col_idx_[src]; }
  VID edge_end(VID src) { return col_idx_[src + 1]; }
  VID get_src(VID idx) const { return src_list_[idx]; }
  VID get_dst(VID idx) const { return dst_list_[idx]; }
  VID get_vnum() const { return (VID)vsize_; }
  VID get_enum() const { return (VID)esize_; }
  VID getMaxDegree() { return max_degree_; }
  VID getOutDegree(VID src) { return col_idx_[src + 1] - col_idx_[src]; }
  bool is_oriented() const { return oriented_; }
  const std::vector<VLABEL>& labels() const { return vlabels_; }

  void Dump(std::ostream& out) {
    out << "vsize: " << vsize_ << " esize: " << esize_ << "\n";
    out << "row_ptr: ";
    for (auto x : row_ptr_) out << x << " ";
    out << "\ncol_idx: ";
    for (auto x : col_idx_) out << x << " ";
    out << "\n";
  }

 private:
  std::vector<uint32_t> offsets_u32() const {
    return std::vector<uint32_t>(col_idx_.begin(), col_idx_.end());
  }
  std::vector<uint32_t> adj_u32() const {
    return std::vector<uint32_t>(row_ptr_.begin(), row_ptr_.end());
  }
  void sort_dedup() {
    std::vector<VID> off(vsize_ + 1, 0), adj;
    adj.reserve(row_ptr_.size());
    for (size_t v = 0; v < vsize_; ++v) {
      std::vector<VID> l(row_ptr_.begin() + col_idx_[v], row_ptr_.begin() + col_idx_[v + 1]);
      std::sort(l.begin(), l.end());
      l.erase(std::unique(l.begin(), l.end()), l.end());
      adj.insert(adj.end(), l.begin(), l.end());
      off[v + 1] = (VID)adj.size();
    }
    col_idx_.swap(off);
    row_ptr_.swap(adj);
    esize_ = row_ptr_.size();
  }
  void release_device() {
    if (dev_) gpm_graph_destroy(dev_);
    dev_ = nullptr;
  }

  size_t vsize_ = 0, esize_ = 0;
  VID max_degree_ = 0;
  bool oriented_ = false;
  std::vector<VLABEL> vlabels_;
  std::vector<VID> row_ptr_;  // adjacency entries (reference naming)
  std::vector<VID> col_idx_;  // offsets (reference naming)
  std::vector<VID> odegs_;
  std::vector<VID> src_list_, dst_list_;
  gpm_graph* dev_ = nullptr;
};

}  // namespace project_AntiRF
