// Labeled subgraph matching over the C ABI: the counting the reference's CPU
// CECI matcher does (sub_ceci/myceci.cpp), with its graph file format
// (sub_ceci/graph/graph.cpp:86-166), for many query graphs per launch.
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "utils.h"

namespace project_AntiRF {

// A labeled graph as read from a sub_ceci text file ("t", "v", "e" lines).
struct LabeledGraph {
  uint32_t n = 0;
  uint32_t n_labels = 0;
  std::vector<uint32_t> offsets, adj, labels;  // symmetric CSR, sorted lists

  // Undirected edges (u < v), each once, as the query descriptor wants them.
  std::vector<uint32_t> edge_pairs() const {
    std::vector<uint32_t> e;
    for (uint32_t u = 0; u < n; ++u)
      for (uint32_t i = offsets[u]; i < offsets[u + 1]; ++i)
        if (u < adj[i]) {
          e.push_back(u);
          e.push_back(adj[i]);
        }
    return e;
  }
};

inline LabeledGraph LoadLabeledGraph(const std::string& path) {
  LabeledGraph g;
  uint32_t *o = nullptr, *a = nullptr, *l = nullptr;
  uint64_t m = 0;
  check_status(gpm_load_labeled(path.c_str(), &g.n, &o, &a, &m, &l, &g.n_labels),
               "LoadLabeledGraph");
  g.offsets.assign(o, o + g.n + 1);
  g.adj.assign(a, a + m);
  g.labels.assign(l, l + g.n);
  gpm_free(o);
  gpm_free(a);
  gpm_free(l);
  return g;
}

// Keeps the data graph resident on the device; Count runs a batch of
// queries in one persistent launch and returns each query's embeddings
// (injective, label- and edge-preserving, non-induced, automorphisms
// counted -- the "#Embeddings" of myceci.cpp:1115).
class Matcher {
 public:
  explicit Matcher(const LabeledGraph& data) {
    check_status(gpm_graph_create(data.offsets.data(), data.n, data.adj.data(), data.adj.size(),
                                  data.labels.data(), std::max(1u, data.n_labels),
                                  GPM_GRAPH_SYMMETRIC, &dev_),
                 "Matcher");
  }
  ~Matcher() {
    if (dev_) gpm_graph_destroy(dev_);
  }
  Matcher(const Matcher&) = delete;
  Matcher& operator=(const Matcher&) = delete;

  std::vector<uint64_t> Count(const std::vector<LabeledGraph>& queries, bool learned_order = false,
                              bool autotune = true) {
    std::vector<std::vector<uint32_t>> edges(queries.size());
    std::vector<gpm_query> qs(queries.size());
    for (size_t i = 0; i < queries.size(); ++i) {
      edges[i] = queries[i].edge_pairs();
      qs[i].n = queries[i].n;
      qs[i].n_edges = (uint32_t)(edges[i].size() / 2);
      qs[i].edges = edges[i].data();
      qs[i].labels = queries[i].labels.data();
    }
    gpm_tuning t{};
    t.autotune = autotune ? 1 : 0;
    t.policy = learned_order ? GPM_SCHED_LEARNED : GPM_SCHED_DEGREE;
    t.use_graph = autotune ? 0 : 1;
    std::vector<uint64_t> out(queries.size(), 0);
    check_status(gpm_match_batch(dev_, qs.data(), (int)qs.size(), &t, out.data()), "Matcher");
    return out;
  }

 private:
  gpm_graph* dev_ = nullptr;
};

}  // namespace project_AntiRF
