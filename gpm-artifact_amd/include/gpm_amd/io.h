// Loader of the reference API (src/graph/io.h:19-153) over gpm_load_mtx.
#pragma once

#include <fstream>
#include <limits>
#include <string>
#include <utility>
#include <vector>

#include "graph.h"

namespace project_AntiRF {

template <typename VID, typename VLABEL>
class Loader {
 public:
  // Loader::Load (io.h:22-33): "mtx" only; other formats abort as in the
  // reference (io.h:142-143).
  void Load(std::string const& path, std::string const& format) {
    if (format == "mtx") {
      uint32_t n = 0, *s = nullptr, *d = nullptr;
      uint64_t ne = 0;
      check_status(gpm_load_mtx(path.c_str(), &n, &s, &d, &ne), "Loader::Load");
      vsize_ = n;
      edges_.clear();
      for (uint64_t i = 0; i < ne; ++i) edges_.emplace_back((VID)s[i], (VID)d[i]);
      gpm_free(s);
      gpm_free(d);
      vlabels_.clear();
    } else if (format == "txt" || format == "bin") {
      fatal("not implemented");
    } else {
      fatal("unknown format " + format);
    }
  }

  // Loader::LoadVLabel (io.h:35-73): "randomK" draws seeded labels in
  // [0, K); otherwise "v label" lines, one per vertex.
  void LoadVLabel(std::string const& path) {
    if (path.rfind("random", 0) == 0) {
      uint32_t* l = nullptr;
      check_status(gpm_gen_labels((uint32_t)vsize_, (uint32_t)std::stoul(path.substr(6)), 1, &l),
                   "Loader::LoadVLabel");
      vlabels_.assign(l, l + vsize_);
      gpm_free(l);
      return;
    }
    std::ifstream fin(path);
    if (!fin.is_open()) fatal("cannot open " + path);
    std::vector<std::pair<long long, long long>> tmp;
    while (fin.good()) {
      long long v, label;
      fin >> v >> label;
      fin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
      if (fin.eof()) break;
      tmp.emplace_back(v, label);
    }
    if (tmp.size() != vsize_) fatal("vertex size not match");
    vlabels_.assign(vsize_, 0);
    long long vmin = std::numeric_limits<long long>::max();
    for (auto& p : tmp) vmin = std::min(vmin, p.first);
    for (auto& p : tmp) vlabels_[p.first - vmin] = (VLABEL)p.second;
  }

  // Loader::Build (io.h:75-81); labels are passed on (the reference drops
  // them by calling Init without use_label).
  void Build(Graph<VID, VLABEL>& hg) {
    std::vector<VID> vids(vsize_);
    for (size_t i = 0; i < vsize_; ++i) vids[i] = (VID)i;
    hg.Init(vids, vlabels_, edges_, !vlabels_.empty());
  }

 private:
  size_t vsize_ = 0;
  std::vector<std::pair<VID, VID>> edges_;
  std::vector<VLABEL> vlabels_;
};

}  // namespace project_AntiRF
