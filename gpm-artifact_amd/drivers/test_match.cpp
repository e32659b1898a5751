// Labeled matching driver, argument-compatible with the reference's CECI
// binary for the options that change the count (-d data graph, -q query
// graph; sub_ceci/matchingcommand.cpp:9-25): several -q may be given and
// are counted in one batch.  "-order learned" selects the cost-based
// matching order.  Prints "#Embeddings: N" per query as myceci.cpp:1115.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gpm_amd/matcher.h"

using namespace project_AntiRF;

int main(int argc, char* argv[]) {
  std::string data;
  std::vector<std::string> qpaths;
  bool learned = false;
  for (int i = 1; i + 1 < argc; i += 2) {
    if (!std::strcmp(argv[i], "-d")) data = argv[i + 1];
    else if (!std::strcmp(argv[i], "-q")) qpaths.push_back(argv[i + 1]);
    else if (!std::strcmp(argv[i], "-order")) learned = !std::strcmp(argv[i + 1], "learned");
  }
  if (data.empty() || qpaths.empty()) {
    std::fprintf(stderr, "Usage: %s -d <data.graph> -q <query.graph> [-q ...] [-order plain|learned]\n",
                 argv[0]);
    return 1;
  }
  LabeledGraph g = LoadLabeledGraph(data);
  std::vector<LabeledGraph> qs;
  for (const auto& p : qpaths) qs.push_back(LoadLabeledGraph(p));
  Matcher m(g);
  const std::vector<uint64_t> counts = m.Count(qs, learned);
  for (size_t i = 0; i < counts.size(); ++i)
    std::printf("#Embeddings: %llu\n", (unsigned long long)counts[i]);
  return 0;
}
