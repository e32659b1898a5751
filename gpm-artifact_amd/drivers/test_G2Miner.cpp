// Driver with the shape of the reference's test/test_G2Miner.cu and
// test/test_patternEnum.cu: pattern count on an undirected mtx graph.
#include <iostream>

#include "gpm_amd/Engine.h"

using namespace project_AntiRF;

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::cerr << "Usage: " << argv[0] << " <graph_path> <pattern_name> [g2miner|patternenum]"
              << std::endl;
    return 1;
  }
  Pattern pattern(argv[2]);
  bool enumerate = argc > 3 && std::string(argv[3]) == "patternenum";
  Loader<int, int> loader;
  Graph<int, int> hg;
  loader.Load(argv[1], "mtx");
  loader.Build(hg);
  Engine<int, int> engine(hg, 1, CF, e_centric, pattern);
  uint64_t result = enumerate ? engine.RunPatternEnum() : engine.RunG2Miner();
  std::cout << "Result: " << result << "\n";
  if (pattern.get_name() == "Motif4") {
    const auto& m = engine.motif_counts();
    for (size_t i = 0; i < m.size(); ++i) std::cout << "Result" << i << ": " << m[i] << "\n";
  }
  return 0;
}
