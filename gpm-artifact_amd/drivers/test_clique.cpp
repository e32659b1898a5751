// k-clique driver (the reference has CFSolver but no driver for it).
#include <cstdlib>
#include <iostream>

#include "gpm_amd/Engine.h"

using namespace project_AntiRF;

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::cerr << "Usage: " << argv[0] << " <graph_path> <k>" << std::endl;
    return 1;
  }
  int k = std::atoi(argv[2]);
  Loader<int, int> loader;
  Graph<int, int> hg;
  loader.Load(argv[1], "mtx");
  loader.Build(hg);
  Engine<int, int> engine(hg, 1, CF, e_centric, k);
  std::cout << "Result: " << engine.RunCF() << "\n";
  return 0;
}
