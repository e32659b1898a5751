// Driver with the shape of the reference's test/test_trianglecounting.cu:
// load an mtx graph, vertex-centric triangle count through Engine::RunTC.
#include <iostream>

#include "gpm_amd/Engine.h"

using namespace project_AntiRF;

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::cerr << "Usage: " << argv[0] << " <graph_path> [v_centric|e_centric]" << std::endl;
    return 1;
  }
  modes cal_mode = (argc > 2 && std::string(argv[2]) == "e_centric") ? e_centric : v_centric;
  Loader<int, int> loader;
  Graph<int, int> hg;
  loader.Load(argv[1], "mtx");
  loader.Build(hg);
  Engine<int, int> engine(hg, 1, TC, cal_mode);
  std::cout << "Pattern 0: " << engine.RunTC() << "\n";
  return 0;
}
