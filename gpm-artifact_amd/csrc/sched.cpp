// Host-side pieces of the learned scheduler that are independent of any
// kernel: the heavy-root threshold used to promote expensive small roots to
// a whole block.
#include <algorithm>
#include <cstdint>
#include <vector>

namespace gpm {

// A root whose predicted cost exceeds 1/16384 of the predicted total would,
// run by one wave, outlast the average wave of a full launch (256 CUs x 32
// waves = 8192 waves) by 2x or more; such roots are given a block instead.
float learned_heavy_cost(const float* cost, uint32_t n) {
  if (!cost || n == 0) return 0.f;
  double total = 0.0;
  for (uint32_t v = 0; v < n; ++v) total += std::max(0.f, cost[v]);
  return (float)(total / 16384.0);
}

}  // namespace gpm
