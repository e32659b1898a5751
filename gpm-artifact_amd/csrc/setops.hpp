// Wave-cooperative sorted-set operations on lists in HBM (the primitives of
// src/graph/set_intersection.cuh and set_difference.cuh, rewritten for
// wave64).  The iterated list is walked 64 entries per step; each lane
// binary-searches its entry in the other list; kept entries are compacted
// with a 64-bit ballot and v_mbcnt (no shared counters, no __syncwarp).
// Every function returns a wave-uniform result.
//
// Semantics, as in the reference:
//   filter<KEEP_IN>(A, B, ub): {x in A : x < ub and (x in B) == KEEP_IN}
//     KEEP_IN = false is A \ B (A iterated, never swapped:
//     set_difference.cuh:98-234); KEEP_IN = true is A & B.
//   An upper bound ub stops the walk at the first entry >= ub (lists are
//   sorted), which is the early exit of set_intersection.cuh:99-110.
#pragma once

#include "device_common.hpp"

namespace gpm {

constexpr uint32_t kNoBound = 0xFFFFFFFFu;

// Materialising filter; returns |out|.
template <bool KEEP_IN>
__device__ __forceinline__ uint32_t wave_filter(const uint32_t* __restrict__ A, uint32_t na,
                                                const uint32_t* __restrict__ B, uint32_t nb,
                                                uint32_t ub, uint32_t* __restrict__ out) {
  const uint32_t lane = lane_id();
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < na; base += kWave) {
    const uint32_t i = base + lane;
    const bool valid = i < na;
    const uint32_t x = valid ? A[i] : kNoBound;
    const bool below = valid && x < ub;
    const bool keep = below && (sorted_contains(B, nb, x) == KEEP_IN);
    const uint64_t m = wave_ballot(keep);
    if (keep) out[cnt + lanes_below(m)] = x;
    cnt += (uint32_t)__popcll(m);
    if (wave_ballot(valid && !below)) break;
  }
  return cnt;
}

// Counting filter; returns the count.
template <bool KEEP_IN>
__device__ __forceinline__ uint32_t wave_filter_count(const uint32_t* __restrict__ A,
                                                      uint32_t na,
                                                      const uint32_t* __restrict__ B,
                                                      uint32_t nb, uint32_t ub) {
  const uint32_t lane = lane_id();
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < na; base += kWave) {
    const uint32_t i = base + lane;
    const bool valid = i < na;
    const uint32_t x = valid ? A[i] : kNoBound;
    const bool below = valid && x < ub;
    const bool keep = below && (sorted_contains(B, nb, x) == KEEP_IN);
    cnt += (uint32_t)__popcll(wave_ballot(keep));
    if (wave_ballot(valid && !below)) break;
  }
  return cnt;
}

// |A & B| restricted to x < ub, iterating the shorter list.
__device__ __forceinline__ uint32_t wave_isect_count(const uint32_t* A, uint32_t na,
                                                     const uint32_t* B, uint32_t nb,
                                                     uint32_t ub) {
  return na <= nb ? wave_filter_count<true>(A, na, B, nb, ub)
                  : wave_filter_count<true>(B, nb, A, na, ub);
}

// A & B materialised (sorted), restricted to x < ub, iterating the shorter.
__device__ __forceinline__ uint32_t wave_isect(const uint32_t* A, uint32_t na, const uint32_t* B,
                                               uint32_t nb, uint32_t ub, uint32_t* out) {
  return na <= nb ? wave_filter<true>(A, na, B, nb, ub, out)
                  : wave_filter<true>(B, nb, A, na, ub, out);
}

__device__ __forceinline__ uint32_t wave_diff(const uint32_t* A, uint32_t na, const uint32_t* B,
                                              uint32_t nb, uint32_t ub, uint32_t* out) {
  return wave_filter<false>(A, na, B, nb, ub, out);
}

__device__ __forceinline__ uint32_t wave_diff_count(const uint32_t* A, uint32_t na,
                                                    const uint32_t* B, uint32_t nb,
                                                    uint32_t ub) {
  return wave_filter_count<false>(A, na, B, nb, ub);
}

}  // namespace gpm
