// Device-side building blocks for gfx950 (CDNA4, wave64): lane/mask helpers
// built on 64-bit ballots and mbcnt, wave reductions, sorted-set searches and
// the work-queue fetch.  Every kernel in libgpm is written against these.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

#define GPM_HIP(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::gpm::fail(GPM_EHIP, "%s failed: %s (%s:%d)", #call,            \
                         hipGetErrorString(e_), __FILE__, __LINE__);          \
  } while (0)

namespace gpm {

constexpr int kWave = 64;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Atomics on global memory that only the calling workgroup reads or writes
// (block-private scratch), at workgroup scope, to be ordered with
// __threadfence_block() rather than __threadfence(): a device-scope fence per
// Motif4 root cost a quarter of the 4-cycle kernel's time (1.85 s against
// 1.39 s on R-MAT scale 22).  The atomics themselves still reach the memory
// side at either scope (every L2 atomic is forwarded in the counters).
__device__ __forceinline__ uint32_t wg_atomic_add(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void wg_atomic_or(unsigned long long* p, unsigned long long v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Number of set bits of `mask` strictly below this lane (v_mbcnt_lo/hi).
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// Declares a value the same in every lane (moves it to an SGPR).  Loop
// bounds of loops that contain __syncthreads() go through this so that the
// compiler emits them as scalar, wave-uniform loops.  Bounds loaded from
// LDS or global memory without it are treated as possibly divergent; the
// block kernels then gave nondeterministic counts once two blocks shared a
// compute unit, and exact counts with it.
__device__ __forceinline__ uint32_t uniform(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint32_t wave_bcast(uint32_t v, int src) {
  return __shfl(v, src, kWave);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Inclusive prefix sum across the wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

// Branch-free membership test in a sorted list (lower-bound walk).
__device__ __forceinline__ bool sorted_contains(const uint32_t* __restrict__ a,
                                                uint32_t n, uint32_t key) {
  if (n == 0) return false;
  uint32_t lo = 0;
  while (n > 1) {
    uint32_t half = n >> 1;
    lo = (a[lo + half] <= key) ? lo + half : lo;
    n -= half;
  }
  return a[lo] == key;
}

// Index of the first element > key in a sorted LDS/global array of length n.
__device__ __forceinline__ uint32_t upper_index(const uint32_t* a, uint32_t n,
                                                uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Multiplicative hash into a power-of-two table.
__device__ __forceinline__ uint32_t hslot(uint32_t key, uint32_t mask) {
  return ((key * 0x9E3779B1u) >> 16) & mask;
}

// Insert `key` (unique) into an open-addressed LDS table; stores `val` in the
// parallel value array.  The table must have room (load factor <= 1/2).
__device__ __forceinline__ void lds_hash_insert(uint32_t* keys, uint32_t* vals,
                                                uint32_t mask, uint32_t key,
                                                uint32_t val) {
  uint32_t h = hslot(key, mask);
  while (true) {
    uint32_t prev = atomicCAS(&keys[h], kEmpty, key);
    if (prev == kEmpty) {
      vals[h] = val;
      return;
    }
    h = (h + 1) & mask;
  }
}

// Returns the stored value, or -1 when absent.
__device__ __forceinline__ int lds_hash_find(const uint32_t* keys,
                                             const uint32_t* vals,
                                             uint32_t mask, uint32_t key) {
  uint32_t h = hslot(key, mask);
  while (true) {
    uint32_t k = keys[h];
    if (k == key) return vals[h];
    if (k == kEmpty) return -1;
    h = (h + 1) & mask;
  }
}

__device__ __forceinline__ bool lds_hash_has(const uint32_t* keys, uint32_t mask,
                                             uint32_t key) {
  uint32_t h = hslot(key, mask);
  while (true) {
    uint32_t k = keys[h];
    if (k == key) return true;
    if (k == kEmpty) return false;
    h = (h + 1) & mask;
  }
}

}  // namespace gpm
