// Shared host-side helpers of libgpm: thread-local error text, status
// plumbing and the deterministic counter-based RNG used by the generators.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpm.h"

namespace gpm {

// Sets the thread-local error message and returns `status`.
int fail(int status, const char* fmt, ...);
void clear_error();

// splitmix64 finaliser: a bijective 64-bit mix.  Used as a counter-based
// generator (value i of stream s is mix(s + i * golden)) so that generators
// produce the same output for any number of threads.
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Sequential splitmix64 stream.
struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

// 53-bit uniform double in [0, 1).
inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Label-partitioned copy of a sorted CSR: every list reordered by
// (label, id); loff[v * (L + 1) + l] is the absolute start of label l in v's
// list (loff[v * (L + 1) + L] == offsets[v + 1]).
void build_label_csr(uint32_t n, uint32_t L, const uint32_t* offsets,
                     const uint32_t* adj, const uint32_t* labels,
                     std::vector<uint32_t>& ladj, std::vector<uint32_t>& loff);

// Topological relabel of a DAG CSR: rank[v] = position of v in a
// topological order (Kahn's algorithm, ready vertices taken in id order);
// new_off/new_adj is the same DAG over the new ids with sorted lists, so
// every edge goes from a lower to a higher id.  Returns false when the graph
// has a cycle (then the outputs are unspecified).
bool topo_relabel(uint32_t n, const uint32_t* offsets, const uint32_t* adj,
                  std::vector<uint32_t>& rank, std::vector<uint32_t>& new_off,
                  std::vector<uint32_t>& new_adj);

// Runs the body of a C-ABI entry point: a C++ exception must not cross the
// C boundary, so host allocation failures become GPM_ENOMEM and any other
// exception GPM_EINVAL, each with a message.
template <typename F>
int guarded(F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(GPM_ENOMEM, "host allocation failed");
  } catch (const std::length_error&) {
    return fail(GPM_ENOMEM, "host allocation too large");
  } catch (const std::exception& e) {
    return fail(GPM_EINVAL, "%s", e.what());
  } catch (...) {
    return fail(GPM_EINVAL, "unexpected exception");
  }
}

template <typename T>
T* xmalloc(size_t count) {
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = sizeof(T);
  return static_cast<T*>(std::malloc(bytes));
}

}  // namespace gpm

#define GPM_CHECK_ARG(cond, ...)                          \
  do {                                                    \
    if (!(cond)) return ::gpm::fail(GPM_EINVAL, __VA_ARGS__); \
  } while (0)
