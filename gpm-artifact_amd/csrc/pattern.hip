// Fixed-pattern counting on a symmetric graph (replaces G2MinerSolver and
// PatternSolver, src/G2MinerSolver.cuh and src/PatternSolver.cuh).
//
// Direct formulation (GPM_FORM_G2MINER): one wave per directed edge
// (v0, v1), fetched from a work queue in chunks; the nested set operations
// of src/gpu_kernels/G2Miner_kernels.cuh run as wave64 filters (setops.hpp)
// with the intermediate lists in a per-wave region of HBM scratch.
//   P1  = induced 3-star           P2 = induced chair (fork)
//   P3  = 3 claw + 4 K1,4          P4 = induced 4-path (v1 < v0 only)
//   P5  = 2 bull                   P6 = 2 tailed-triangle + 2 (diamond +
//   P7  = 2 K1,1,3                      pendant at a degree-3 vertex)
//   P8  = 3 (K4 + pendant)
//   Motif4 = induced {3-star, 4-path, tailed-triangle, 4-cycle, diamond, K4}
// (the identities are pinned against a brute-force census in the tests).
//
// Counting formulation (GPM_FORM_PATTERN, src/PatternSolver.cuh and
// src/gpu_kernels/new_edge_warp.cuh):
//  * Motif4, P1 (= induced 3-star) and P4 (= induced 4-path) come from exact
//    counting identities instead of enumeration: non-induced counts from
//    degrees, per-edge triangle counts, 4-cycles and 4-cliques, then the
//    fixed inclusion-exclusion over the six connected 4-vertex shapes (the
//    reference's frequency x count - correction carried to closed form).
//  * P2, P3, P7, P8: frequency x count - correction per directed edge
//    (new_edge_warp.cuh:157-578).  The last vertex of the pattern is not
//    enumerated: the pairs (x, v2) the inner loop would visit are counted
//    once, weighted by how many choices of the remaining vertex accept v2
//    (its "frequency"), and the weighted count over-counts exactly the
//    triples where that vertex also fails for x, which a second pass counts
//    and subtracts.  Both sums go to counters 1 and 2; the pattern count is
//    h[1] - h[2] (PatternSolver.cuh:251-254).
//  * P5, P6 (new_edge_warp.cuh:580-918), and P4 under the block prefilter
//    (:920-1082): for each middle vertex, the non-adjacent pairs between two
//    candidate lists are counted from the shorter list, each entry
//    contributing |other| - |other & N(entry)|; the count is h[1].
//  * GPM_FORM_BLOCK_PREFILTER (BLOCK_OPT, PatternSolver.cuh:109-151,
//    new_edge_warp.cuh:647-679): P4, P5 and P6 first test a pair against a
//    64-bit block mask per vertex (bit b set when the vertex has a neighbour
//    in id block b, blocks of 2^shift ids, 64 blocks cover all ids) and
//    binary-search only pairs the mask cannot rule out.
#include <algorithm>
#include <vector>

#include "graph_dev.hpp"
#include "setops.hpp"

using namespace gpm;

namespace {

constexpr int kBlock = 256;
constexpr int kLists = 5;

struct Lists {
  uint32_t* l[kLists];
};

__device__ __forceinline__ void wave_fence() { __threadfence_block(); }

template <int PID>
__device__ __forceinline__ void edge_work(const uint32_t* __restrict__ off,
                                          const uint32_t* __restrict__ adj, uint32_t v0,
                                          uint32_t v1, Lists L, uint64_t* c) {
  const uint32_t* N0 = adj + off[v0];
  const uint32_t d0 = off[v0 + 1] - off[v0];
  const uint32_t* N1 = adj + off[v1];
  const uint32_t d1 = off[v1 + 1] - off[v1];
  auto N = [&](uint32_t v) { return adj + off[v]; };
  auto D = [&](uint32_t v) { return off[v + 1] - off[v]; };
  uint32_t *L0 = L.l[0], *L1 = L.l[1], *L2 = L.l[2], *L3 = L.l[3], *L4 = L.l[4];
  if (PID == 1) {  // G2Miner_kernels.cuh:22-42
    const uint32_t n0 = wave_diff(N0, d0, N1, d1, v1, L0);
    wave_fence();
    for (uint32_t i = 0; i < n0; ++i) {
      const uint32_t v2 = L0[i];
      c[0] += wave_diff_count(L0, n0, N(v2), D(v2), v2);
    }
  } else if (PID == 2) {  // :68-103
    const uint32_t n0 = wave_diff(N0, d0, N1, d1, kNoBound, L0);
    const uint32_t n1 = wave_diff(N1, d1, N0, d0, kNoBound, L1);
    wave_fence();
    for (uint32_t i = 0; i < n1; ++i) {
      const uint32_t v4 = L1[i];
      const uint32_t n2 = wave_diff(L0, n0, N(v4), D(v4), kNoBound, L2);
      wave_fence();
      for (uint32_t j = 0; j < n2; ++j) {
        const uint32_t v2 = L2[j];
        c[0] += wave_diff_count(L2, n2, N(v2), D(v2), v2);
      }
      wave_fence();
    }
  } else if (PID == 3) {  // :126-157
    const uint32_t n0 = wave_diff(N0, d0, N1, d1, kNoBound, L0);
    wave_fence();
    for (uint32_t i = 0; i < n0; ++i) {
      const uint32_t v2 = L0[i];
      const uint32_t n1 = wave_diff(L0, n0, N(v2), D(v2), v2, L1);
      wave_fence();
      for (uint32_t j = 0; j < n1; ++j) {
        const uint32_t v3 = L1[j];
        c[0] += wave_diff_count(L1, n1, N(v3), D(v3), v3);
      }
      wave_fence();
    }
  } else if (PID == 4) {  // :174-197
    if (v1 >= v0) return;
    const uint32_t n0 = wave_diff(N0, d0, N1, d1, kNoBound, L0);
    const uint32_t n1 = wave_diff(N1, d1, N0, d0, kNoBound, L1);
    wave_fence();
    for (uint32_t j = 0; j < n0; ++j) {
      const uint32_t v2 = L0[j];
      c[0] += wave_diff_count(L1, n1, N(v2), D(v2), kNoBound);
    }
  } else if (PID == 5) {  // :219-259
    const uint32_t n0 = wave_diff(N0, d0, N1, d1, kNoBound, L0);
    const uint32_t n1 = wave_diff(N1, d1, N0, d0, kNoBound, L1);
    const uint32_t n2 = wave_isect(N0, d0, N1, d1, kNoBound, L2);
    wave_fence();
    for (uint32_t i = 0; i < n2; ++i) {
      const uint32_t v2 = L2[i];
      const uint32_t n3 = wave_diff(L0, n0, N(v2), D(v2), kNoBound, L3);
      const uint32_t n4 = wave_diff(L1, n1, N(v2), D(v2), kNoBound, L4);
      wave_fence();
      for (uint32_t j = 0; j < n3; ++j) {
        const uint32_t v3 = L3[j];
        c[0] += wave_diff_count(L4, n4, N(v3), D(v3), kNoBound);
      }
      wave_fence();
    }
  } else if (PID == 6) {  // :281-316
    const uint32_t n0 = wave_isect(N0, d0, N1, d1, kNoBound, L0);
    const uint32_t n1 = wave_diff(N1, d1, N0, d0, kNoBound, L1);
    wave_fence();
    for (uint32_t i = 0; i < n0; ++i) {
      const uint32_t v2 = L0[i];
      const uint32_t n2 = wave_diff(L0, n0, N(v2), D(v2), kNoBound, L2);
      const uint32_t n3 = wave_diff(L1, n1, N(v2), D(v2), kNoBound, L3);
      wave_fence();
      for (uint32_t j = 0; j < n2; ++j) {
        const uint32_t v3 = L2[j];
        c[0] += wave_diff_count(L3, n3, N(v3), D(v3), kNoBound);
      }
      wave_fence();
    }
  } else if (PID == 7) {  // :339-370
    const uint32_t n0 = wave_isect(N0, d0, N1, d1, kNoBound, L0);
    wave_fence();
    for (uint32_t i = 0; i < n0; ++i) {
      const uint32_t v2 = L0[i];
      const uint32_t n1 = wave_diff(L0, n0, N(v2), D(v2), v2, L1);
      wave_fence();
      for (uint32_t j = 0; j < n1; ++j) {
        const uint32_t v3 = L1[j];
        c[0] += wave_diff_count(L1, n1, N(v3), D(v3), v3);
      }
      wave_fence();
    }
  } else if (PID == 8) {  // :395-429
    const uint32_t n0 = wave_isect(N0, d0, N1, d1, kNoBound, L0);
    const uint32_t n1 = wave_diff(N1, d1, N0, d0, kNoBound, L1);
    wave_fence();
    for (uint32_t i = 0; i < n1; ++i) {
      const uint32_t v4 = L1[i];
      const uint32_t n2 = wave_diff(L0, n0, N(v4), D(v4), kNoBound, L2);
      wave_fence();
      for (uint32_t j = 0; j < n2; ++j) {
        const uint32_t v3 = L2[j];
        c[0] += wave_isect_count(L2, n2, N(v3), D(v3), v3);
      }
      wave_fence();
    }
  } else if (PID == 9) {  // Motif4_G2Miner, :454-548
    // 3-star over every directed edge.
    const uint32_t nl = wave_diff(N0, d0, N1, d1, v1, L0);
    wave_fence();
    for (uint32_t i = 0; i < nl; ++i) {
      const uint32_t v2 = L0[i];
      c[0] += wave_diff_count(L0, nl, N(v2), D(v2), v2);
    }
    if (v1 > v0) return;
    // Diamond and tailed-triangle over I = N(v0) & N(v1).
    const uint32_t ni = wave_isect(N0, d0, N1, d1, kNoBound, L1);
    wave_fence();
    for (uint32_t i = 0; i < ni; ++i) {
      const uint32_t v2 = L1[i];
      c[4] += wave_diff_count(L1, ni, N(v2), D(v2), v2);
      const uint32_t nt = wave_diff(N(v2), D(v2), N0, d0, kNoBound, L2);
      wave_fence();
      c[2] += wave_diff_count(L2, nt, N1, d1, kNoBound);
      wave_fence();
    }
    // 4-clique over {x in I, x < v1}.
    const uint32_t nb = wave_isect(N0, d0, N1, d1, v1, L2);
    wave_fence();
    for (uint32_t i = 0; i < nb; ++i) {
      const uint32_t v2 = L2[i];
      c[5] += wave_isect_count(L2, nb, N(v2), D(v2), v2);
    }
    wave_fence();
    // 4-path.
    const uint32_t na = wave_diff(N0, d0, N1, d1, kNoBound, L3);
    const uint32_t nq = wave_diff(N1, d1, N0, d0, kNoBound, L4);
    wave_fence();
    for (uint32_t i = 0; i < na; ++i) {
      const uint32_t v2 = L3[i];
      c[1] += wave_diff_count(L4, nq, N(v2), D(v2), kNoBound);
    }
    wave_fence();
    // 4-cycle: A = {x in N(v1)\N(v0), x < v0}, B = {y in N(v0)\N(v1), y < v1}.
    const uint32_t nA = wave_diff(N1, d1, N0, d0, v0, L3);
    const uint32_t nB = wave_diff(N0, d0, N1, d1, v1, L4);
    wave_fence();
    for (uint32_t i = 0; i < nB; ++i) {
      const uint32_t v2 = L4[i];
      c[3] += wave_isect_count(L3, nA, N(v2), D(v2), v0);
    }
    wave_fence();
  }
}

// Block prefilter of the counting formulation: bm[v] has bit (w >> shift)
// set for every neighbour w of v.
struct BlockMask {
  const unsigned long long* bm;  // nullptr: no prefilter
  uint32_t shift;
};

// Pairs (a in A, b in B) with a and b not adjacent, counted from the shorter
// list: each entry s of it adds |other| - |other & N(s)|.  With a block
// prefilter and a short other list, a pair is counted as non-adjacent
// outright when s has no neighbour in b's id block, and binary-searched
// otherwise (new_edge_warp.cuh:647-679).  Wave-uniform result.
__device__ __forceinline__ uint64_t nonadjacent_pairs(const uint32_t* __restrict__ off,
                                                      const uint32_t* __restrict__ adj,
                                                      const uint32_t* A, uint32_t na,
                                                      const uint32_t* B, uint32_t nb,
                                                      BlockMask blk) {
  if (na > nb) {
    const uint32_t* t = A;
    A = B;
    B = t;
    const uint32_t tn = na;
    na = nb;
    nb = tn;
  }
  uint64_t c = 0;
  const uint32_t lane = lane_id();
  for (uint32_t i = 0; i < na; ++i) {
    const uint32_t s = uniform(A[i]);
    const uint32_t* Ns = adj + off[s];
    const uint32_t ds = uniform(off[s + 1] - off[s]);
    if (blk.bm && nb < ds) {
      const unsigned long long mask = blk.bm[s];
      uint32_t k = 0;
      for (uint32_t j = lane; j < nb; j += kWave) {
        const uint32_t b = B[j];
        k += ((mask >> (b >> blk.shift)) & 1ull) == 0ull ? 1u : (sorted_contains(Ns, ds, b) ? 0u : 1u);
      }
      c += wave_sum_u32(k);
    } else {
      c += nb - wave_isect_count(B, nb, Ns, ds, kNoBound);
    }
  }
  return c;
}

// Counts of x in L (sorted, n entries) around position i, against N(v):
// below = |{x in L[0..i) : x not in N(v)}|, above = |{x in L(i..n) : x not
// in N(v)}| (KEEP_IN = false) or the same with "in N(v)" (KEEP_IN = true).
template <bool KEEP_IN>
__device__ __forceinline__ void split_counts(const uint32_t* L, uint32_t n, uint32_t i,
                                             const uint32_t* Nv, uint32_t dv, uint32_t* below,
                                             uint32_t* above) {
  const uint32_t lane = lane_id();
  uint32_t lo = 0, hi = 0;
  for (uint32_t base = 0; base < n; base += kWave) {
    const uint32_t j = base + lane;
    if (j < n && j != i) {
      const bool hit = sorted_contains(Nv, dv, L[j]) == KEEP_IN;
      if (hit) (j < i ? lo : hi) += 1;
    }
  }
  *below = wave_sum_u32(lo);
  *above = wave_sum_u32(hi);
}

// Correction pass shared by P2, P3, P7, P8: for every w in W, the entries
// key of K with key > w whose adjacency to w is `ADJ` (K sorted).
template <bool ADJ>
__device__ __forceinline__ uint64_t correction(const uint32_t* __restrict__ off,
                                               const uint32_t* __restrict__ adj, const uint32_t* W,
                                               uint32_t nw, const uint32_t* K, uint32_t nk) {
  uint64_t c = 0;
  for (uint32_t t = 0; t < nw; ++t) {
    const uint32_t w = uniform(W[t]);
    const uint32_t s = uniform(upper_index(K, nk, w));
    if (s >= nk) break;  // W is sorted: later w have no larger keys either
    const uint32_t* Nw = adj + off[w];
    const uint32_t dw = off[w + 1] - off[w];
    c += ADJ ? wave_isect_count(K + s, nk - s, Nw, dw, kNoBound)
             : wave_diff_count(K + s, nk - s, Nw, dw, kNoBound);
  }
  return c;
}

// Counting formulation per directed edge (v0, v1); c[1] += frequency-weighted
// counts, c[2] += corrections (P2, P3, P7, P8), or c[1] += pair counts (P4,
// P5, P6).
template <int PID>
__device__ __forceinline__ void edge_work_counting(const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ adj, uint32_t v0,
                                                   uint32_t v1, Lists L, BlockMask blk,
                                                   uint64_t* c) {
  const uint32_t* N0 = adj + off[v0];
  const uint32_t d0 = off[v0 + 1] - off[v0];
  const uint32_t* N1 = adj + off[v1];
  const uint32_t d1 = off[v1 + 1] - off[v1];
  auto N = [&](uint32_t v) { return adj + off[v]; };
  auto D = [&](uint32_t v) { return off[v + 1] - off[v]; };
  uint32_t *L0 = L.l[0], *L1 = L.l[1], *L2 = L.l[2], *L3 = L.l[3];
  if (PID == 2 || PID == 8) {
    // P2 (:157-262): L0 = N(v0)\N(v1), L1 = N(v1)\N(v0); the frequency of v2
    //   in L0 is #{v4 in L1 : v2 not adj v4}; weighted pairs (x < v2, x, v2
    //   in L0, x not adj v2).  Correction: v4 in L1, x in L0 & N(v4),
    //   key in L0 \ N(v4), key > x, key not adj x.
    // P8 (:471-578): L0 = N(v0) & N(v1) instead, the weighted pairs are
    //   adjacent (x < v2, x adj v2), and the correction counts key adj x.
    constexpr bool kAdj = PID == 8;
    const uint32_t n0 = kAdj ? wave_isect(N0, d0, N1, d1, kNoBound, L0)
                             : wave_diff(N0, d0, N1, d1, kNoBound, L0);
    const uint32_t n1 = wave_diff(N1, d1, N0, d0, kNoBound, L1);
    wave_fence();
    for (uint32_t i = 0; i < n0; ++i) {
      const uint32_t v2 = uniform(L0[i]);
      const uint32_t freq = n1 - wave_isect_count(L1, n1, N(v2), D(v2), kNoBound);
      if (!freq) continue;
      uint32_t below, above;
      split_counts<kAdj>(L0, n0, i, N(v2), D(v2), &below, &above);
      c[1] += (uint64_t)freq * below;
    }
    for (uint32_t i = 0; i < n1; ++i) {
      const uint32_t v4 = uniform(L1[i]);
      const uint32_t n2 = wave_diff(L0, n0, N(v4), D(v4), kNoBound, L2);
      const uint32_t n3 = wave_isect(L0, n0, N(v4), D(v4), kNoBound, L3);
      wave_fence();
      c[2] += correction<kAdj>(off, adj, L3, n3, L2, n2);
      wave_fence();
    }
  } else if (PID == 3 || PID == 7) {
    // P3 (:264-364): L0 = N(v0)\N(v1); for v2 in L0, (x below v2, y above
    //   v2) with x, y not adj v2 are counted as below x above; the triples
    //   where x adj y are counted again as (w < key < m, w adj m, key not
    //   adj m, key not adj w) and subtracted.
    // P7 (:367-469): the same over L0 = N(v0) & N(v1).
    const uint32_t n0 = PID == 7 ? wave_isect(N0, d0, N1, d1, kNoBound, L0)
                                 : wave_diff(N0, d0, N1, d1, kNoBound, L0);
    wave_fence();
    for (uint32_t i = 0; i < n0; ++i) {
      const uint32_t v2 = uniform(L0[i]);
      uint32_t below, above;
      split_counts<false>(L0, n0, i, N(v2), D(v2), &below, &above);
      c[1] += (uint64_t)below * above;
      const uint32_t n1 = wave_diff(L0, n0, N(v2), D(v2), v2, L1);  // key < m, not adj m
      const uint32_t n2 = wave_isect(L0, n0, N(v2), D(v2), kNoBound, L2);  // w adj m
      wave_fence();
      c[2] += correction<false>(off, adj, L2, n2, L1, n1);
      wave_fence();
    }
  } else if (PID == 4) {
    // P4 (:920-1082): v1 < v0 only; non-adjacent pairs between
    // N(v0)\N(v1) and N(v1)\N(v0).
    if (v1 >= v0) return;
    const uint32_t n0 = wave_diff(N0, d0, N1, d1, kNoBound, L0);
    const uint32_t n1 = wave_diff(N1, d1, N0, d0, kNoBound, L1);
    wave_fence();
    c[1] += nonadjacent_pairs(off, adj, L0, n0, L1, n1, blk);
  } else if (PID == 5 || PID == 6) {
    // P5 (:580-757): for v2 in I = N(v0) & N(v1), non-adjacent pairs between
    //   (N(v0)\N(v1))\N(v2) and (N(v1)\N(v0))\N(v2).
    // P6 (:759-918): between I\N(v2) and (N(v1)\N(v0))\N(v2).
    const uint32_t ni = wave_isect(N0, d0, N1, d1, kNoBound, L0);
    const uint32_t na = PID == 5 ? wave_diff(N0, d0, N1, d1, kNoBound, L1) : 0u;
    const uint32_t nb = wave_diff(N1, d1, N0, d0, kNoBound, L2);
    wave_fence();
    uint32_t* A = L.l[3];
    uint32_t* B = L.l[4];
    for (uint32_t i = 0; i < ni; ++i) {
      const uint32_t v2 = uniform(L0[i]);
      const uint32_t n3 = PID == 5 ? wave_diff(L1, na, N(v2), D(v2), kNoBound, A)
                                   : wave_diff(L0, ni, N(v2), D(v2), kNoBound, A);
      const uint32_t n4 = wave_diff(L2, nb, N(v2), D(v2), kNoBound, B);
      wave_fence();
      c[1] += nonadjacent_pairs(off, adj, A, n3, B, n4, blk);
      wave_fence();
    }
  }
}

template <int PID, int FORM>
__global__ __launch_bounds__(kBlock) void pattern_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ src, uint64_t m, uint32_t chunk,
    unsigned long long* __restrict__ queue, uint32_t* __restrict__ scratch, uint64_t stride,
    BlockMask blk, unsigned long long* __restrict__ out) {
  const uint32_t lane = lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  Lists L;
  for (int i = 0; i < kLists; ++i) L.l[i] = scratch + (wave * kLists + i) * stride;
  uint64_t c[6] = {0, 0, 0, 0, 0, 0};
  while (true) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(queue, (unsigned long long)chunk);
    base = __shfl(base, 0, kWave);
    if (base >= m) break;
    const uint64_t stop = min((unsigned long long)m, base + chunk);
    for (uint64_t e = base; e < stop; ++e) {
      const uint32_t v0 = uniform(src[e]), v1 = uniform(adj[e]);
      if (v0 == v1) continue;
      if (FORM == GPM_FORM_G2MINER)
        edge_work<PID>(off, adj, v0, v1, L, c);
      else
        edge_work_counting<PID>(off, adj, v0, v1, L, blk, c);
    }
  }
  if (lane == 0)
    for (int i = 0; i < 6; ++i)
      if (c[i]) atomicAdd(&out[i], (unsigned long long)c[i]);
}

// bm[v] = OR over neighbours w of v of bit (w >> shift); one wave per vertex.
__global__ void block_mask_kernel(const uint32_t* __restrict__ off,
                                  const uint32_t* __restrict__ adj, uint32_t n, uint32_t shift,
                                  unsigned long long* __restrict__ bm) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const uint32_t lane = lane_id();
  for (uint32_t v = wave; v < n; v += nwaves) {
    unsigned long long m = 0;
    for (uint32_t i = off[v] + lane; i < off[v + 1]; i += kWave) m |= 1ull << (adj[i] >> shift);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m |= __shfl_xor(m, o, kWave);
    if (lane == 0) bm[v] = m;
  }
}

__global__ void zero_u64_kernel(unsigned long long* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0ull;
}

// Builds g->bmask (64 id blocks of 2^shift ids cover [0, n)) if missing.
int ensure_block_mask(gpm_graph* g) {
  if (g->bmask) return GPM_OK;
  uint32_t shift = 0;
  while (((uint64_t)std::max<uint32_t>(g->n, 1) - 1) >> shift >= 64) ++shift;
  if (hipMalloc(&g->bmask, (size_t)std::max<uint32_t>(g->n, 1) * 8) != hipSuccess) {
    g->bmask = nullptr;
    return fail(GPM_ENOMEM, "block mask (%u vertices)", g->n);
  }
  g->bmask_shift = shift;
  const unsigned blocks = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>(((uint64_t)g->n * kWave + kBlock - 1) / kBlock,
                            (uint64_t)device_cus(g) * 16));
  hipLaunchKernelGGL(block_mask_kernel, dim3(blocks), dim3(kBlock), 0, g->stream, g->off, g->adj,
                     g->n, shift, g->bmask);
  GPM_HIP(hipGetLastError());
  GPM_HIP(hipStreamSynchronize(g->stream));
  return GPM_OK;
}

template <int PID, int FORM>
int launch_pattern(gpm_graph* g, BlockMask blk, unsigned long long* out) {
  int st = ensure_src(g);
  if (st != GPM_OK) return st;
  const int cus = device_cus(g);
  // Waves resident at once bound the scratch: kLists lists of max_deg+1
  // entries per wave.
  const uint64_t stride = (uint64_t)g->max_deg + 1;
  uint64_t waves = std::min<uint64_t>((uint64_t)cus * 16, std::max<uint64_t>(1, g->m));
  const uint64_t cap_bytes = 4ull << 30;  // at most 4 GiB of list scratch
  while (waves > 64 && waves * kLists * stride * 4 > cap_bytes) waves /= 2;
  const unsigned blocks = (unsigned)((waves * kWave + kBlock - 1) / kBlock);
  const uint64_t nwaves = (uint64_t)blocks * kBlock / kWave;
  const size_t list_bytes = nwaves * kLists * stride * 4;
  if ((st = ensure_scratch(g, list_bytes + 64)) != GPM_OK) return st;
  uint32_t* lists = static_cast<uint32_t*>(g->scratch);
  unsigned long long* queue = g->counters + 8;
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, g->stream, queue, 1);
  const uint32_t chunk = 4;
  hipLaunchKernelGGL((pattern_kernel<PID, FORM>), dim3(blocks), dim3(kBlock), 0, g->stream,
                     g->off, g->adj, g->src, g->m, chunk, queue, lists, stride, blk, out);
  GPM_HIP(hipGetLastError());
  timing_mark(g, FORM == GPM_FORM_G2MINER ? "pattern_direct" : "pattern_counting");
  return GPM_OK;
}

template <int FORM>
int launch_pid(gpm_graph* g, int pid, BlockMask blk, unsigned long long* out) {
  switch (pid) {
    case 1: return launch_pattern<1, FORM>(g, blk, out);
    case 2: return launch_pattern<2, FORM>(g, blk, out);
    case 3: return launch_pattern<3, FORM>(g, blk, out);
    case 4: return launch_pattern<4, FORM>(g, blk, out);
    case 5: return launch_pattern<5, FORM>(g, blk, out);
    case 6: return launch_pattern<6, FORM>(g, blk, out);
    case 7: return launch_pattern<7, FORM>(g, blk, out);
    case 8: return launch_pattern<8, FORM>(g, blk, out);
    default: return launch_pattern<9, FORM>(g, blk, out);
  }
}

}  // namespace

namespace gpm {
// Counting formulation of Motif4 (motif_count.hip): induced counts of the six
// connected 4-vertex shapes from non-induced counts.
int motif4_by_identities(gpm_graph* g, uint64_t out[6]);
}  // namespace gpm

extern "C" int gpm_pattern(gpm_graph* g, int pattern_id, int formulation, uint64_t out[6],
                           uint64_t* scalar) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g && out && scalar, "null pointer");
    GPM_CHECK_ARG(g->flags & GPM_GRAPH_SYMMETRIC, "pattern counting needs a SYMMETRIC graph");
    GPM_CHECK_ARG(pattern_id >= GPM_P1 && pattern_id <= GPM_MOTIF4, "%d not support!", pattern_id);
    GPM_CHECK_ARG(formulation == GPM_FORM_G2MINER || formulation == GPM_FORM_PATTERN ||
                      formulation == GPM_FORM_BLOCK_PREFILTER,
                  "unknown formulation %d", formulation);
    int st;
    uint64_t h[6] = {0, 0, 0, 0, 0, 0};
    const bool counting = formulation != GPM_FORM_G2MINER;
    const bool prefilter = formulation == GPM_FORM_BLOCK_PREFILTER &&
                           (pattern_id == GPM_P4 || pattern_id == GPM_P5 || pattern_id == GPM_P6);
    const bool identities =
        counting && (pattern_id == GPM_MOTIF4 || pattern_id == GPM_P1 ||
                     (pattern_id == GPM_P4 && !prefilter));
    // Counting kernels whose result is h[1] alone (PatternSolver.cuh:251-254).
    const bool h1_only = pattern_id == GPM_P4 || pattern_id == GPM_P5 || pattern_id == GPM_P6;
    if (identities) {
      uint64_t m4[6];
      if ((st = motif4_by_identities(g, m4)) != GPM_OK) return st;
      if (pattern_id == GPM_MOTIF4)
        for (int i = 0; i < 6; ++i) h[i] = m4[i];
      else
        h[0] = pattern_id == GPM_P1 ? m4[0] : m4[1];
    } else {
      BlockMask blk{nullptr, 0};
      if (prefilter) {
        if ((st = ensure_block_mask(g)) != GPM_OK) return st;
        blk = BlockMask{g->bmask, g->bmask_shift};
      }
      if ((st = timing_begin(g)) != GPM_OK) return st;
      if ((st = zero_counters(g, 8)) != GPM_OK) return st;
      st = counting ? launch_pid<GPM_FORM_PATTERN>(g, pattern_id, blk, g->counters)
                    : launch_pid<GPM_FORM_G2MINER>(g, pattern_id, blk, g->counters);
      if (st != GPM_OK) return st;
      if ((st = timing_end(g)) != GPM_OK) return st;
      unsigned long long hh[6];
      GPM_HIP(hipMemcpyAsync(hh, g->counters, sizeof(hh), hipMemcpyDeviceToHost, g->stream));
      GPM_HIP(hipStreamSynchronize(g->stream));
      for (int i = 0; i < 6; ++i) h[i] = hh[i];
      if (counting) h[0] = h1_only ? h[1] : h[1] - h[2];  // modulo 2^64, exact when >= 0
    }
    for (int i = 0; i < 6; ++i) out[i] = h[i];
    if (pattern_id == GPM_MOTIF4) {
      // G2MinerSolver returns counters[0] (3-star, G2MinerSolver.cuh:166);
      // PatternSolver returns h[1] - h[2] (PatternSolver.cuh:251-254), which
      // after its Motif4 sequence is 4-path minus tailed-triangle.
      *scalar = formulation == GPM_FORM_G2MINER ? h[0] : h[1] - h[2];
    } else {
      *scalar = h[0];
    }
    return GPM_OK;
  });
}
