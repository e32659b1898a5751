// k-clique counting (3 <= k <= 5) on a degree-oriented DAG (replaces
// src/CFSolver.cuh).
//
// The reference runs one 32-lane warp per oriented edge (v0, v1) and
// materialises W1 = N+(v0) & N+(v1), W2 = W1 & N+(v2), ... in per-warp global
// scratch (CFSolver.cuh:15-97).  Here every root vertex v0 instead builds the
// DAG induced on its out-list S = N+(v0) once, as adjacency bitmaps over the
// positions of S ("rows": bit j of row i says S[j] is in N+(S[i])), and then
// counts (k-1)-cliques of that local DAG with AND + popcount:
//     k=3: sum_i |row_i|
//     k=4: sum_i sum_{j in row_i} |row_i & row_j|
//     k=5: sum_i sum_{j in row_i} sum_{l in row_i & row_j} |row_i & row_j & row_l|
// Row i equals the reference's W1 for the edge (v0, S[i]); row_i & row_j is its
// W2.  Rows are built by streaming the lists N+(S[i]) through an LDS hash of S
// (a flattened walk: all lanes read consecutive entries, whatever the list
// lengths), or, for lists much longer than S, by binary-searching each S[j]
// in N+(S[i]) and taking the 64-bit ballot as the row word.
//
// The kernels read the DAG relabelled in a topological order (gpm_graph's
// toff/tadj, built once per handle), so S = N+(v0) is sorted topologically
// and every row is upper triangular: row i only holds positions > i.  For
// k=5 the popcount of row_i & row_j & row_l then starts at l's word, which
// skips about half to two thirds of the words.  Counts do not depend on ids.
//
// Three tiers by |S| (tiers.hpp): one wave per root with rows in registers
// (|S| <= 64, one 64-bit word per row, lane i owns row i); one 256-thread
// block per root with rows in LDS (|S| <= 512); one block per root with rows
// in a per-block HBM region (larger |S|).
//
// MODE selects what is accumulated: kCount (the clique count), kBytes (the
// algorithmic bytes of the reference's edge-rooted plan, see gpm.h), kWork
// (per-root plan entries, the cost-model training target).
#include <algorithm>
#include <vector>

#include "graph_dev.hpp"
#include "tiers.hpp"

using namespace gpm;

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kWaveTable = 128;
constexpr uint32_t kBlockCap = 512;
// Threads per block of the block tiers.  Rows are shared by the whole block,
// so wider blocks raise the waves per SIMD that hide LDS latency without
// more LDS: the wide tier fits one block per CU, the LDS tier three.
constexpr int kLdsBlock = 512;
constexpr int kWideBlock = 1024;
constexpr int kHbmBlock = 1024;
// Consecutive list entries one thread streams in the row build.
constexpr uint32_t kSeg = 8;
enum { kCount = 0, kBytes = 1, kWork = 2, kAll = 3 };
// kAll (with K = 5): the 3-, 4- and 5-clique counts of every root from one
// build of its rows; out[0..2] receive them.

// A list N+(S[i]) much longer than S is not streamed: each S[t] is looked up
// in it by binary search instead (d searches of log2(len) steps against len
// streamed entries).  Thresholds measured on R-MAT scale 20: lists longer
// than max(64, 2d) for the wave tier and max(64, 4d) for the block tiers.
__device__ __forceinline__ bool use_probe(uint32_t len, uint32_t d) {
  return len > max(64u, 4u * d);
}
__device__ __forceinline__ bool use_probe_wave(uint32_t len, uint32_t d) {
  return len > max(64u, 2u * d);
}

__device__ __forceinline__ uint32_t ctz64(uint64_t x) { return (uint32_t)__builtin_ctzll(x); }

// Position of the k-th (0-based) set bit of x; x must have more than k bits.
__device__ __forceinline__ uint32_t select_bit(uint64_t x, uint32_t k) {
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t half = 32; half >= 1; half >>= 1) {
    const uint64_t lo_mask = (half == 64) ? ~0ull : ((1ull << half) - 1);
    const uint32_t c = (uint32_t)__popcll(x & lo_mask);
    if (k >= c) {
      k -= c;
      x >>= half;
      pos += half;
    }
  }
  return pos;
}

// ------------------------------------------------------------ wave tier
// Wave w takes the chunks of `chunk` roots w, w + waves, ... of the tier's
// list (ordered by degree or predicted cost, so the interleave evens the
// waves' loads out).  No task queue: a queue fetch is an atomic on one
// address, and those serialise (about 10 ns each), which bounded this tier
// at ~1 M roots per fused pass.
template <int K, int MODE>
__global__ __launch_bounds__(kBlock) void clique_wave_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ roots, uint32_t n_roots, uint32_t chunk,
    unsigned long long* __restrict__ out, unsigned long long* __restrict__ work) {
  __shared__ uint32_t s_keys[kWavesPerBlock][kWaveTable];
  __shared__ uint32_t s_vals[kWavesPerBlock][kWaveTable];
  __shared__ unsigned long long s_rows[kWavesPerBlock][kWave];
  __shared__ uint32_t s_pre[kWavesPerBlock][kWave];
  __shared__ uint32_t s_beg[kWavesPerBlock][kWave];
  __shared__ uint32_t s_len[kWavesPerBlock][kWave];
  __shared__ uint32_t s_sv[kWavesPerBlock][kWave];
  const uint32_t lane = lane_id();
  const uint32_t w = threadIdx.x / kWave;
  uint32_t* keys = s_keys[w];
  uint32_t* vals = s_vals[w];
  unsigned long long* rows = s_rows[w];
  uint32_t* pre = s_pre[w];
  uint32_t* beg = s_beg[w];
  uint32_t* lens = s_len[w];
  uint32_t* svals = s_sv[w];
  constexpr uint32_t hmask = kWaveTable - 1;
  uint64_t acc = 0, acc3 = 0, acc4 = 0;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / kWave;
  for (uint64_t base = wave * chunk; base < n_roots; base += nwaves * chunk) {
    const uint32_t stop = (uint32_t)min<uint64_t>(base + chunk, n_roots);
    for (uint32_t r = (uint32_t)base; r < stop; ++r) {
      const uint32_t u = uniform(roots[r]);
      const uint32_t sb = uniform(off[u]), d = uniform(off[u + 1]) - sb;  // d <= 64 by tiering
      const bool act = lane < d;
      const uint32_t s = act ? adj[sb + lane] : kEmpty;
      keys[lane] = kEmpty;
      keys[lane + kWave] = kEmpty;
      rows[lane] = 0ull;
      __builtin_amdgcn_wave_barrier();
      if (act) lds_hash_insert(keys, vals, hmask, s, lane);
      uint32_t b = 0, len = 0;
      if (act) {
        b = off[s];
        len = off[s + 1] - b;
      }
      lens[lane] = len;
      const bool probe = act && use_probe_wave(len, d);
      const uint32_t incl = wave_incl_scan(probe ? 0u : len);
      pre[lane] = incl;
      beg[lane] = b;
      const uint32_t total = wave_bcast(incl, kWave - 1);
      __builtin_amdgcn_wave_barrier();
      for (uint32_t t = 0; t < total; t += kWave) {
        const uint32_t f = t + lane;
        if (f < total) {
          const uint32_t j = upper_index(pre, kWave, f);
          const uint32_t excl = j ? pre[j - 1] : 0u;
          const uint32_t x = adj[beg[j] + (f - excl)];
          const int idx = lds_hash_find(keys, vals, hmask, x);
          if (idx >= 0) atomicOr(&rows[j], 1ull << idx);
        }
      }
      // Probed lists: the (probed row, element of S) pairs are dealt out
      // over the lanes, so all binary searches of a root run side by side.
      const uint64_t pm = wave_ballot(probe);
      if (pm) {
        svals[lane] = s;
        __builtin_amdgcn_wave_barrier();
        const uint32_t np = (uint32_t)__popcll(pm);
        for (uint32_t f = lane; f < np * d; f += kWave) {
          const uint32_t k = f / d, t = f - k * d;
          const uint32_t i = select_bit(pm, k);
          if (sorted_contains(adj + beg[i], lens[i], svals[t])) atomicOr(&rows[i], 1ull << t);
        }
      }
      __builtin_amdgcn_wave_barrier();
      const uint64_t row = act ? rows[lane] : 0ull;
      uint64_t c = 0;
      if (MODE == kAll) acc3 += __popcll(row);
      if (MODE == kCount || MODE == kAll) {
        if (K == 3) {
          c = __popcll(row);
        } else {
          // The (i, j in row_i) pairs are dealt out evenly over the lanes:
          // pair p belongs to the row whose prefix range holds p, and j is
          // the (p - start)-th set bit of that row.
          const uint32_t pincl = wave_incl_scan((uint32_t)__popcll(row));
          pre[lane] = pincl;
          const uint32_t npairs = wave_bcast(pincl, kWave - 1);
          __builtin_amdgcn_wave_barrier();
          for (uint32_t t = 0; t < npairs; t += kWave) {
            const uint32_t p = t + lane;
            if (p < npairs) {
              const uint32_t i = upper_index(pre, kWave, p);
              const uint32_t excl = i ? pre[i - 1] : 0u;
              const uint64_t ri = rows[i];
              const uint32_t j = select_bit(ri, p - excl);
              const uint64_t mij = ri & rows[j];
              if (MODE == kAll) acc4 += __popcll(mij);
              if (K == 4) {
                c += __popcll(mij);
              } else {
                uint64_t mm = mij;
                while (mm) {
                  const uint32_t l = ctz64(mm);
                  mm &= mm - 1;
                  c += __popcll(mij & rows[l]);
                }
              }
            }
          }
        }
      } else if (act) {
        // Edge (u, S[lane]) of the reference plan.
        const uint64_t w1 = __popcll(row);
        c = (uint64_t)d + len + (K >= 4 ? w1 : 0);
        if (K >= 4) {
          uint64_t bits = row;
          while (bits) {
            const uint32_t j = ctz64(bits);
            bits &= bits - 1;
            if (K == 4) {
              c += w1 + lens[j];
            } else {
              const uint64_t w2m = row & rows[j];
              const uint64_t w2 = __popcll(w2m);
              c += w1 + lens[j] + w2;
              uint64_t mm = w2m;
              while (mm) {
                const uint32_t l = ctz64(mm);
                mm &= mm - 1;
                c += w2 + lens[l];
              }
            }
          }
        }
        c *= 4;  // bytes
      }
      if (MODE == kWork) {
        const uint64_t tot = wave_sum_u64(c);
        if (lane == 0) work[u] = tot / 4;
      } else {
        acc += c;
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (MODE == kAll) {
    acc3 = wave_sum_u64(acc3);
    acc4 = wave_sum_u64(acc4);
    acc = wave_sum_u64(acc);
    if (lane == 0) {
      if (acc3) atomicAdd(out, (unsigned long long)acc3);
      if (acc4) atomicAdd(out + 1, (unsigned long long)acc4);
      if (acc) atomicAdd(out + 2, (unsigned long long)acc);
    }
  } else if (MODE != kWork) {
    acc = wave_sum_u64(acc);
    if (lane == 0 && acc) atomicAdd(out, (unsigned long long)acc);
  }
}

// ------------------------------------------------------ block tiers

__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t* s_wsum) {
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) s_wsum[w] = incl;
  __syncthreads();
  uint32_t add = 0;
  for (uint32_t i = 0; i < w; ++i) add += s_wsum[i];
  __syncthreads();
  return incl + add;
}

// Per-root counting (or accounting) over W-word rows; thread-owned rows.
// K=4/5 counting with the (i, j in row_i) pairs dealt out evenly over the
// block: rows are taken BS at a time, a block scan of their popcounts
// gives each pair an owner row, and j is found by popcount over the words.
//
// MAXW > 0 (rows in LDS, W <= MAXW): rows are read MAXW words at a time
// with no per-word branch, so all loads of a row are in flight together and
// one wait covers them (words past W belong to the next row or to the tail
// of the block's LDS, and are masked out by m[x] = 0 for x >= W).
// MAXW == 0 is the generic path for rows in HBM.
template <int K, int MAXW, int BS, bool ALL = false>
__device__ __forceinline__ uint64_t count_pairs(const unsigned long long* rows, uint32_t d,
                                                uint32_t W, uint32_t RS, uint32_t part,
                                                uint32_t n_parts, uint32_t* s_pre,
                                                uint32_t* s_wsum, uint64_t& c3, uint64_t& c4) {
  uint64_t c = 0;
  // This task counts rows part, part + n_parts, ...: nr of them.
  const uint32_t nr = d > part ? (d - part + n_parts - 1) / n_parts : 0u;
  for (uint32_t c0 = 0; c0 < nr; c0 += BS) {
    const uint32_t t0 = c0 + threadIdx.x;
    uint32_t pc = 0;
    if (t0 < nr) {
      const unsigned long long* r0 = rows + (size_t)(part + t0 * n_parts) * RS;
      if (MAXW > 0) {
#pragma unroll
        for (int w = 0; w < MAXW; ++w) {
          const unsigned long long x = r0[w];
          pc += w < (int)W ? (uint32_t)__popcll(x) : 0u;
        }
      } else {
        for (uint32_t w = 0; w < W; ++w) pc += (uint32_t)__popcll(r0[w]);
      }
    }
    if (ALL) c3 += pc;
    const uint32_t incl = block_incl_scan(pc, s_wsum);
    s_pre[threadIdx.x] = incl;
    __syncthreads();
    const uint32_t total = uniform(s_pre[BS - 1]);
    for (uint32_t p = threadIdx.x; p < total; p += BS) {
      const uint32_t jj = upper_index(s_pre, BS, p);
      uint32_t kth = p - (jj ? s_pre[jj - 1] : 0u);
      if (MAXW > 0) {
        // LDS rows: row index < 1024 and stride < 32, so the 24-bit multiply
        // (full rate; a 32-bit multiply issues at quarter rate) is exact.
        const unsigned long long* ri = rows + __umul24(part + (c0 + jj) * n_parts, RS);
        uint64_t wi[MAXW > 0 ? MAXW : 1];
#pragma unroll
        for (int x = 0; x < MAXW; ++x) wi[x] = x < (int)W ? ri[x] : 0ull;
        // j = position of the kth set bit of row i.
        uint32_t j = 0;
        bool found = false;
#pragma unroll
        for (int x = 0; x < MAXW; ++x) {
          const uint32_t pw = (uint32_t)__popcll(wi[x]);
          if (!found && kth < pw) {
            j = x * 64 + select_bit(wi[x], kth);
            found = true;
          } else if (!found) {
            kth -= pw;
          }
        }
        const unsigned long long* rj = rows + __umul24(j, RS);
        uint64_t m[MAXW > 0 ? MAXW : 1];
#pragma unroll
        for (int x = 0; x < MAXW; ++x) m[x] = wi[x] & rj[x];
        if (K == 4) {
#pragma unroll
          for (int x = 0; x < MAXW; ++x) c += __popcll(m[x]);
        } else {
          if (ALL)
#pragma unroll
            for (int x = 0; x < MAXW; ++x) c4 += __popcll(m[x]);
#pragma unroll
          for (int x = 0; x < MAXW; ++x) {
            uint64_t mm = m[x];
            while (mm) {
              const uint32_t l = x * 64 + ctz64(mm);
              mm &= mm - 1;
              // Row l holds only positions > l (rows are upper triangular),
              // so its words below x are zero and are not read.  Up to 8
              // words, a word of row l is read only by lanes whose m word is
              // non-zero (fewer LDS reads; 33.1 -> 29.5 ms per fused pass on
              // R-MAT s20); the 16-word tier measured faster with every
              // word read under one wait.
              // One 32-bit sum per l (at most 16 x 64 bits), one 64-bit add.
              const unsigned long long* rl = rows + __umul24(l, RS);
              uint32_t cl = 0;
              if (MAXW <= 8) {
#pragma unroll
                for (int y = x; y < MAXW; ++y)
                  if (m[y]) cl += (uint32_t)__popcll(m[y] & rl[y]);
              } else {
                uint64_t a[MAXW > 0 ? MAXW : 1];
#pragma unroll
                for (int y = x; y < MAXW; ++y) a[y] = rl[y];
#pragma unroll
                for (int y = x; y < MAXW; ++y) cl += (uint32_t)__popcll(m[y] & a[y]);
              }
              c += cl;
            }
          }
        }
      } else {
        const unsigned long long* ri = rows + (size_t)(part + (c0 + jj) * n_parts) * RS;
        uint32_t w = 0;
        while (true) {
          const uint32_t pw = (uint32_t)__popcll(ri[w]);
          if (kth < pw) break;
          kth -= pw;
          ++w;
        }
        const uint32_t j = w * 64 + select_bit(ri[w], kth);
        const unsigned long long* rj = rows + (size_t)j * RS;
        if (K == 4) {
          for (uint32_t x = 0; x < W; ++x) c += __popcll(ri[x] & rj[x]);
        } else {
          for (uint32_t x = 0; x < W; ++x) {
            uint64_t mm = ri[x] & rj[x];
            if (ALL) c4 += __popcll(mm);
            while (mm) {
              const uint32_t l = x * 64 + ctz64(mm);
              mm &= mm - 1;
              const unsigned long long* rl = rows + (size_t)l * RS;
              for (uint32_t y = x; y < W; ++y) c += __popcll(ri[y] & rj[y] & rl[y]);
            }
          }
        }
      }
    }
    __syncthreads();
  }
  return c;
}

template <int K, int MODE, int BS>
__device__ __forceinline__ uint64_t count_rows(const unsigned long long* rows, uint32_t d,
                                               uint32_t W, uint32_t RS, uint32_t part,
                                               uint32_t n_parts, const uint32_t* lens) {
  uint64_t c = 0;
  for (uint32_t i = part + threadIdx.x * n_parts; i < d; i += BS * n_parts) {
    const unsigned long long* ri = rows + (size_t)i * RS;
    if (MODE == kCount) {
      for (uint32_t w = 0; w < W; ++w) c += __popcll(ri[w]);  // K == 3 only
    } else {
      uint64_t w1 = 0;
      for (uint32_t w = 0; w < W; ++w) w1 += __popcll(ri[w]);
      uint64_t e = (uint64_t)d + lens[i] + (K >= 4 ? w1 : 0);
      if (K >= 4) {
        for (uint32_t w = 0; w < W; ++w) {
          uint64_t bits = ri[w];
          while (bits) {
            const uint32_t j = w * 64 + ctz64(bits);
            bits &= bits - 1;
            if (K == 4) {
              e += w1 + lens[j];
            } else {
              const unsigned long long* rj = rows + (size_t)j * RS;
              uint64_t w2 = 0;
              for (uint32_t x = 0; x < W; ++x) w2 += __popcll(ri[x] & rj[x]);
              e += w1 + lens[j] + w2;
              for (uint32_t x = 0; x < W; ++x) {
                uint64_t mm = ri[x] & rj[x];
                while (mm) {
                  const uint32_t l = x * 64 + ctz64(mm);
                  mm &= mm - 1;
                  e += w2 + lens[l];
                }
              }
            }
          }
        }
      }
      c += 4 * e;
    }
  }
  return c;
}

// Block per root.  LDS_ROWS: rows, S, lens, begins and the hash table live in
// dynamic LDS (|S| <= cap, W <= MAXW words per row).  Otherwise rows live in this block's
// HBM region `grows` (cap*W words) and membership in S is a binary search.
template <int K, int MODE, bool LDS_ROWS, int MAXW, int BS>
__global__ __launch_bounds__(BS) void clique_block_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ roots, const uint32_t* __restrict__ parts, uint32_t n_roots,
    uint32_t cap, uint32_t table, unsigned int* __restrict__ queue,
    unsigned long long* __restrict__ grows, uint32_t* __restrict__ gmeta,
    unsigned long long* __restrict__ out, unsigned long long* __restrict__ work) {
  extern __shared__ unsigned long long s_dyn[];
  __shared__ uint32_t s_pre[BS];
  __shared__ uint32_t s_wsum[(BS / kWave)];
  __shared__ uint32_t s_root;
  __shared__ uint32_t s_np;
  __shared__ int s_any_probe;
  __shared__ unsigned long long s_red[(BS / kWave)];
  const uint32_t Wcap = (cap + 63) / 64;
  // LDS rows of the 8-word tier are padded to an odd stride (in 8-byte words):
  // the rows read together sit at data-dependent indices, and an even stride
  // maps them onto few banks.
  constexpr bool kPad = LDS_ROWS && MAXW == 8;
  const uint32_t RScap = kPad ? (Wcap | 1u) : Wcap;
  unsigned long long* rows;
  uint32_t *lens, *begs, *keys;
  uint32_t* vals;
  if (LDS_ROWS) {
    rows = s_dyn;
    lens = reinterpret_cast<uint32_t*>(rows + (size_t)cap * RScap);
    begs = lens + cap;
    keys = begs + cap;
    vals = keys + table;
  } else {
    rows = grows + (size_t)blockIdx.x * cap * Wcap;
    lens = gmeta + (size_t)blockIdx.x * 2 * cap;
    begs = lens + cap;
    keys = nullptr;
    vals = nullptr;
  }
  const uint32_t hmask = table - 1;
  uint64_t acc = 0, acc3 = 0, acc4 = 0;
  while (true) {
    if (threadIdx.x == 0) s_root = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t r = uniform(s_root);
    __syncthreads();
    if (r >= n_roots) break;
    const uint32_t u = uniform(roots[r]);
    const uint32_t pp = parts ? uniform(parts[r]) : (1u << 16);
    const uint32_t part = pp & 0xFFFFu, n_parts = pp >> 16;
    const uint32_t sb = uniform(off[u]), d = uniform(off[u + 1]) - sb;
    const uint32_t* S = adj + sb;
    const uint32_t W = (d + 63) / 64;
    const uint32_t RS = kPad ? (W | 1u) : W;
    if (threadIdx.x == 0) s_any_probe = 0;
    for (uint32_t i = threadIdx.x; i < d; i += BS) {
      const uint32_t v = S[i];
      const uint32_t b = off[v];
      const uint32_t len = off[v + 1] - b;
      lens[i] = len;
      begs[i] = b;
    }
    for (uint32_t i = threadIdx.x; i < d * RS; i += BS) rows[i] = 0ull;
    if (LDS_ROWS)
      for (uint32_t i = threadIdx.x; i < table; i += BS) keys[i] = kEmpty;
    __syncthreads();
    // HBM rows: drop stale L1 lines of the previous root before reading.
    if (!LDS_ROWS) __threadfence();
    if (LDS_ROWS)
      for (uint32_t i = threadIdx.x; i < d; i += BS)
        lds_hash_insert(keys, vals, hmask, S[i], i);
    __syncthreads();
    // Flattened walk over the lists that are streamed.
    for (uint32_t c0 = 0; c0 < d; c0 += BS) {
      const uint32_t i = c0 + threadIdx.x;
      uint32_t wl = 0;
      if (i < d) {
        const uint32_t len = lens[i];
        if (use_probe(len, d))
          s_any_probe = 1;
        else
          wl = len;
      }
      const uint32_t incl = block_incl_scan(wl, s_wsum);
      s_pre[threadIdx.x] = incl;
      __syncthreads();
      const uint32_t total = uniform(s_pre[BS - 1]);
      // Each thread takes kSeg consecutive entries: one owner search for the
      // first, then the owner row advances along the prefix sums.
      for (uint32_t f0 = threadIdx.x * kSeg; f0 < total; f0 += BS * kSeg) {
        uint32_t j = upper_index(s_pre, BS, f0);
        uint32_t excl = j ? s_pre[j - 1] : 0u;
        uint32_t end = s_pre[j];
#pragma unroll
        for (uint32_t k = 0; k < kSeg; ++k) {
          const uint32_t f = f0 + k;
          if (f >= total) break;
          while (f >= end) {  // next non-empty row; ends since f < total
            excl = end;
            end = s_pre[++j];
          }
          const uint32_t row_i = c0 + j;
          const uint32_t x = adj[begs[row_i] + (f - excl)];
          int idx;
          if (LDS_ROWS) {
            idx = lds_hash_find(keys, vals, hmask, x);
          } else {
            const uint32_t p = upper_index(S, d, x);
            idx = (p > 0 && S[p - 1] == x) ? (int)(p - 1) : -1;
          }
          // Rows are this block's alone (LDS, or its slice of HBM scratch).
          if (idx >= 0) wg_atomic_or(&rows[(size_t)row_i * RS + (idx >> 6)], 1ull << (idx & 63));
        }
      }
      __syncthreads();
    }
    // Probed lists: the ballot of "S[t] in N+(S[i])" over 64 consecutive t
    // is row word t/64.
    if (uniform((uint32_t)s_any_probe)) {
      // Probed rows of each BS-row chunk are listed in s_pre, then the
      // (probed row, element of S) pairs are dealt out over the block.
      for (uint32_t c0 = 0; c0 < d; c0 += BS) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t flag = (i < d && use_probe(lens[i], d)) ? 1u : 0u;
        const uint32_t incl = block_incl_scan(flag, s_wsum);
        if (flag) s_pre[incl - 1] = i;
        if (threadIdx.x == BS - 1) s_np = incl;
        __syncthreads();
        const uint32_t np = uniform(s_np);
        for (uint32_t f = threadIdx.x; f < np * d; f += BS) {
          const uint32_t k = f / d, t = f - k * d;
          const uint32_t ri = s_pre[k];
          if (sorted_contains(adj + begs[ri], lens[ri], S[t]))
            wg_atomic_or(&rows[(size_t)ri * RS + (t >> 6)], 1ull << (t & 63));
        }
        __syncthreads();
      }
    }
    __syncthreads();
    if (!LDS_ROWS) __threadfence();
    uint64_t c;
    if ((MODE == kCount || MODE == kAll) && K >= 4) {
      // Narrowest register width that holds this root's rows.
      constexpr bool ALL = MODE == kAll;
      // One instantiation per word count: the k = 5 loop reads MAXW words
      // of every row it visits, so a narrower width saves LDS reads.
#define GPM_PAIRS(NW)                                                                     \
  count_pairs<K, (MAXW == 8 ? ((NW) <= 8 ? (NW) : 8) : (MAXW == 16 ? (NW) : MAXW)), BS, ALL>( \
      rows, d, W, RS, part, n_parts, s_pre, s_wsum, acc3, acc4)
      if (MAXW == 8) {
        switch (W) {
          case 0: case 1: case 2: c = GPM_PAIRS(2); break;
          case 3: c = GPM_PAIRS(3); break;
          case 4: c = GPM_PAIRS(4); break;
          case 5: c = GPM_PAIRS(5); break;
          case 6: c = GPM_PAIRS(6); break;
          case 7: c = GPM_PAIRS(7); break;
          default: c = GPM_PAIRS(8); break;
        }
      } else if (MAXW == 16) {
        switch (W) {
          case 0: case 1: case 2: case 3: case 4: case 5: case 6: case 7: case 8: case 9:
            c = GPM_PAIRS(9); break;
          case 10: c = GPM_PAIRS(10); break;
          case 11: c = GPM_PAIRS(11); break;
          case 12: c = GPM_PAIRS(12); break;
          case 13: c = GPM_PAIRS(13); break;
          case 14: c = GPM_PAIRS(14); break;
          case 15: c = GPM_PAIRS(15); break;
          default: c = GPM_PAIRS(16); break;
        }
      } else {
        c = count_pairs<K, MAXW, BS, ALL>(rows, d, W, RS, part, n_parts, s_pre, s_wsum, acc3,
                                          acc4);
      }
#undef GPM_PAIRS
    } else {
      c = count_rows<K, MODE, BS>(rows, d, W, RS, part, n_parts, lens);
    }
    if (MODE == kWork) {
      c = wave_sum_u64(c);
      if (lane_id() == 0) s_red[threadIdx.x / kWave] = c;
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < (BS / kWave); ++i) t += s_red[i];
        work[u] = t / 4;
      }
    } else {
      acc += c;
    }
    __syncthreads();
  }
  if (MODE == kAll) {
    acc3 = wave_sum_u64(acc3);
    acc4 = wave_sum_u64(acc4);
    acc = wave_sum_u64(acc);
    if (lane_id() == 0) {
      if (acc3) atomicAdd(out, (unsigned long long)acc3);
      if (acc4) atomicAdd(out + 1, (unsigned long long)acc4);
      if (acc) atomicAdd(out + 2, (unsigned long long)acc);
    }
  } else if (MODE != kWork) {
    acc = wave_sum_u64(acc);
    if (lane_id() == 0 && acc) atomicAdd(out, (unsigned long long)acc);
  }
}

__global__ void zero_u32_kernel(unsigned int* p, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

// LDS hash slots for a tier of this cap (a power of two, >= 2 cap).
uint32_t table_for(uint32_t cap) {
  uint32_t table = 1;
  while (table < 2 * cap) table <<= 1;
  return table;
}

size_t block_lds_bytes(uint32_t cap, uint32_t table, bool pad) {
  const uint32_t W = (cap + 63) / 64;
  return (size_t)cap * (pad ? (W | 1u) : W) * 8 + (size_t)cap * 8 + (size_t)table * 8;
}

// Largest out-degree whose rows fit one block's LDS on this device (at most
// 1024 = 16 words per row); kBlockCap when the device allows only 64 KiB.
uint32_t lds_wide_cap() {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0;
    hipDeviceProp_t prop;
    size_t limit = 64 * 1024;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      limit = prop.sharedMemPerBlock;
    uint32_t cap = kBlockCap;
    for (uint32_t c = 1024; c > kBlockCap; c -= 64) {
      if (block_lds_bytes(c, table_for(c), false) + 4096 <= limit) {
        cap = c;
        break;
      }
    }
    cached = (int)cap;
  }
  return (uint32_t)cached;
}

// Blocks of an LDS-rows tier of this cap resident per CU: limited by LDS
// (160 KiB per CU) and by 32 waves per CU.
int lds_blocks_per_cu(uint32_t cap, int block_threads, bool pad) {
  const size_t lds = block_lds_bytes(cap, table_for(cap), pad);
  const int by_waves = std::max(1, 32 / (block_threads / kWave));
  return std::min(by_waves, std::max<int>(1, (int)(160 * 1024 / (lds + 2048))));
}

// One LDS-rows tier: rows of up to MAXW words, cap = its largest out-degree.
template <int K, int MODE, int MAXW, int BS>
int launch_lds_tier(gpm_graph* g, hipStream_t stream, const uint32_t* roots,
                    const uint32_t* parts, uint32_t n_roots, uint32_t cap,
                    unsigned int* queue, unsigned long long* out, unsigned long long* work,
                    const char* name) {
  const int cus = device_cus(g);
  const uint32_t table = table_for(cap);
  const size_t lds = block_lds_bytes(cap, table, MAXW == 8);
  auto kern = clique_block_kernel<K, MODE, true, MAXW, BS>;
  if (lds > 48 * 1024)
    GPM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const unsigned blocks =
      (unsigned)std::min<uint64_t>(n_roots, (uint64_t)cus * lds_blocks_per_cu(cap, BS, MAXW == 8));
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(BS), lds, stream, g->toff, g->tadj, roots,
                     parts, n_roots, cap, table, queue, nullptr, nullptr, out, work);
  GPM_HIP(hipGetLastError());
  if ((MODE == kCount || MODE == kAll) && stream == g->stream) timing_mark(g, name);
  return GPM_OK;
}

template <int K, int MODE>
int launch_wave_tier(gpm_graph* g, hipStream_t stream, const RootTiers* t, unsigned int* q,
                     unsigned long long* out, unsigned long long* work) {
  const int cus = device_cus(g);
  const uint32_t chunk = t->wave_chunk;
  const uint64_t waves_wanted = (t->n_wave + chunk - 1) / chunk;
  const uint64_t waves = std::min<uint64_t>(waves_wanted, (uint64_t)cus * 32);
  const unsigned blocks =
      (unsigned)std::max<uint64_t>(1, (waves + kWavesPerBlock - 1) / kWavesPerBlock);
  hipLaunchKernelGGL((clique_wave_kernel<K, MODE>), dim3(blocks), dim3(kBlock), 0, stream,
                     g->toff, g->tadj, t->d_wave, t->n_wave, chunk, out, work);
  GPM_HIP(hipGetLastError());
  if ((MODE == kCount || MODE == kAll) && stream == g->stream) timing_mark(g, "clique_wave");
  return GPM_OK;
}

// Launches every non-empty tier on the handle's stream, one after another.
// (Running the rows-in-HBM and wave tiers on a second stream beside the LDS
// tiers was measured slower: 254 ms against 230 ms for the learned 5-clique
// count on R-MAT s22, from contention with the LDS tiers.)
template <int K, int MODE>
int launch_tiers(gpm_graph* g, const RootTiers* t, unsigned long long* out,
                 unsigned long long* work) {
  const int cus = device_cus(g);
  unsigned int* q = t->queues;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, g->stream, q, 4);
  GPM_HIP(hipGetLastError());
  if (t->n_big) {
    const uint32_t cap = t->max_big_deg;
    const uint32_t W = (cap + 63) / 64;
    const unsigned blocks = (unsigned)std::min<uint64_t>(t->n_big, (uint64_t)cus);
    const size_t row_bytes = (size_t)blocks * cap * W * 8;
    const size_t meta_bytes = (size_t)blocks * 2 * cap * 4;
    int st = ensure_scratch(g, row_bytes + meta_bytes + 256);
    if (st != GPM_OK) return st;
    unsigned long long* grows = static_cast<unsigned long long*>(g->scratch);
    uint32_t* gmeta = reinterpret_cast<uint32_t*>(static_cast<char*>(g->scratch) + row_bytes);
    hipLaunchKernelGGL((clique_block_kernel<K, MODE, false, 0, kHbmBlock>), dim3(blocks),
                       dim3(kHbmBlock), 0, g->stream, g->toff, g->tadj, t->d_big, t->d_big_parts,
                       t->n_big, cap, 0u, q + 2, grows, gmeta, out, work);
    GPM_HIP(hipGetLastError());
    if (MODE == kCount || MODE == kAll) timing_mark(g, "clique_hbm_block");
  }
  if (t->n_block2) {
    int st = launch_lds_tier<K, MODE, 16, kWideBlock>(g, g->stream, t->d_block2, t->d_block2_parts,
                                                      t->n_block2, t->max_block2_deg, q + 3, out,
                                                      work, "clique_lds_block_wide");
    if (st != GPM_OK) return st;
  }
  if (t->n_block) {
    int st = launch_lds_tier<K, MODE, 8, kLdsBlock>(g, g->stream, t->d_block, t->d_block_parts,
                                                    t->n_block, t->max_block_deg, q + 1, out, work,
                                                    "clique_lds_block");
    if (st != GPM_OK) return st;
  }
  if (t->n_wave) {
    int st = launch_wave_tier<K, MODE>(g, g->stream, t, q, out, work);
    if (st != GPM_OK) return st;
  }
  return GPM_OK;
}

template <int MODE>
int launch_k(gpm_graph* g, int k, const RootTiers* t, unsigned long long* out,
             unsigned long long* work) {
  if (MODE == kAll) return launch_tiers<5, MODE>(g, t, out, work);
  switch (k) {
    case 3: return launch_tiers<3, MODE>(g, t, out, work);
    case 4: return launch_tiers<4, MODE>(g, t, out, work);
    default: return launch_tiers<5, MODE>(g, t, out, work);
  }
}

TierRequest request_from(const gpm_sched* s, uint32_t min_deg) {
  TierRequest r;
  r.min_deg = min_deg;
  r.wave_cap = 64;
  r.block_cap = kBlockCap;
  r.block2_cap = lds_wide_cap();
  if (s) {
    r.root_begin = s->root_begin;
    r.root_end = s->root_end;
    r.policy = s->policy;
    r.cost = s->root_cost;
    r.chunk = s->chunk;
    r.root_list = s->root_list;
    r.n_root_list = s->root_list ? s->n_root_list : 0;
    if (s->wave_cap && s->wave_cap < 64) r.wave_cap = s->wave_cap;
  }
  return r;
}

}  // namespace

namespace gpm {
// Heavy-root threshold for the learned policy: wave roots whose predicted
// cost exceeds it are served by a whole block.
float learned_heavy_cost(const float* cost, uint32_t n);
}

namespace {

// Tier request of one k-clique call.  LEARNED uses the caller's predicted
// costs; DEGREE_SPLIT uses the same ordering, promotion and splitting with
// the plain degree estimate cost(v) = d+(v)^2 (the root's pair count), so
// the two differ only in where the cost comes from; DEGREE keeps the
// reference-style degree order without splitting.
int prepare_request(gpm_graph* g, const gpm_sched* sched, uint32_t min_deg, TierRequest* req,
                    std::vector<float>* deg_cost) {
  GPM_CHECK_ARG(!sched || sched->policy == GPM_SCHED_DEGREE ||
                    sched->policy == GPM_SCHED_LEARNED || sched->policy == GPM_SCHED_DEGREE_SPLIT,
                "unknown scheduling policy");
  GPM_CHECK_ARG(!sched || sched->policy != GPM_SCHED_LEARNED || sched->root_cost,
                "LEARNED policy needs root_cost");
  int st = ensure_topo(g);
  if (st != GPM_OK) return st;
  *req = request_from(sched, min_deg);
  req->rank = g->h_rank.data();
  if (req->policy == GPM_SCHED_DEGREE_SPLIT) {
    deg_cost->resize(g->n);
    for (uint32_t v = 0; v < g->n; ++v) {
      const float d = (float)(g->h_off[v + 1] - g->h_off[v]);
      (*deg_cost)[v] = d * d;
    }
    req->cost = deg_cost->data();
    req->policy = GPM_SCHED_LEARNED;
  }
  if (req->policy == GPM_SCHED_LEARNED) {
    req->heavy_cost = learned_heavy_cost(req->cost, g->n);
    const uint32_t cus = (uint32_t)device_cus(g);
    req->resident_block = cus * lds_blocks_per_cu(req->block_cap, kLdsBlock, true);
    req->resident_block2 = cus * lds_blocks_per_cu(req->block2_cap, kWideBlock, false);
    req->resident_big = cus;
  }
  return GPM_OK;
}

}  // namespace

extern "C" {

int gpm_clique(gpm_graph* g, int k, const gpm_sched* sched, uint64_t* out) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g && out, "null pointer");
    GPM_CHECK_ARG(g->flags & GPM_GRAPH_ORIENTED, "k-clique counting needs an ORIENTED graph");
    GPM_CHECK_ARG(k >= 3 && k <= 5, "k must be 3, 4 or 5 (got %d)", k);
    TierRequest req;
    std::vector<float> deg_cost;
    int st = prepare_request(g, sched, (uint32_t)(k - 1), &req, &deg_cost);
    if (st != GPM_OK) return st;
    const RootTiers* t = nullptr;
    if ((st = get_tiers(g, req, &t)) != GPM_OK) return st;
    if ((st = timing_begin(g)) != GPM_OK) return st;
    if ((st = zero_counters(g, 1)) != GPM_OK) return st;
    if ((st = launch_k<kCount>(g, k, t, g->counters, nullptr)) != GPM_OK) return st;
    if ((st = timing_end(g)) != GPM_OK) return st;
    unsigned long long h = 0;
    GPM_HIP(hipMemcpyAsync(&h, g->counters, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    GPM_HIP(hipStreamSynchronize(g->stream));
    *out = h;
    return GPM_OK;
  });
}

int gpm_clique_all(gpm_graph* g, const gpm_sched* sched, uint64_t out[3]) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g && out, "null pointer");
    GPM_CHECK_ARG(g->flags & GPM_GRAPH_ORIENTED, "k-clique counting needs an ORIENTED graph");
    TierRequest req;
    std::vector<float> deg_cost;
    int st = prepare_request(g, sched, 2u, &req, &deg_cost);  // triangles need d+ >= 2
    if (st != GPM_OK) return st;
    const RootTiers* t = nullptr;
    if ((st = get_tiers(g, req, &t)) != GPM_OK) return st;
    if ((st = timing_begin(g)) != GPM_OK) return st;
    if ((st = zero_counters(g, 3)) != GPM_OK) return st;
    if ((st = launch_k<kAll>(g, 5, t, g->counters, nullptr)) != GPM_OK) return st;
    if ((st = timing_end(g)) != GPM_OK) return st;
    unsigned long long h[3] = {0, 0, 0};
    GPM_HIP(hipMemcpyAsync(h, g->counters, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    GPM_HIP(hipStreamSynchronize(g->stream));
    for (int i = 0; i < 3; ++i) out[i] = h[i];
    return GPM_OK;
  });
}

int gpm_clique_tier_caps(uint32_t* wave_cap, uint32_t* lds_cap, uint32_t* wide_cap) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(wave_cap && lds_cap && wide_cap, "null pointer");
    *wave_cap = 64;
    *lds_cap = kBlockCap;
    *wide_cap = lds_wide_cap();
    return GPM_OK;
  });
}

int gpm_clique_plan_bytes(gpm_graph* g, int k, uint64_t* bytes) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g && bytes, "null pointer");
    GPM_CHECK_ARG(g->flags & GPM_GRAPH_ORIENTED, "needs an ORIENTED graph");
    GPM_CHECK_ARG(k >= 3 && k <= 5, "k must be 3, 4 or 5 (got %d)", k);
    if (!g->plan_bytes[k]) {
      int st = ensure_topo(g);
      if (st != GPM_OK) return st;
      TierRequest req = request_from(nullptr, 1);
      req.rank = g->h_rank.data();
      const RootTiers* t = nullptr;
      st = get_tiers(g, req, &t);
      if (st != GPM_OK) return st;
      if ((st = zero_counters(g, 1)) != GPM_OK) return st;
      if ((st = launch_k<kBytes>(g, k, t, g->counters, nullptr)) != GPM_OK) return st;
      unsigned long long h = 0;
      GPM_HIP(hipMemcpyAsync(&h, g->counters, sizeof(h), hipMemcpyDeviceToHost, g->stream));
      GPM_HIP(hipStreamSynchronize(g->stream));
      g->plan_bytes[k] = h + 4ull * (g->n + 1) + 8ull * g->m;
    }
    *bytes = g->plan_bytes[k];
    return GPM_OK;
  });
}

int gpm_clique_root_work(gpm_graph* g, int k, uint64_t* work) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(g && work, "null pointer");
    GPM_CHECK_ARG(g->flags & GPM_GRAPH_ORIENTED, "needs an ORIENTED graph");
    GPM_CHECK_ARG(k >= 3 && k <= 5, "k must be 3, 4 or 5 (got %d)", k);
    int st = ensure_topo(g);
    if (st != GPM_OK) return st;
    TierRequest req = request_from(nullptr, 1);
    req.rank = g->h_rank.data();
    const RootTiers* t = nullptr;
    st = get_tiers(g, req, &t);
    if (st != GPM_OK) return st;
    unsigned long long* d_work = nullptr;
    GPM_HIP(hipMalloc(&d_work, (size_t)std::max(g->n, 1u) * 8));
    if (hipMemsetAsync(d_work, 0, (size_t)std::max(g->n, 1u) * 8, g->stream) != hipSuccess) {
      (void)hipFree(d_work);
      return fail(GPM_EHIP, "clearing the work log failed");
    }
    st = launch_k<kWork>(g, k, t, g->counters, d_work);
    // The kernels log by topological id; report by vertex id.
    std::vector<uint64_t> by_rank(g->n);
    if (st == GPM_OK) {
      if (hipMemcpyAsync(by_rank.data(), d_work, (size_t)g->n * 8, hipMemcpyDeviceToHost,
                         g->stream) != hipSuccess ||
          hipStreamSynchronize(g->stream) != hipSuccess)
        st = fail(GPM_EHIP, "work log copy failed");
    }
    hipFree(d_work);
    if (st == GPM_OK)
      for (uint32_t v = 0; v < g->n; ++v) work[v] = by_rank[g->h_rank[v]];
    return st;
  });
}

}  // extern "C"
