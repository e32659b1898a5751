// Batched sorted-set operations (gpm_setops): one wave per list pair, with
// the membership test picked by `variant` -- the primitives of
// src/graph/search.cuh, set_intersection.cuh and set_difference.cuh, each
// rewritten for wave64:
//
//   GPM_SETV_BINARY     each lane binary-searches its entry of A in B
//                       (search.cuh:20-35).
//   GPM_SETV_TWO_PHASE  64 pivots B[k*|B|/64] are staged in LDS; a lane
//                       bisects the pivots in LDS, then searches one 1/64
//                       slice of B in HBM (binary_search_2phase,
//                       search.cuh:54-87; 64 pivots instead of 32, and no
//                       read of B[0] when B is empty).
//   GPM_SETV_MERGE      merge path: the merged order of A and B is cut into
//                       64 equal diagonal segments, one per lane; each lane
//                       finds its cut by a binary search on the diagonal and
//                       merges its segment sequentially (no per-entry search).
//   GPM_SETV_HASH       B goes into an open-addressed LDS table of the wave
//                       (|B| <= kHashCap, else binary search), lanes probe it.
//   GPM_SETV_BITMAP     B goes into an LDS bitmap over [B[0], B[|B|-1]]
//                       (span <= kBitmapBits, else binary search).
//
// Operations (A is iterated, B probed; A \ B never swaps the operands,
// set_difference.cuh:98-234; A & B iterates A as given):
//   ISECT_COUNT / ISECT  |{x in A & B : x < ub}| / the list
//   DIFF_COUNT / DIFF    |{x in A \ B : x < ub}| / the list
//   FUSE                 one pass giving A \ B, A & B and B \ A, with the
//                        matched entries of B marked as they are found
//                        (fuse_set, set_difference.cuh:236-325); returns
//                        the three sizes.
// Materialised lists are sorted and written at the pair's offset into the
// caller's buffers.
#include <algorithm>
#include <vector>

#include "device_common.hpp"
#include "setops.hpp"

using namespace gpm;

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kHashCap = 512;              // B entries an LDS table takes
constexpr uint32_t kHashSlots = 2 * kHashCap;   // load factor <= 1/2
constexpr uint32_t kBitmapBits = 32 * 1024;     // span of ids an LDS bitmap covers
constexpr uint32_t kNoMatch = 0xFFFFFFFFu;

struct WaveLds {
  uint32_t pivots[kWave];
  uint32_t keys[kHashSlots];
  uint32_t vals[kHashSlots];
};  // the bitmap reuses keys+vals (2 * kHashSlots words = kBitmapBits bits)
static_assert(2 * kHashSlots * 32 >= kBitmapBits, "bitmap does not fit the LDS area");

// Position of key in sorted b[lo, hi), or kNoMatch.
__device__ __forceinline__ uint32_t search_range(const uint32_t* __restrict__ b, uint32_t lo,
                                                 uint32_t hi, uint32_t key) {
  uint32_t n = hi - lo;
  if (n == 0) return kNoMatch;
  uint32_t at = lo;
  while (n > 1) {
    const uint32_t half = n >> 1;
    at = b[at + half] <= key ? at + half : at;
    n -= half;
  }
  return b[at] == key ? at : kNoMatch;
}

// Probe state of one wave for one pair: the variant actually used (table
// variants fall back to binary search when B does not fit) and its LDS.
struct Probe {
  int variant;
  const uint32_t* b;
  uint32_t nb;
  uint32_t base;  // bitmap: id of bit 0
  WaveLds* lds;
};

__device__ Probe make_probe(int variant, const uint32_t* b, uint32_t nb, WaveLds* lds) {
  const uint32_t lane = lane_id();
  Probe p{variant, b, nb, 0, lds};
  if (variant == GPM_SETV_HASH && nb > kHashCap) p.variant = GPM_SETV_BINARY;
  if (variant == GPM_SETV_BITMAP && (nb == 0 || b[nb - 1] - b[0] >= kBitmapBits))
    p.variant = nb == 0 ? GPM_SETV_HASH : GPM_SETV_BINARY;  // empty B: an empty table
  if (p.variant == GPM_SETV_TWO_PHASE) {
    if (nb) lds->pivots[lane] = b[(uint32_t)(((uint64_t)lane * nb) / kWave)];
  } else if (p.variant == GPM_SETV_HASH) {
    for (uint32_t i = lane; i < kHashSlots; i += kWave) lds->keys[i] = kEmpty;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane; i < nb; i += kWave) lds_hash_insert(lds->keys, lds->vals, kHashSlots - 1, b[i], i);
  } else if (p.variant == GPM_SETV_BITMAP) {
    uint32_t* bits = lds->keys;
    p.base = b[0];
    for (uint32_t i = lane; i < 2 * kHashSlots; i += kWave) bits[i] = 0u;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane; i < nb; i += kWave) {
      const uint32_t o = b[i] - p.base;
      atomicOr(&bits[o >> 5], 1u << (o & 31));
    }
  }
  __builtin_amdgcn_wave_barrier();
  return p;
}

// Position of x in B, or kNoMatch (bitmap: 0 for a hit, positions unknown).
__device__ __forceinline__ uint32_t probe(const Probe& p, uint32_t x) {
  switch (p.variant) {
    case GPM_SETV_TWO_PHASE: {
      if (p.nb == 0) return kNoMatch;
      const uint32_t* piv = p.lds->pivots;
      uint32_t k = 0;  // last pivot <= x (pivot 0 is B[0])
#pragma unroll
      for (uint32_t step = kWave / 2; step > 0; step >>= 1)
        k = piv[k + step] <= x ? k + step : k;
      if (piv[0] > x) return kNoMatch;
      const uint32_t lo = (uint32_t)(((uint64_t)k * p.nb) / kWave);
      const uint32_t hi = (uint32_t)(((uint64_t)(k + 1) * p.nb) / kWave);
      return search_range(p.b, lo, max(hi, lo + 1u), x);
    }
    case GPM_SETV_HASH: {
      const int v = lds_hash_find(p.lds->keys, p.lds->vals, kHashSlots - 1, x);
      return v < 0 ? kNoMatch : (uint32_t)v;
    }
    case GPM_SETV_BITMAP: {
      if (x < p.base || x - p.base >= kBitmapBits) return kNoMatch;
      const uint32_t o = x - p.base;
      return (p.lds->keys[o >> 5] >> (o & 31)) & 1u ? 0u : kNoMatch;
    }
    default:
      return search_range(p.b, 0, p.nb, x);
  }
}

// Merge-path cut of diagonal d between sorted A (na) and B (nb): the number
// of A entries among the first d entries of the merged order (A before B on
// ties, so a match A[i] == B[j] is consumed A first).
__device__ __forceinline__ uint32_t merge_cut(const uint32_t* a, uint32_t na, const uint32_t* b,
                                              uint32_t nb, uint32_t d) {
  uint32_t lo = d > nb ? d - nb : 0, hi = min(d, na);
  while (lo < hi) {
    const uint32_t i = (lo + hi) >> 1;  // i A entries, d - i B entries taken?
    if (a[i] <= b[d - i - 1])
      lo = i + 1;
    else
      hi = i;
  }
  return lo;
}

// Merge-path pass: every lane walks its diagonal segment and calls
// f(i, match) for each A entry i it consumes, match = position in B or
// kNoMatch.  An A entry equal to the next unconsumed B entry is a match
// (that B entry may lie past the segment).
template <typename F>
__device__ __forceinline__ void merge_walk(const uint32_t* a, uint32_t na, const uint32_t* b,
                                           uint32_t nb, F&& f) {
  const uint32_t lane = lane_id();
  const uint64_t total = (uint64_t)na + nb;
  const uint32_t d0 = (uint32_t)(total * lane / kWave), d1 = (uint32_t)(total * (lane + 1) / kWave);
  uint32_t i = merge_cut(a, na, b, nb, d0), j = d0 - i;
  for (uint32_t d = d0; d < d1; ++d) {
    if (i < na && (j >= nb || a[i] <= b[j])) {
      f(i, (j < nb && a[i] == b[j]) ? j : kNoMatch);
      ++i;
    } else {
      ++j;
    }
  }
}

__global__ __launch_bounds__(kBlock) void setops_kernel(
    int op, int variant, uint32_t n_pairs, const uint64_t* __restrict__ a_off,
    const uint32_t* __restrict__ a_all, const uint64_t* __restrict__ b_off,
    const uint32_t* __restrict__ b_all, const uint32_t* __restrict__ ubs,
    unsigned long long* __restrict__ counts, uint32_t* __restrict__ out_a,
    uint32_t* __restrict__ out_ab, uint32_t* __restrict__ out_b) {
  __shared__ WaveLds s_lds[kWavesPerBlock];
  WaveLds* lds = &s_lds[threadIdx.x / kWave];
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const bool keep_in = op == GPM_SETOP_ISECT_COUNT || op == GPM_SETOP_ISECT;
  const bool write = op == GPM_SETOP_ISECT || op == GPM_SETOP_DIFF;
  for (uint32_t pr = wave; pr < n_pairs; pr += nwaves) {
    const uint64_t ab = a_off[pr], bb = b_off[pr];
    const uint32_t* A = a_all + ab;
    const uint32_t* B = b_all + bb;
    uint32_t na = uniform((uint32_t)(a_off[pr + 1] - ab));
    const uint32_t nb = uniform((uint32_t)(b_off[pr + 1] - bb));
    if (op == GPM_SETOP_FUSE) {
      // Marks of matched B entries live in out_b until B \ A is compacted.
      for (uint32_t i = lane; i < nb; i += kWave) out_b[bb + i] = 0u;
      __builtin_amdgcn_wave_barrier();
      uint32_t c_dif = 0, c_int = 0;
      if (variant == GPM_SETV_MERGE) {
        // Per-lane sizes first (merge order is A order), then the writes.
        uint32_t kd = 0, ki = 0;
        merge_walk(A, na, B, nb, [&](uint32_t, uint32_t m) { (m == kNoMatch ? kd : ki) += 1; });
        const uint32_t ed = wave_incl_scan(kd) - kd, ei = wave_incl_scan(ki) - ki;
        c_dif = uniform(wave_bcast(ed + kd, kWave - 1));
        c_int = uniform(wave_bcast(ei + ki, kWave - 1));
        uint32_t wd = ed, wi = ei;
        merge_walk(A, na, B, nb, [&](uint32_t i, uint32_t m) {
          if (m == kNoMatch) {
            out_a[ab + wd++] = A[i];
          } else {
            out_ab[ab + wi++] = A[i];
            out_b[bb + m] = 1u;
          }
        });
      } else {
        const Probe p = make_probe(variant == GPM_SETV_BITMAP ? GPM_SETV_BINARY : variant, B, nb, lds);
        for (uint32_t base = 0; base < na; base += kWave) {
          const uint32_t i = base + lane;
          const bool valid = i < na;
          const uint32_t x = valid ? A[i] : 0u;
          const uint32_t m = valid ? probe(p, x) : kNoMatch;
          const uint64_t dm = wave_ballot(valid && m == kNoMatch), im = wave_ballot(valid && m != kNoMatch);
          if (valid && m == kNoMatch) out_a[ab + c_dif + lanes_below(dm)] = x;
          if (valid && m != kNoMatch) {
            out_ab[ab + c_int + lanes_below(im)] = x;
            out_b[bb + m] = 1u;
          }
          c_dif += (uint32_t)__popcll(dm);
          c_int += (uint32_t)__popcll(im);
        }
        __builtin_amdgcn_wave_barrier();
      }
      __threadfence_block();
      // B \ A: the unmarked entries of B, compacted in place (writes never
      // pass the reads of the same window).
      uint32_t c_ba = 0;
      for (uint32_t base = 0; base < nb; base += kWave) {
        const uint32_t i = base + lane;
        const bool valid = i < nb;
        const bool keep = valid && out_b[bb + i] == 0u;
        const uint32_t x = keep ? B[i] : 0u;
        const uint64_t km = wave_ballot(keep);
        __builtin_amdgcn_wave_barrier();
        if (keep) out_b[bb + c_ba + lanes_below(km)] = x;
        c_ba += (uint32_t)__popcll(km);
        __threadfence_block();
      }
      if (lane == 0) {
        counts[3 * (uint64_t)pr] = c_dif;
        counts[3 * (uint64_t)pr + 1] = c_int;
        counts[3 * (uint64_t)pr + 2] = c_ba;
      }
      continue;
    }
    // Upper bound: only x < ub of A take part (the walk stops there).
    const uint32_t ub = ubs ? ubs[pr] : kNoBound;
    if (ub != kNoBound) na = uniform(upper_index(A, na, ub - 1u) * (ub > 0u));
    uint32_t cnt = 0;
    if (variant == GPM_SETV_MERGE) {
      uint32_t k = 0;
      merge_walk(A, na, B, nb, [&](uint32_t, uint32_t m) { k += (m != kNoMatch) == keep_in; });
      const uint32_t excl = wave_incl_scan(k) - k;
      cnt = uniform(wave_bcast(excl + k, kWave - 1));
      if (write) {
        uint32_t w = excl;
        merge_walk(A, na, B, nb, [&](uint32_t i, uint32_t m) {
          if ((m != kNoMatch) == keep_in) out_a[ab + w++] = A[i];
        });
      }
    } else {
      const Probe p = make_probe(variant, B, nb, lds);
      for (uint32_t base = 0; base < na; base += kWave) {
        const uint32_t i = base + lane;
        const bool valid = i < na;
        const uint32_t x = valid ? A[i] : 0u;
        const bool keep = valid && ((probe(p, x) != kNoMatch) == keep_in);
        const uint64_t km = wave_ballot(keep);
        if (write && keep) out_a[ab + cnt + lanes_below(km)] = x;
        cnt += (uint32_t)__popcll(km);
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) counts[pr] = cnt;
  }
}

template <typename T>
int upload_vec(const T* src, size_t count, T** dst, std::vector<void*>& owned) {
  *dst = nullptr;
  if (!count) return GPM_OK;
  if (hipMalloc(dst, count * sizeof(T)) != hipSuccess) {
    *dst = nullptr;
    return fail(GPM_ENOMEM, "device buffer of %zu bytes", count * sizeof(T));
  }
  owned.push_back(*dst);
  GPM_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return GPM_OK;
}

int run_setops(int op, int variant, uint32_t n_pairs, const uint64_t* a_off, const uint32_t* a,
               const uint64_t* b_off, const uint32_t* b, const uint32_t* ub, uint64_t* counts,
               uint32_t* out_a, uint32_t* out_ab, uint32_t* out_b, std::vector<void*>& owned) {
  GPM_CHECK_ARG(op >= GPM_SETOP_ISECT_COUNT && op <= GPM_SETOP_FUSE, "unknown set op %d", op);
  GPM_CHECK_ARG(variant >= GPM_SETV_BINARY && variant <= GPM_SETV_BITMAP, "unknown variant %d",
                variant);
  GPM_CHECK_ARG(a_off && b_off && counts, "null pointer");
  if (n_pairs == 0) return GPM_OK;
  const uint64_t na = a_off[n_pairs], nb = b_off[n_pairs];
  GPM_CHECK_ARG((a || na == 0) && (b || nb == 0), "null list data");
  for (uint32_t i = 0; i < n_pairs; ++i) {
    GPM_CHECK_ARG(a_off[i] <= a_off[i + 1] && b_off[i] <= b_off[i + 1], "offsets decrease at %u", i);
    GPM_CHECK_ARG(a_off[i + 1] - a_off[i] < (1ull << 32) && b_off[i + 1] - b_off[i] < (1ull << 32),
                  "list %u longer than 2^32 - 1", i);
  }
  for (uint64_t i = 0; i < na; ++i)
    GPM_CHECK_ARG(i == 0 || a[i - 1] < a[i] || std::upper_bound(a_off, a_off + n_pairs + 1, i)[-1] == i,
                  "list A is not strictly increasing at entry %llu", (unsigned long long)i);
  for (uint64_t i = 0; i < nb; ++i)
    GPM_CHECK_ARG(i == 0 || b[i - 1] < b[i] || std::upper_bound(b_off, b_off + n_pairs + 1, i)[-1] == i,
                  "list B is not strictly increasing at entry %llu", (unsigned long long)i);
  const bool write = op == GPM_SETOP_ISECT || op == GPM_SETOP_DIFF || op == GPM_SETOP_FUSE;
  GPM_CHECK_ARG(!write || out_a || na == 0, "null output");
  GPM_CHECK_ARG(op != GPM_SETOP_FUSE || ((out_ab || na == 0) && (out_b || nb == 0)),
                "FUSE needs out_ab and out_b");
  int st;
  uint64_t *d_aoff, *d_boff;
  uint32_t *d_a, *d_b, *d_ub = nullptr, *d_oa = nullptr, *d_oab = nullptr, *d_ob = nullptr;
  if ((st = upload_vec(a_off, (size_t)n_pairs + 1, &d_aoff, owned)) != GPM_OK) return st;
  if ((st = upload_vec(b_off, (size_t)n_pairs + 1, &d_boff, owned)) != GPM_OK) return st;
  if ((st = upload_vec(a, na, &d_a, owned)) != GPM_OK) return st;
  if ((st = upload_vec(b, nb, &d_b, owned)) != GPM_OK) return st;
  if (ub && op != GPM_SETOP_FUSE && (st = upload_vec(ub, n_pairs, &d_ub, owned)) != GPM_OK) return st;
  const size_t ncnt = (size_t)n_pairs * (op == GPM_SETOP_FUSE ? 3 : 1);
  unsigned long long* d_cnt = nullptr;
  if (hipMalloc(&d_cnt, ncnt * 8) != hipSuccess) return fail(GPM_ENOMEM, "counts");
  owned.push_back(d_cnt);
  auto alloc_out = [&](uint64_t count, uint32_t** p) -> int {
    if (!count) return GPM_OK;
    if (hipMalloc(p, count * 4) != hipSuccess) {
      *p = nullptr;
      return fail(GPM_ENOMEM, "output buffer");
    }
    owned.push_back(*p);
    return GPM_OK;
  };
  if (write && (st = alloc_out(na, &d_oa)) != GPM_OK) return st;
  if (op == GPM_SETOP_FUSE && ((st = alloc_out(na, &d_oab)) != GPM_OK ||
                               (st = alloc_out(nb, &d_ob)) != GPM_OK))
    return st;
  int dev = 0, cus = 256;
  GPM_HIP(hipGetDevice(&dev));
  GPM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint64_t blocks = std::max<uint64_t>(
      1, std::min<uint64_t>(((uint64_t)n_pairs + kWavesPerBlock - 1) / kWavesPerBlock,
                            (uint64_t)cus * 8));
  hipLaunchKernelGGL(setops_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, 0, op, variant,
                     n_pairs, d_aoff, d_a ? d_a : reinterpret_cast<uint32_t*>(d_aoff), d_boff,
                     d_b ? d_b : reinterpret_cast<uint32_t*>(d_boff), d_ub, d_cnt, d_oa, d_oab,
                     d_ob);
  GPM_HIP(hipGetLastError());
  GPM_HIP(hipDeviceSynchronize());
  std::vector<unsigned long long> hc(ncnt);
  GPM_HIP(hipMemcpy(hc.data(), d_cnt, ncnt * 8, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < ncnt; ++i) counts[i] = hc[i];
  if (d_oa) GPM_HIP(hipMemcpy(out_a, d_oa, na * 4, hipMemcpyDeviceToHost));
  if (d_oab) GPM_HIP(hipMemcpy(out_ab, d_oab, na * 4, hipMemcpyDeviceToHost));
  if (d_ob) GPM_HIP(hipMemcpy(out_b, d_ob, nb * 4, hipMemcpyDeviceToHost));
  return GPM_OK;
}

}  // namespace

extern "C" int gpm_setops(int op, int variant, uint32_t n_pairs, const uint64_t* a_off,
                          const uint32_t* a, const uint64_t* b_off, const uint32_t* b,
                          const uint32_t* ub, uint64_t* counts, uint32_t* out_a,
                          uint32_t* out_ab, uint32_t* out_b) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    std::vector<void*> owned;
    const int st = run_setops(op, variant, n_pairs, a_off, a, b_off, b, ub, counts, out_a, out_ab,
                              out_b, owned);
    for (void* p : owned) (void)hipFree(p);
    return st;
  });
}
