// Work estimator of the learned scheduler: per-vertex features and a small
// MLP (4 -> 64 -> 64 -> 1, ReLU) evaluated on gfx950 matrix cores in BF16
// with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// One wave evaluates 16 vertices per step.  Layer 1 is one 16x32 by 32x16
// MFMA per 16 hidden units (the 4 features padded to K = 32 with zeros);
// its 16x64 output is rounded to BF16 and passed through LDS to re-enter as
// the A operand of layer 2 (two K = 32 chunks per 16 hidden units); layer 3
// is a dot product reduced across the 16 lanes that hold a row.  Weights
// stay in registers for the whole launch.
//
// MFMA operand layout (16x16x32, wave64): lane l supplies A[l%16][8(l/16)+i]
// and B[8(l/16)+i][l%16] for i = 0..7, and receives D[4(l/16)+i][l%16] for
// i = 0..3.
#include <cmath>
#include <vector>

#include "graph_dev.hpp"

using namespace gpm;

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kHidden = 64;
constexpr int kFeat = 4;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs here are finite).
__device__ __forceinline__ __bf16 to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return __builtin_bit_cast(__bf16, (uint16_t)(u >> 16));
}

struct Mlp {
  const float* mean;  // 4
  const float* std;   // 4
  const float* w1;    // 64 x 4 (out x in)
  const float* b1;    // 64
  const float* w2;    // 64 x 64
  const float* b2;    // 64
  const float* w3;    // 64
  const float* b3;    // 1
};

__global__ __launch_bounds__(kBlock) void mlp_kernel(const float* __restrict__ feat, uint32_t n,
                                                     Mlp m, float* __restrict__ out) {
  __shared__ __bf16 s_h1[kWavesPerBlock][16][kHidden + 8];
  const uint32_t lane = lane_id();
  const uint32_t w = threadIdx.x / kWave;
  const uint32_t r = lane % 16, grp = lane / 16;
  // Registers: B operands of layers 1 and 2, biases and w3 for this lane's
  // output columns n = 16t + r.
  bf16x8 b1v[4], b2v[4][2];
  float bias1[4], bias2[4], w3v[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int col = 16 * t + r;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = 8 * grp + i;
      b1v[t][i] = to_bf16(k < kFeat ? m.w1[col * kFeat + k] : 0.f);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 8; ++i) b2v[t][c][i] = to_bf16(m.w2[col * kHidden + 32 * c + 8 * grp + i]);
    bias1[t] = m.b1[col];
    bias2[t] = m.b2[col];
    w3v[t] = m.w3[col];
  }
  float mean[kFeat], inv_std[kFeat];
#pragma unroll
  for (int k = 0; k < kFeat; ++k) {
    mean[k] = m.mean[k];
    inv_std[k] = 1.f / m.std[k];
  }
  const float bias3 = m.b3[0];
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  for (uint32_t base = (blockIdx.x * kWavesPerBlock + w) * 16; base < n; base += nwaves * 16) {
    // Layer 1.
    bf16x8 a;
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = to_bf16(0.f);
    const uint32_t v = base + r;
    if (grp == 0 && v < n) {
#pragma unroll
      for (int k = 0; k < kFeat; ++k) a[k] = to_bf16((log1pf(feat[4 * v + k]) - mean[k]) * inv_std[k]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      floatx4 c = {0.f, 0.f, 0.f, 0.f};
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1v[t], c, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) s_h1[w][4 * grp + i][16 * t + r] = to_bf16(fmaxf(c[i] + bias1[t], 0.f));
    }
    __builtin_amdgcn_wave_barrier();
    // Layer 2: A rows from LDS.
    bf16x8 a2[2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 8; ++i) a2[c][i] = s_h1[w][r][32 * c + 8 * grp + i];
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      floatx4 c = {0.f, 0.f, 0.f, 0.f};
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[0], b2v[t][0], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[1], b2v[t][1], c, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) part[i] += fmaxf(c[i] + bias2[t], 0.f) * w3v[t];
    }
    // Layer 3: sum over the 16 lanes holding the same rows.
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) part[i] += __shfl_xor(part[i], o, kWave);
    if (r == 0)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t row = base + 4 * grp + i;
        if (row < n) out[row] = part[i] + bias3;
      }
    __builtin_amdgcn_wave_barrier();
  }
}

// feat[4v + {0,1,2,3}] = {deg, out-degree in dag, 2-hop bound, label}.
__global__ __launch_bounds__(kBlock) void features_kernel(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ doff, const uint32_t* __restrict__ dadj,
    const uint32_t* __restrict__ labels, uint32_t n, float* __restrict__ feat) {
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  for (uint32_t v = wave; v < n; v += nwaves) {
    const uint32_t deg = off[v + 1] - off[v];
    uint32_t dplus = 0;
    uint64_t hop = 0;
    if (doff) {
      // Oriented 2-hop: the entries the k-clique row build of v streams.
      const uint32_t b = doff[v], e = doff[v + 1];
      dplus = e - b;
      for (uint32_t i = b + lane; i < e; i += kWave) {
        const uint32_t u = dadj[i];
        hop += doff[u + 1] - doff[u];
      }
    } else {
      const uint32_t b = off[v], e = off[v + 1];
      for (uint32_t i = b + lane; i < e; i += kWave) {
        const uint32_t u = adj[i];
        hop += off[u + 1] - off[u];
      }
    }
    hop = wave_sum_u64(hop);
    if (lane == 0) {
      feat[4 * (size_t)v + 0] = (float)deg;
      feat[4 * (size_t)v + 1] = (float)dplus;
      feat[4 * (size_t)v + 2] = (float)hop;
      feat[4 * (size_t)v + 3] = labels ? (float)labels[v] : 0.f;
    }
  }
}

Mlp unpack(const float* d_w) {
  Mlp m;
  m.mean = d_w;
  m.std = d_w + 4;
  m.w1 = d_w + 8;
  m.b1 = m.w1 + 64 * 4;
  m.w2 = m.b1 + 64;
  m.b2 = m.w2 + 64 * 64;
  m.w3 = m.b2 + 64;
  m.b3 = m.w3 + 64;
  return m;
}

}  // namespace

namespace gpm {

int mlp_forward(const float* d_feat, uint32_t n, const float* d_packed, float* d_out, int cus,
                hipStream_t stream) {
  if (!n) return GPM_OK;
  const uint64_t groups = (n + 15) / 16;
  const unsigned blocks = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((groups + kWavesPerBlock - 1) / kWavesPerBlock, (uint64_t)cus * 8));
  hipLaunchKernelGGL(mlp_kernel, dim3(blocks), dim3(kBlock), 0, stream, d_feat, n,
                     unpack(d_packed), d_out);
  GPM_HIP(hipGetLastError());
  return GPM_OK;
}

}  // namespace gpm

extern "C" {

int gpm_vertex_features(gpm_graph* sym, gpm_graph* dag, float* feat) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(sym && feat, "null pointer");
    GPM_CHECK_ARG(!dag || dag->n == sym->n, "sym and dag differ in n");
    const uint32_t n = sym->n;
    float* d_feat = nullptr;
    GPM_HIP(hipMalloc(&d_feat, (size_t)std::max(n, 1u) * 16));
    const int cus = device_cus(sym);
    const unsigned blocks = (unsigned)std::max<uint64_t>(
        1, std::min<uint64_t>(((uint64_t)n * kWave + kBlock - 1) / kBlock, (uint64_t)cus * 16));
    hipLaunchKernelGGL(features_kernel, dim3(blocks), dim3(kBlock), 0, sym->stream, sym->off,
                       sym->adj, dag ? dag->off : nullptr, dag ? dag->adj : nullptr, sym->labels, n,
                       d_feat);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(feat, d_feat, (size_t)n * 16, hipMemcpyDeviceToHost, sym->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(sym->stream);
    (void)hipFree(d_feat);
    if (e != hipSuccess) return fail(GPM_EHIP, "vertex features: %s", hipGetErrorString(e));
    return GPM_OK;
  });
}

int gpm_mlp_predict(const float* feat, uint32_t n, const float* norm_mean, const float* norm_std,
                    const float* w1, const float* b1, const float* w2, const float* b2,
                    const float* w3, const float* b3, float* out) {
  return ::gpm::guarded([&]() -> int {
    clear_error();
    GPM_CHECK_ARG(feat && norm_mean && norm_std && w1 && b1 && w2 && b2 && w3 && b3 && out,
                  "null pointer");
    for (int k = 0; k < kFeat; ++k)
      GPM_CHECK_ARG(norm_std[k] > 0.f && std::isfinite(norm_std[k]) && std::isfinite(norm_mean[k]),
                    "normalisation std must be finite and > 0");
    for (uint32_t i = 0; i < 4 * n; ++i)
      GPM_CHECK_ARG(feat[i] > -1.f && std::isfinite(feat[i]), "features must be finite and > -1");
    const size_t wcount = kMlpWeights;
    std::vector<float> packed;
    packed.reserve(wcount);
    auto put = [&](const float* p, size_t c) { packed.insert(packed.end(), p, p + c); };
    put(norm_mean, 4);
    put(norm_std, 4);
    put(w1, 64 * 4);
    put(b1, 64);
    put(w2, 64 * 64);
    put(b2, 64);
    put(w3, 64);
    put(b3, 1);
    int dev = 0;
    GPM_HIP(hipGetDevice(&dev));
    float *d_w = nullptr, *d_feat = nullptr, *d_out = nullptr;
    hipStream_t stream = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&d_w, wcount * 4);
    if (e == hipSuccess) e = hipMalloc(&d_feat, (size_t)std::max(n, 1u) * 16);
    if (e == hipSuccess) e = hipMalloc(&d_out, (size_t)std::max(n, 1u) * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(d_w, packed.data(), wcount * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && n)
      e = hipMemcpyAsync(d_feat, feat, (size_t)n * 16, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && n) {
      const Mlp m = unpack(d_w);
      hipDeviceProp_t prop;
      int cus = 256;
      if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
      const uint64_t groups = (n + 15) / 16;
      const unsigned blocks = (unsigned)std::max<uint64_t>(
          1, std::min<uint64_t>((groups + kWavesPerBlock - 1) / kWavesPerBlock, (uint64_t)cus * 8));
      hipLaunchKernelGGL(mlp_kernel, dim3(blocks), dim3(kBlock), 0, stream, d_feat, n, m, d_out);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n * 4, hipMemcpyDeviceToHost, stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(d_w);
    (void)hipFree(d_feat);
    (void)hipFree(d_out);
    if (stream) (void)hipStreamDestroy(stream);
    if (e != hipSuccess) return fail(GPM_EHIP, "mlp predict: %s", hipGetErrorString(e));
    return GPM_OK;
  });
}

}  // extern "C"
