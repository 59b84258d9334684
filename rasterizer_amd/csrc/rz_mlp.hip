// Batch-scoring MLP kernels for gfx950: 16 -> 64 -> 64 -> 1 with GELU.
//
// Every matrix product runs on v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation).  A wave takes 32 samples at a time (two 16-row tiles).
// Operands are staged in LDS with the reduction dimension contiguous, so a
// lane reads its 8 consecutive k values with one 16-byte load:
//   A tile (16 x 32): lane l reads row l%16, k = 8*(l/16) .. +7
//   B tile (32 x 16): lane l reads column l%16, k = 8*(l/16) .. +7 (stored [n][k])
//   D tile (16 x 16): lane l holds rows 4*(l/16) .. +3 of column l%16
// Work-groups walk the samples in steps of 128 together (4 waves x 32) so
// that block barriers stay uniform.
#include <hip/hip_runtime.h>

#include "rz_kernels.h"

namespace rzd {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kPadK32 = 40;  // row stride (bf16) of 32-wide operands: 80 B rows avoid bank-aligned rows
constexpr int kPadK64 = 72;  // row stride of 64-wide operands (144 B)

__device__ __forceinline__ f32x4 mfma_tile(const __bf16* A, int lda, const __bf16* B, int ldb, f32x4 c,
                                           uint32_t lane) {
  const int r = int(lane & 15u), kq = int(lane >> 4) * 8;
  const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + r * lda + kq);
  const bf16x8 b = *reinterpret_cast<const bf16x8*>(B + r * ldb + kq);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// Weights shared by a work-group, as BF16 operands.
struct MlpWeights {
  __bf16 w1t[64 * kPadK32];  // [out][in], in padded with zeros to 32
  __bf16 w2t[64 * kPadK64];  // [out][in]  (layer 2 forward: B[n = out][k = in])
  __bf16 w2n[64 * kPadK64];  // [in][out]  (dH1 = dZ2 W2^T: B[n = in][k = out])
};

// The forward pass's weights and one wave's forward working set (the
// scoring kernel: 14 KB + 4 x 7 KB of LDS, so that three work-groups share a
// CU).
struct MlpWeightsFwd {
  __bf16 w1t[64 * kPadK32];
  __bf16 w2t[64 * kPadK64];
};
struct MlpWaveFwd {
  __bf16 xn[32 * kPadK32];
  __bf16 h1[32 * kPadK64];
};

// One wave's 32-sample working set (training).
struct MlpWave {
  __bf16 xn[32 * kPadK32];    // [sample][feature], features 16-31 zero
  __bf16 xnt[16 * kPadK32];   // [feature][sample]
  __bf16 h1[32 * kPadK64];    // [sample][hidden]
  __bf16 h1t[64 * kPadK32];   // [hidden][sample]
  __bf16 dz2[32 * kPadK64];   // [sample][hidden]
  __bf16 dz2t[64 * kPadK32];  // [hidden][sample]
  __bf16 dz1t[64 * kPadK32];  // [hidden][sample]
};

template <class WeightsT>
__device__ void stage_weights(WeightsT& W, const float* __restrict__ w) {
  for (int i = int(threadIdx.x); i < 64 * 32; i += 256) {
    const int o = i >> 5, k = i & 31;
    W.w1t[o * kPadK32 + k] = __bf16(k < 16 ? w[kMlpW1 + k * 64 + o] : 0.0f);
  }
  for (int i = int(threadIdx.x); i < 64 * 64; i += 256) {
    const int a = i >> 6, b = i & 63;
    W.w2t[a * kPadK64 + b] = __bf16(w[kMlpW2 + b * 64 + a]);  // a = out, b = in
    if constexpr (sizeof(WeightsT) == sizeof(MlpWeights))
      W.w2n[a * kPadK64 + b] = __bf16(w[kMlpW2 + a * 64 + b]);  // a = in, b = out
  }
}

// Normalised inputs of the wave's 32 samples (zeros past n) into xn and,
// for the backward pass (TRANSPOSED), xnt.  Sample s of the wave is row
// base + s of x, or row rows[base + s] when a row list is given (n = its
// length).
template <bool TRANSPOSED, class WaveT>
__device__ void stage_inputs(WaveT& S, const MlpParams& p, uint32_t base, uint32_t n, uint32_t lane) {
  const uint32_t s = lane >> 1, f0 = (lane & 1u) * 8u;
  const bool in = base + s < n;
  const uint32_t g = !in ? 0u : p.rows ? p.rows[base + s] : base + s;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    const uint32_t f = f0 + j;
    const float x = in ? (p.x[size_t(g) * 16 + f] - p.norm[f]) * p.norm[16 + f] : 0.0f;
    const __bf16 b = __bf16(x);
    S.xn[s * kPadK32 + f] = b;
    S.xn[s * kPadK32 + 16 + f] = __bf16(0.0f);
    if constexpr (TRANSPOSED) S.xnt[f * kPadK32 + s] = b;
  }
}

struct Acts {
  f32x4 z1[2][4], z2[2][4];
  f32x4 h2[2][4];
};

// Forward of the wave's 32 samples: z1, z2 (pre-activations), h2 in
// registers, h1 (BF16) in LDS, and for the backward pass (TRANSPOSED) also
// transposed.
template <bool TRANSPOSED, class WeightsT, class WaveT>
__device__ void forward_wave(const WeightsT& W, WaveT& S, const float* __restrict__ w, uint32_t lane, Acts& a) {
  const uint32_t col = lane & 15u, rq = (lane >> 4) * 4u;
#pragma unroll
  for (int rt = 0; rt < 2; ++rt)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
      c = mfma_tile(S.xn + 16 * rt * kPadK32, kPadK32, W.w1t + 16 * t * kPadK32, kPadK32, c, lane);
      const float bias = w[kMlpB1 + 16 * t + col];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = c[i] + bias;
        a.z1[rt][t][i] = z;
        const __bf16 h = __bf16(gelu(z));
        const uint32_t row = 16u * rt + rq + i, hid = 16u * t + col;
        S.h1[row * kPadK64 + hid] = h;
        if constexpr (TRANSPOSED) S.h1t[hid * kPadK32 + row] = h;
      }
    }
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the wave's LDS writes are done
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int rt = 0; rt < 2; ++rt)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kk = 0; kk < 2; ++kk)
        c = mfma_tile(S.h1 + 16 * rt * kPadK64 + 32 * kk, kPadK64, W.w2t + 16 * t * kPadK64 + 32 * kk, kPadK64, c,
                      lane);
      const float bias = w[kMlpB2 + 16 * t + col];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = c[i] + bias;
        a.z2[rt][t][i] = z;
        a.h2[rt][t][i] = gelu(z);
      }
    }
}

// Sum over the 16 lanes of the lane's quarter (lanes 16q .. 16q+15).
__device__ __forceinline__ float sum16(float v) {
  v += __shfl_xor(v, 8);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 1);
  return v;
}

}  // namespace

__global__ void __launch_bounds__(256) rz_mlp_forward_kernel(MlpParams p) {
  __shared__ __attribute__((aligned(16))) MlpWeightsFwd W;
  __shared__ __attribute__((aligned(16))) MlpWaveFwd Sw[4];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  MlpWaveFwd& S = Sw[wave];
  stage_weights(W, p.w);
  __syncthreads();
  const uint32_t col = lane & 15u, rq = (lane >> 4) * 4u;
  const float b3 = p.w[kMlpB3];
  const uint32_t n = p.rows ? min(*p.nRows, p.n) : p.n;
  for (uint32_t base0 = blockIdx.x * 128u; base0 < n; base0 += gridDim.x * 128u) {
    const uint32_t base = base0 + 32u * wave;
    stage_inputs<false>(S, p, base, n, lane);
    __syncthreads();
    Acts a;
    forward_wave<false>(W, S, p.w, lane, a);
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float part = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) part += a.h2[rt][t][i] * p.w[kMlpW3 + 16 * t + col];
        const float sc = sum16(part) + b3;
        const uint32_t s = base + 16u * rt + rq + i;
        if (col == 0 && s < n) p.scores[p.rows ? p.rows[s] : s] = sc;
      }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) rz_mlp_backward_kernel(MlpParams p) {
  __shared__ __attribute__((aligned(16))) MlpWeights W;
  __shared__ __attribute__((aligned(16))) MlpWave Sw[4];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  MlpWave& S = Sw[wave];
  stage_weights(W, p.w);
  __syncthreads();
  const uint32_t col = lane & 15u, rq = (lane >> 4) * 4u;
  f32x4 dW2[4][4], dW1[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    dW1[m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int n = 0; n < 4; ++n) dW2[m][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  }
  float dW3[4] = {0.0f, 0.0f, 0.0f, 0.0f}, db2[4] = {0.0f, 0.0f, 0.0f, 0.0f}, db1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  float db3 = 0.0f;
  for (uint32_t base0 = blockIdx.x * 128u; base0 < p.n; base0 += gridDim.x * 128u) {
    const uint32_t base = base0 + 32u * wave;
    stage_inputs<true>(S, p, base, p.n, lane);
    __syncthreads();
    Acts a;
    forward_wave<true>(W, S, p.w, lane, a);
    // Layer 3 and layer 2's activation: dz2 = ds * W3 * gelu'(z2).
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t row = 16u * rt + rq + i, g = base + row;
        const float ds = g < p.n ? p.dscore[g] : 0.0f;
        if (col == 0) db3 += ds;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const uint32_t hid = 16u * t + col;
          dW3[t] += ds * a.h2[rt][t][i];
          const float dz = ds * p.w[kMlpW3 + hid] * gelu_grad(a.z2[rt][t][i]);
          db2[t] += dz;
          const __bf16 b = __bf16(dz);
          S.dz2[row * kPadK64 + hid] = b;
          S.dz2t[hid * kPadK32 + row] = b;
        }
      }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    // dW2[in][out] += H1^T dZ2 over the 32 samples.
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n)
        dW2[m][n] = mfma_tile(S.h1t + 16 * m * kPadK32, kPadK32, S.dz2t + 16 * n * kPadK32, kPadK32, dW2[m][n], lane);
    // dH1 = dZ2 W2^T, then dz1 = dH1 * gelu'(z1).
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
          c = mfma_tile(S.dz2 + 16 * rt * kPadK64 + 32 * kk, kPadK64, W.w2n + 16 * t * kPadK64 + 32 * kk, kPadK64, c,
                        lane);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float dz = c[i] * gelu_grad(a.z1[rt][t][i]);
          db1[t] += dz;
          const uint32_t row = 16u * rt + rq + i, hid = 16u * t + col;
          S.dz1t[hid * kPadK32 + row] = __bf16(dz);
        }
      }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    // dW1[feature][hidden] += Xn^T dZ1 over the 32 samples.
#pragma unroll
    for (int n = 0; n < 4; ++n) dW1[n] = mfma_tile(S.xnt, kPadK32, S.dz1t + 16 * n * kPadK32, kPadK32, dW1[n], lane);
    __syncthreads();
  }
  // Flush the wave's sums.
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t in = 16u * m + rq + i, out = 16u * n + col;
        atomicAdd(&p.grad[kMlpW2 + in * 64 + out], dW2[m][n][i]);
      }
#pragma unroll
  for (int n = 0; n < 4; ++n)
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(&p.grad[kMlpW1 + (rq + i) * 64 + 16 * n + col], dW1[n][i]);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    // Lanes of the four quarters hold partial sums for the same column.
    float v3 = dW3[t], v2 = db2[t], v1 = db1[t];
    v3 += __shfl_xor(v3, 16);
    v3 += __shfl_xor(v3, 32);
    v2 += __shfl_xor(v2, 16);
    v2 += __shfl_xor(v2, 32);
    v1 += __shfl_xor(v1, 16);
    v1 += __shfl_xor(v1, 32);
    if (lane < 16) {
      atomicAdd(&p.grad[kMlpW3 + 16 * t + col], v3);
      atomicAdd(&p.grad[kMlpB2 + 16 * t + col], v2);
      atomicAdd(&p.grad[kMlpB1 + 16 * t + col], v1);
    }
  }
  db3 = sum16(db3);
  db3 += __shfl_xor(db3, 16);
  db3 += __shfl_xor(db3, 32);
  if (lane == 0) atomicAdd(&p.grad[kMlpB3], db3);
}

__global__ void rz_mlp_pair_loss_kernel(const float* __restrict__ scores, const uint32_t* __restrict__ pairs,
                                        uint32_t nPairs, float* __restrict__ dscore, float* __restrict__ loss) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (k < nPairs) {
    const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    const float d = scores[i] - scores[j];
    // softplus(-d) = log(1 + exp(-d)); its derivative in d is -sigmoid(-d).
    l = fmaxf(-d, 0.0f) + log1pf(expf(-fabsf(d)));
    const float sg = 1.0f / (1.0f + expf(d));
    const float inv = 1.0f / float(nPairs);
    atomicAdd(&dscore[i], -sg * inv);
    atomicAdd(&dscore[j], sg * inv);
    l *= inv;
  }
  // One atomic per wave for the loss.
  for (int o = 32; o >= 1; o >>= 1) l += __shfl_xor(l, o);
  if ((threadIdx.x & 63u) == 0u) atomicAdd(loss, l);
}

__global__ void rz_mlp_adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                   float* __restrict__ v, uint32_t n, float lr, float beta1, float beta2, float eps,
                                   float corr1, float corr2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / corr1) / (sqrtf(vi / corr2) + eps);
}

__global__ void rz_fill_f32_kernel(float* p, uint32_t n, float v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static inline uint32_t groups_of(uint32_t n, uint32_t per) { return (n + per - 1) / per; }

uint32_t mlp_forward_max_groups() {
  // The scoring kernel's work-groups that fit the device at once (3 per CU
  // on gfx950): a grid of exactly that many runs its grid-stride loop in
  // one round, where a larger grid would leave a partial second round.
  static const uint32_t cap = [] {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, rz_mlp_forward_kernel, 256, 0) != hipSuccess ||
        cus <= 0 || per <= 0)
      return 1024u;
    return uint32_t(cus) * uint32_t(per);
  }();
  return cap;
}

hipError_t launch_mlp_forward(const MlpParams& p, hipStream_t stream) {
  if (p.n == 0) return hipSuccess;
  const uint32_t cap = mlp_forward_max_groups();
  const uint32_t g = groups_of(p.n, 128) < cap ? groups_of(p.n, 128) : cap;
  hipLaunchKernelGGL(rz_mlp_forward_kernel, dim3(g), dim3(256), 0, stream, p);
  return hipGetLastError();
}

hipError_t launch_mlp_backward(const MlpParams& p, hipStream_t stream) {
  if (p.n == 0) return hipSuccess;
  const uint32_t g = groups_of(p.n, 128) < 512 ? groups_of(p.n, 128) : 512;
  hipLaunchKernelGGL(rz_mlp_backward_kernel, dim3(g), dim3(256), 0, stream, p);
  return hipGetLastError();
}

hipError_t launch_mlp_pair_loss(const float* scores, const uint32_t* pairs, uint32_t nPairs, float* dscore,
                                float* loss, hipStream_t stream) {
  if (nPairs == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_mlp_pair_loss_kernel, dim3(groups_of(nPairs, 256)), dim3(256), 0, stream, scores, pairs,
                     nPairs, dscore, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float* w, const float* grad, float* m, float* v, uint32_t n, float lr, float beta1,
                           float beta2, float eps, float corr1, float corr2, hipStream_t stream) {
  hipLaunchKernelGGL(rz_mlp_adam_kernel, dim3(groups_of(n, 256)), dim3(256), 0, stream, w, grad, m, v, n, lr, beta1,
                     beta2, eps, corr1, corr2);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, uint32_t n, float v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_fill_f32_kernel, dim3(groups_of(n, 256)), dim3(256), 0, stream, p, n, v);
  return hipGetLastError();
}

}  // namespace rzd
