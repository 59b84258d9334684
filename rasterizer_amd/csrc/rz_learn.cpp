// The batch-scoring MLP (C ABI in rasterizer_amd.h, rz_mlp_*): parameters on
// the device, forward scores, pairwise-ranking gradients and Adam steps, all
// through the gfx950 kernels of rz_mlp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "rasterizer_amd.h"
#include "rz_internal.h"
#include "rz_kernels.h"

using namespace rzd;

namespace {

// SplitMix64 (Steele, Lea, Flood 2014) for deterministic initial weights.
uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

float uniform(uint64_t& s) { return float(splitmix(s) >> 40) * (1.0f / 16777216.0f); }

void mlp_free(rz_mlp* m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  hipFree(m->d_w);
  hipFree(m->d_norm);
  hipFree(m->d_grad);
  hipFree(m->d_m);
  hipFree(m->d_v);
  hipFree(m->d_x);
  hipFree(m->d_scores);
  hipFree(m->d_ds);
  hipFree(m->d_pairs);
  hipFree(m->d_loss);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int ensure_samples(rz_mlp* m, uint32_t n) {
  if (n <= m->capX) return RZ_OK;
  hipFree(m->d_x);
  hipFree(m->d_scores);
  hipFree(m->d_ds);
  m->d_x = m->d_scores = m->d_ds = nullptr;
  m->capX = 0;
  if (hipMalloc(reinterpret_cast<void**>(&m->d_x), size_t(n) * kMlpIn * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_scores), size_t(n) * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_ds), size_t(n) * 4) != hipSuccess)
    return RZ_ERR_OUT_OF_MEMORY;
  m->capX = n;
  return RZ_OK;
}

int ensure_pairs(rz_mlp* m, uint32_t n) {
  if (n <= m->capPairs) return RZ_OK;
  hipFree(m->d_pairs);
  m->d_pairs = nullptr;
  m->capPairs = 0;
  if (hipMalloc(reinterpret_cast<void**>(&m->d_pairs), size_t(n) * 8) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
  m->capPairs = n;
  return RZ_OK;
}

MlpParams params_of(const rz_mlp* m, uint32_t n) {
  MlpParams p{};
  p.w = m->d_w;
  p.norm = m->d_norm;
  p.x = m->d_x;
  p.n = n;
  p.scores = m->d_scores;
  p.dscore = m->d_ds;
  p.grad = m->d_grad;
  return p;
}

// Loss and gradients of the pairs on the uploaded samples into d_grad.
int run_gradients(rz_mlp* m, uint32_t n, const float* x, uint32_t nPairs, const uint32_t* pairs) {
  for (uint32_t k = 0; k < 2 * nPairs; ++k)
    if (pairs[k] >= n) return RZ_ERR_INVALID_ARGUMENT;
  int s = ensure_samples(m, std::max(n, 1u));
  if (s != RZ_OK) return s;
  s = ensure_pairs(m, std::max(nPairs, 1u));
  if (s != RZ_OK) return s;
  RZ_HIP(hipMemcpyAsync(m->d_x, x, size_t(n) * kMlpIn * 4, hipMemcpyHostToDevice, m->stream));
  if (nPairs) RZ_HIP(hipMemcpyAsync(m->d_pairs, pairs, size_t(nPairs) * 8, hipMemcpyHostToDevice, m->stream));
  const MlpParams p = params_of(m, n);
  RZ_HIP(launch_fill_f32(m->d_grad, kMlpParams, 0.0f, m->stream));
  RZ_HIP(launch_fill_f32(m->d_ds, n, 0.0f, m->stream));
  RZ_HIP(launch_fill_f32(m->d_loss, 1, 0.0f, m->stream));
  RZ_HIP(launch_mlp_forward(p, m->stream));
  RZ_HIP(launch_mlp_pair_loss(m->d_scores, m->d_pairs, nPairs, m->d_ds, m->d_loss, m->stream));
  RZ_HIP(launch_mlp_backward(p, m->stream));
  return RZ_OK;
}

}  // namespace

extern "C" int rz_mlp_create(uint32_t seed, rz_mlp** out) {
  if (!out) return RZ_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RZ_ERR_DEVICE;
  rz_mlp* m = new (std::nothrow) rz_mlp();
  if (!m) return RZ_ERR_OUT_OF_MEMORY;
  auto fail = [&](int code) {
    mlp_free(m);
    return code;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return fail(RZ_ERR_DEVICE);
  const size_t pb = kMlpParams * sizeof(float);
  if (hipMalloc(reinterpret_cast<void**>(&m->d_w), pb) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_grad), pb) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_m), pb) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_v), pb) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_norm), 32 * sizeof(float)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_loss), sizeof(float)) != hipSuccess)
    return fail(RZ_ERR_OUT_OF_MEMORY);
  // Glorot-uniform weights, zero biases, identity normalisation.
  std::vector<float> w(kMlpParams, 0.0f);
  uint64_t st = seed;
  auto fill = [&](int off, int fanIn, int fanOut) {
    const float a = std::sqrt(6.0f / float(fanIn + fanOut));
    for (int i = 0; i < fanIn * fanOut; ++i) w[off + i] = (2.0f * uniform(st) - 1.0f) * a;
  };
  fill(kMlpW1, kMlpIn, kMlpHidden);
  fill(kMlpW2, kMlpHidden, kMlpHidden);
  fill(kMlpW3, kMlpHidden, 1);
  std::vector<float> norm(32, 0.0f);
  for (int f = 0; f < 16; ++f) norm[16 + f] = 1.0f;
  if (hipMemcpy(m->d_w, w.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_norm, norm.data(), 32 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->d_m, 0, pb) != hipSuccess || hipMemset(m->d_v, 0, pb) != hipSuccess)
    return fail(RZ_ERR_DEVICE);
  *out = m;
  return RZ_OK;
}

extern "C" int rz_mlp_destroy(rz_mlp* m) {
  if (!m) return RZ_ERR_INVALID_ARGUMENT;
  mlp_free(m);
  return RZ_OK;
}

extern "C" int rz_mlp_set_uses_labels(rz_mlp* m, int usesLabels) {
  if (!m) return RZ_ERR_INVALID_ARGUMENT;
  m->usesLabels = usesLabels != 0;
  return RZ_OK;
}

extern "C" int rz_mlp_set_params(rz_mlp* m, const float* params, const float* norm) {
  if (!m) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(m->stream));
  if (params) RZ_HIP(hipMemcpy(m->d_w, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  if (norm) RZ_HIP(hipMemcpy(m->d_norm, norm, 32 * sizeof(float), hipMemcpyHostToDevice));
  if (params) {
    RZ_HIP(hipMemset(m->d_m, 0, kMlpParams * sizeof(float)));
    RZ_HIP(hipMemset(m->d_v, 0, kMlpParams * sizeof(float)));
    m->step = 0;
  }
  return RZ_OK;
}

extern "C" int rz_mlp_get_params(rz_mlp* m, float* params, float* norm) {
  if (!m) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(m->stream));
  if (params) RZ_HIP(hipMemcpy(params, m->d_w, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  if (norm) RZ_HIP(hipMemcpy(norm, m->d_norm, 32 * sizeof(float), hipMemcpyDeviceToHost));
  return RZ_OK;
}

extern "C" int rz_mlp_forward(rz_mlp* m, uint32_t n, const float* x, float* scores) {
  if (!m || (n && (!x || !scores))) return RZ_ERR_INVALID_ARGUMENT;
  if (n == 0) return RZ_OK;
  int s = ensure_samples(m, n);
  if (s != RZ_OK) return s;
  RZ_HIP(hipMemcpyAsync(m->d_x, x, size_t(n) * kMlpIn * 4, hipMemcpyHostToDevice, m->stream));
  RZ_HIP(launch_mlp_forward(params_of(m, n), m->stream));
  RZ_HIP(hipMemcpyAsync(scores, m->d_scores, size_t(n) * 4, hipMemcpyDeviceToHost, m->stream));
  RZ_HIP(hipStreamSynchronize(m->stream));
  return RZ_OK;
}

extern "C" int rz_mlp_gradients(rz_mlp* m, uint32_t n, const float* x, uint32_t nPairs, const uint32_t* pairs,
                                float* grad, float* loss) {
  if (!m || !x || (nPairs && !pairs) || n == 0) return RZ_ERR_INVALID_ARGUMENT;
  int s = run_gradients(m, n, x, nPairs, pairs);
  if (s != RZ_OK) return s;
  if (grad) RZ_HIP(hipMemcpyAsync(grad, m->d_grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (loss) RZ_HIP(hipMemcpyAsync(loss, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  RZ_HIP(hipStreamSynchronize(m->stream));
  return RZ_OK;
}

extern "C" int rz_mlp_train_step(rz_mlp* m, uint32_t n, const float* x, uint32_t nPairs, const uint32_t* pairs,
                                 float lr, float* loss) {
  if (!m || !x || (nPairs && !pairs) || n == 0) return RZ_ERR_INVALID_ARGUMENT;
  int s = run_gradients(m, n, x, nPairs, pairs);
  if (s != RZ_OK) return s;
  m->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, float(m->step)), c2 = 1.0f - std::pow(b2, float(m->step));
  RZ_HIP(launch_mlp_adam(m->d_w, m->d_grad, m->d_m, m->d_v, kMlpParams, lr, b1, b2, 1e-8f, c1, c2, m->stream));
  if (loss) RZ_HIP(hipMemcpyAsync(loss, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  RZ_HIP(hipStreamSynchronize(m->stream));
  return RZ_OK;
}
