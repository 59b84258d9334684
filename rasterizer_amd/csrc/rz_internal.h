// Definitions shared by the library's C ABI translation units (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdlib>
#include <string>

#include "rasterizer_amd.h"

#define RZ_HIP(call)                                                \
  do {                                                              \
    const hipError_t rz_hip_error_ = (call);                        \
    if (rz_hip_error_ != hipSuccess) {                              \
      rzi::report_hip_error(rz_hip_error_, #call, __FILE__, __LINE__); \
      return RZ_ERR_DEVICE;                                         \
    }                                                               \
  } while (0)

// A baked occluder's device copy: Occluder::m_vertexData (Occluder.h:17-18)
// plus its reference box and bounds.
struct rz_occluder {
  rz_context* ctx;
  uint32_t* d_packets;
  uint32_t packetCount;
  float refMin[4], refMax[4], boundsMin[4], boundsMax[4];
};

// The batch-scoring MLP's device state (rz_learn.cpp).
struct rz_mlp {
  hipStream_t stream = nullptr;
  float* d_w = nullptr;     // kMlpParams
  float* d_norm = nullptr;  // 32
  float* d_grad = nullptr;
  float* d_m = nullptr;
  float* d_v = nullptr;
  float* d_x = nullptr;
  float* d_scores = nullptr;
  float* d_ds = nullptr;
  uint32_t capX = 0;
  uint32_t* d_pairs = nullptr;
  uint32_t capPairs = 0;
  float* d_loss = nullptr;
  uint32_t step = 0;
  // The model reads feature 10, the previous step's label of the batch; when
  // it does not (trained with that feature held at 0), multi-view steps
  // ordered by it skip the label tracking of the draw.
  bool usesLabels = true;
};

namespace rzi {

// Prints a failed HIP call to stderr when the environment variable RZ_DEBUG
// is set (the C ABI itself only returns RZ_ERR_DEVICE).
void report_hip_error(hipError_t error, const char* call, const char* file, int line);

// The coverage table and this host's RCPPS table, built once per process.
struct HostTables {
  int status = RZ_OK;
  uint64_t cov[4096];
  uint32_t rcp[4096];
  int rsqrtStatus = RZ_OK;  // the device bake needs this host's RSQRTPS structure
  uint32_t rsqrt[8192];
};
const HostTables& host_tables();

// Autotuning cache (rz_tune.cpp): "<device>,<shape>" -> configuration text.
std::string device_key(int device);
bool tune_lookup(const char* path, const std::string& key, std::string& value);
bool tune_store(const char* path, const std::string& key, const std::string& value);
// Cache path from the environment (RZ_TUNE_CACHE), or null.
inline const char* tune_cache_env() {
  const char* p = std::getenv("RZ_TUNE_CACHE");
  return (p && *p) ? p : nullptr;
}

}  // namespace rzi
