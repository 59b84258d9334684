// Occluder::bake (Occluder.cpp:7-181) on the device, for occluders baked at
// run time: per-quad normals, the reference's k-means over six axis seeds,
// the stable regroup by cluster, the 11/11/10-bit packing against the
// reference box, and the bounds.  One wave per occluder.  Bit-exact with the
// host bake: SSE arithmetic is IEEE single precision (compiled without
// contraction), RSQRTPS comes from the host's table, and every sum that the
// reference makes in order is made in the same order.
#include <hip/hip_runtime.h>

#include "rz_kernels.h"
#include "rz_x86.h"

namespace rzd {

namespace {

// First occurrence of the smallest (largest) value in vertex order, as the
// reference's sequential _mm_min_ps(v, acc) (_mm_max_ps) keeps it: a value
// replaces the accumulator only when strictly smaller; NaNs never do.
struct Ext {
  float v;
  uint32_t i;
};
__device__ __forceinline__ Ext pick_min(Ext a, Ext b) { return (b.v < a.v || (b.v == a.v && b.i < a.i)) ? b : a; }
__device__ __forceinline__ Ext pick_max(Ext a, Ext b) { return (b.v > a.v || (b.v == a.v && b.i < a.i)) ? b : a; }
template <bool MIN>
__device__ __forceinline__ float wave_extreme(Ext e) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    Ext o{__shfl_xor(e.v, off), uint32_t(__shfl_xor(int(e.i), off))};
    e = MIN ? pick_min(e, o) : pick_max(e, o);
  }
  return e.v;
}

}  // namespace

__global__ void __launch_bounds__(64) rz_bake_kernel(BakeParams p) {
  const uint32_t o = blockIdx.x, lane = threadIdx.x;
  const uint32_t v0 = p.offsets[o], n = p.offsets[o + 1] - v0, quads = n / 4;
  const float4* V = reinterpret_cast<const float4*>(p.vertices) + v0;
  float4* N = p.normals + v0 / 4;
  uint32_t* C = p.cluster + v0 / 4;
  uint32_t* ord = p.order + v0 / 4;

  // Quad normals: the two triangle normals summed, then normalized
  // (Occluder.cpp:11-21).
  for (uint32_t q = lane; q < quads; q += 64) {
    const float4 P0 = V[4 * q], P1 = V[4 * q + 1], P2 = V[4 * q + 2], P3 = V[4 * q + 3];
    const V3 a = cross3(sub3(P1, P0), sub3(P2, P0)), b = cross3(sub3(P2, P0), sub3(P3, P0));
    const V3 s = approx_normalize({a.x + b.x, a.y + b.y, a.z + b.z}, p.rsqrt);
    N[q] = make_float4(s.x, s.y, s.z, 0.0f);
    C[q] = 0u;
  }
  __syncthreads();

  // K-means over six axis seeds, at most 10 rounds (Occluder.cpp:23-78):
  // each quad takes the last seed with the largest dot product; each seed
  // becomes the normalized sum of its quads' normals, summed in quad order
  // (lane k sums seed k).
  V3 seed[6] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, -1, 0}, {0, 0, -1}, {-1, 0, 0}};
  bool changed = true;
  for (int round = 0; round < 10 && changed; ++round) {
    bool mine = false;
    for (uint32_t q = lane; q < quads; q += 64) {
      const float4 nq = N[q];
      const V3 v{nq.x, nq.y, nq.z};
      float best = -__builtin_inff();
      uint32_t pick = 0;
#pragma unroll
      for (uint32_t k = 0; k < 6; ++k) {
        const float d = dp3(seed[k], v);
        if (d >= best) {
          best = d;
          pick = k;
        }
      }
      if (C[q] != pick) {
        C[q] = pick;
        mine = true;
      }
    }
    changed = __ballot(mine) != 0ull;
    __syncthreads();
    V3 sum{0.0f, 0.0f, 0.0f};
    if (lane < 6)
      for (uint32_t q = 0; q < quads; ++q)
        if (C[q] == lane) {
          const float4 nq = N[q];
          sum = {sum.x + nq.x, sum.y + nq.y, sum.z + nq.z};
        }
    const V3 c = approx_normalize(sum, p.rsqrt);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      seed[k] = {__shfl(c.x, k), __shfl(c.y, k), __shfl(c.z, k)};
    __syncthreads();
  }

  // Stable regroup by cluster (Occluder.cpp:80-93).
  uint32_t at = 0;
  for (uint32_t k = 0; k < 6; ++k)
    for (uint32_t base = 0; base < quads; base += 64) {
      const uint32_t q = base + lane;
      const bool in = q < quads && C[q] == k;
      const uint64_t m = __ballot(in);
      const uint32_t pre = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
      if (in) ord[at + pre] = q;
      at += uint32_t(__popcll(m));
    }
  __syncthreads();

  // Packing (Occluder.cpp:95-157): word (4b + k) * 8 + l is vertex k of
  // regrouped quad 8b + l, quantized to 11/11/10 bits of the reference box.
  const float* ref = p.ref + 8 * o;
  const float lo[3] = {ref[0], ref[1], ref[2]};
  const float inv[3] = {1.0f / (ref[4] - ref[0]), 1.0f / (ref[5] - ref[1]), 1.0f / (ref[6] - ref[2])};
  for (uint32_t w = lane; w < n; w += 64) {
    const uint32_t b = w >> 5, k = (w >> 3) & 3u, l = w & 7u;
    const float4 vx = V[4 * ord[8 * b + l] + k];
    const float tx = (vx.x - lo[0]) * inv[0], ty = (vx.y - lo[1]) * inv[1], tz = (vx.z - lo[2]) * inv[2];
    const uint32_t qx = uint32_t(cvtt_x86(__builtin_fmaf(tx, 2047.0f, 0.5f)) - 1024);
    const uint32_t qy = uint32_t(cvtt_x86(__builtin_fmaf(ty, 2047.0f, 0.5f)));
    const uint32_t qz = uint32_t(cvtt_x86(__builtin_fmaf(tz, 1023.0f, 0.5f)));
    p.packets[v0 + w] = (qx << 21) | (qy << 10) | qz;
  }

  // Bounds over the input vertices, w set to 1, and the centre
  // (Occluder.cpp:159-178).
  Ext mn[3], mx[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    mn[a] = {__builtin_inff(), 0xFFFFFFFFu};
    mx[a] = {-__builtin_inff(), 0xFFFFFFFFu};
  }
  for (uint32_t i = lane; i < n; i += 64) {
    const float4 vx = V[i];
    const float c[3] = {vx.x, vx.y, vx.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (c[a] < mn[a].v) mn[a] = {c[a], i};
      if (c[a] > mx[a].v) mx[a] = {c[a], i};
    }
  }
  float bmin[3], bmax[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    bmin[a] = wave_extreme<true>(mn[a]);
    bmax[a] = wave_extreme<false>(mx[a]);
  }
  if (lane == 0) {
    float* out = p.bounds + 12 * o;
    for (int a = 0; a < 3; ++a) {
      out[a] = bmin[a];
      out[4 + a] = bmax[a];
      out[8 + a] = (bmax[a] + bmin[a]) * 0.5f;
    }
    out[3] = out[7] = 1.0f;
    out[11] = (1.0f + 1.0f) * 0.5f;
  }
}

hipError_t launch_bake(const BakeParams& p, uint32_t nOccluders, hipStream_t stream) {
  if (nOccluders == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_bake_kernel, dim3(nOccluders), dim3(64), 0, stream, p);
  return hipGetLastError();
}

}  // namespace rzd
