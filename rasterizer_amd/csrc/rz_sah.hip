// SurfaceAreaHeuristic::generateBatches (SurfaceAreaHeuristic.cpp:10-104) on
// the device, for occluder sets built at run time.
//
// The reference recurses depth first; each call sorts its index range stably
// by box centre along x, then y, then z, scores every split position that is
// a multiple of the granularity with prefix and suffix box areas, and sorts
// the range once more along the winning axis.  Here one level of the
// recursion is one round over all open ranges ("segments") at once:
//   - each stable sort is one device-wide radix sort on (segment, centre
//     key), so segments stay where they are and sort independently;
//   - prefix and suffix boxes are segmented scans over the boxes of chunks
//     of gran positions, since every split falls between chunks (min and max
//     are exact, so neither the chunking nor the scan order changes them);
//   - the winning (cost, axis, index) of a segment is a 64-bit atomicMin,
//     which keeps the reference's first strict minimum in (axis, index) order.
// Segments stay in the index array in the order the depth-first recursion
// would emit them, so the batches are the final segments, left to right.
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan_by_key.hpp>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "rasterizer_amd.h"
#include "rz_kernels.h"

extern "C" __device__ uint32_t __ockl_wfred_min_u32(uint32_t);
extern "C" __device__ uint32_t __ockl_wfred_max_u32(uint32_t);

namespace rzd {

namespace {

constexpr uint32_t kMaxSahItems = 1u << 28;
constexpr float kMaxSahCoordinate = 1099511627776.0f;  // 2^40

struct Ext6 {
  float mn[3], mx[3];
};

struct MergeExt {
  __device__ Ext6 operator()(const Ext6& a, const Ext6& b) const {
    Ext6 r;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      r.mn[j] = fminf(a.mn[j], b.mn[j]);
      r.mx[j] = fmaxf(a.mx[j], b.mx[j]);
    }
    return r;
  }
};

__device__ __forceinline__ Ext6 load_ext(const float* __restrict__ boxes, uint32_t e) {
  const float4 lo = reinterpret_cast<const float4*>(boxes)[2 * size_t(e)];
  const float4 hi = reinterpret_cast<const float4*>(boxes)[2 * size_t(e) + 1];
  return {{lo.x, lo.y, lo.z}, {hi.x, hi.y, hi.z}};
}

// Aabb::surfaceArea (VectorMath.h:57-62): DPPS(ext, ext.yzx, 0x7F), summed
// (x + y) + (z + the masked lane's 0).
__device__ __forceinline__ float surface_area(const Ext6& b) {
  const float e0 = b.mx[0] - b.mn[0], e1 = b.mx[1] - b.mn[1], e2 = b.mx[2] - b.mn[2];
  const float p0 = e0 * e1, p1 = e1 * e2, p2 = e2 * e0;
  return (p0 + p1) + (p2 + 0.0f);
}

// Sort key of a centre coordinate: ascending in float order, with -0 and +0
// equal, as the reference's _mm_cmplt_ps comparator sees them.
__device__ __forceinline__ uint32_t centre_key(float c) {
  const uint32_t u = __float_as_uint(c == 0.0f ? 0.0f : c);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void __launch_bounds__(256) sah_check_kernel(const float* __restrict__ boxes, uint32_t n,
                                                        uint32_t* __restrict__ bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Ext6 b = load_ext(boxes, i);
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    ok = ok && fabsf(b.mn[j]) <= kMaxSahCoordinate && fabsf(b.mx[j]) <= kMaxSahCoordinate && b.mn[j] <= b.mx[j];
  if (!ok) *bad = 1u;
}

__global__ void __launch_bounds__(256) sah_iota_kernel(uint32_t* __restrict__ idx, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) idx[i] = i;
}

// Segment of every position: the last segment start at or before it.
__global__ void __launch_bounds__(256) sah_segof_kernel(const uint32_t* __restrict__ segStart, uint32_t nSeg,
                                                        uint32_t n, uint32_t* __restrict__ segOf) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t lo = 0, hi = nSeg - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (segStart[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  segOf[i] = lo;
}

// Radix keys (segment << 32 | centre key along the axis).  Closed segments
// get a constant key and keep their order; axis < 0 takes each open
// segment's winning axis.
__global__ void __launch_bounds__(256) sah_keys_kernel(const float* __restrict__ boxes, const uint32_t* __restrict__ idx,
                                                       const uint32_t* __restrict__ segOf,
                                                       const uint32_t* __restrict__ segOpen,
                                                       const uint64_t* __restrict__ segBest, int axis, uint32_t n,
                                                       uint64_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = segOf[i];
  uint32_t low = 0;
  if (segOpen[s]) {
    const int a = axis >= 0 ? axis : int((segBest[s] >> 30) & 3u);
    const uint32_t e = idx[i];
    low = centre_key(boxes[8 * size_t(e) + a] + boxes[8 * size_t(e) + 4 + a]);  // Aabb::getCenter
  }
  keys[i] = (uint64_t(s) << 32) | low;
}

// Boxes of each chunk of gran consecutive positions, in sorted order,
// forwards and reversed (the suffix scan runs over the reversed array).
// Every segment starts at a multiple of gran (the root at 0, each split gran
// times an integer further on), so no chunk spans two segments.
__global__ void __launch_bounds__(256) sah_chunks_kernel(const float* __restrict__ boxes,
                                                         const uint32_t* __restrict__ idx,
                                                         const uint32_t* __restrict__ segOf, uint32_t n,
                                                         uint32_t gran, uint32_t nChunks, Ext6* __restrict__ fwd,
                                                         Ext6* __restrict__ rev, uint32_t* __restrict__ chunkSeg,
                                                         uint32_t* __restrict__ revSeg) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nChunks) return;
  const uint32_t i0 = c * gran, i1 = min(i0 + gran, n);
  Ext6 b = load_ext(boxes, idx[i0]);
  const MergeExt merge;
  for (uint32_t i = i0 + 1; i < i1; ++i) b = merge(b, load_ext(boxes, idx[i]));
  const uint32_t sg = segOf[i0];
  fwd[c] = b;
  rev[nChunks - 1 - c] = b;
  chunkSeg[c] = sg;
  revSeg[nChunks - 1 - c] = sg;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(v)), off));
    const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(v >> 32)), off));
    const uint64_t o = (uint64_t(hi) << 32) | lo;
    v = o < v ? o : v;
  }
  return v;
}

// sahSplit's cost loop (SurfaceAreaHeuristic.cpp:46-64) for one axis: split
// position k of a segment of len items costs area(items [0, k)) * k +
// area(items [k, len)) * (len - k), for k = gran, 2 gran, ... < len - gran.
// The key orders by cost (non-negative, -0 folded into +0), then axis, then
// k, so the minimum is the reference's first strict minimum.  Keys are
// reduced per wave and per work-group while they share a segment, so a large
// segment takes one atomic per work-group rather than one per candidate.
__global__ void __launch_bounds__(256) sah_cost_kernel(const Ext6* __restrict__ fwdScan,
                                                       const Ext6* __restrict__ revScan,
                                                       const uint32_t* __restrict__ chunkSeg,
                                                       const uint32_t* __restrict__ segStart,
                                                       const uint32_t* __restrict__ segLen,
                                                       const uint32_t* __restrict__ segOpen, uint32_t nChunks,
                                                       uint32_t gran, uint32_t axis, uint64_t* __restrict__ segBest) {
  __shared__ uint64_t waveBest[4];
  __shared__ uint32_t waveSeg[4];
  // One thread per chunk c: the split before position c * gran, whose left
  // part ends with chunk c - 1 and whose right part starts with chunk c.
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const bool inRange = c < nChunks;
  const uint32_t s = inRange ? chunkSeg[c] : 0u;
  uint64_t key = ~0ull;
  if (inRange && segOpen[s]) {
    const uint32_t k = c * gran - segStart[s], len = segLen[s];
    if (k >= gran && k + gran < len) {
      const float left = surface_area(fwdScan[c - 1]) * float(int(k));
      const float right = surface_area(revScan[nChunks - 1 - c]) * float(int(len - k));
      float cost = left + right;
      if (cost == 0.0f) cost = 0.0f;
      key = (uint64_t(__float_as_uint(cost)) << 32) | (uint64_t(axis) << 30) | k;
    }
  }
  // Lanes are consecutive chunks, so lane 0 is in range whenever any is.
  const uint32_t s0 = uint32_t(__builtin_amdgcn_readfirstlane(int(s)));
  const bool uniform = __ballot(inRange && s != s0) == 0ull;
  uint64_t wk = ~0ull;
  if (uniform) {
    wk = wave_min_u64(key);
  } else if (key != ~0ull) {
    atomicMin(reinterpret_cast<unsigned long long*>(&segBest[s]), static_cast<unsigned long long>(key));
  }
  if (lane == 0) {
    waveBest[wave] = wk;
    waveSeg[wave] = s0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t cur = ~0ull;
    uint32_t curSeg = 0;
    for (uint32_t w = 0; w < 4; ++w) {
      if (waveBest[w] == ~0ull) continue;
      if (cur != ~0ull && waveSeg[w] != curSeg) {
        atomicMin(reinterpret_cast<unsigned long long*>(&segBest[curSeg]), static_cast<unsigned long long>(cur));
        cur = ~0ull;
      }
      curSeg = waveSeg[w];
      cur = waveBest[w] < cur ? waveBest[w] : cur;
    }
    if (cur != ~0ull)
      atomicMin(reinterpret_cast<unsigned long long*>(&segBest[curSeg]), static_cast<unsigned long long>(cur));
  }
}

uint32_t grid_of(uint32_t n) { return (n + 255u) / 256u; }

template <typename T>
hipError_t alloc(T** p, size_t count) {
  *p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(p), (count ? count : 1) * sizeof(T));
}

struct SahBuffers {
  uint32_t* idx[2] = {nullptr, nullptr};
  uint64_t* keys[2] = {nullptr, nullptr};
  uint32_t *segOf = nullptr, *chunkSeg = nullptr, *revSeg = nullptr, *segStart = nullptr, *segLen = nullptr,
           *segOpen = nullptr;
  uint32_t* bad = nullptr;
  uint64_t* segBest = nullptr;
  Ext6 *fwd = nullptr, *fwdScan = nullptr, *rev = nullptr, *revScan = nullptr;
  void* temp = nullptr;
  ~SahBuffers() {
    for (void* p : {(void*)idx[0], (void*)idx[1], (void*)keys[0], (void*)keys[1], (void*)segOf, (void*)chunkSeg,
                    (void*)revSeg,
                    (void*)segStart, (void*)segLen, (void*)segOpen, (void*)bad, (void*)segBest, (void*)fwd,
                    (void*)fwdScan, (void*)rev, (void*)revScan, temp})
      if (p) (void)hipFree(p);
  }
};

}  // namespace

#define SAH_HIP(call)                                  \
  do {                                                 \
    if ((call) != hipSuccess) return RZ_ERR_DEVICE;    \
  } while (0)

int sah_batches_run(hipStream_t stream, uint32_t n, const float* d_boxes, uint32_t target, uint32_t gran,
                    uint32_t* d_indices, uint32_t* batchOffsets, uint32_t* nBatches) {
  if (gran == 0 || target <= 2 * uint64_t(gran)) return RZ_ERR_INVALID_ARGUMENT;
  if (n > kMaxSahItems) return RZ_ERR_TOO_LARGE;
  batchOffsets[0] = 0;
  if (n == 0) {
    *nBatches = 0;
    return RZ_OK;
  }
  SahBuffers B;
  if (alloc(&B.bad, 1) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
  uint32_t bad = 0;
  SAH_HIP(hipMemsetAsync(B.bad, 0, 4, stream));
  sah_check_kernel<<<grid_of(n), 256, 0, stream>>>(d_boxes, n, B.bad);
  SAH_HIP(hipGetLastError());
  SAH_HIP(hipMemcpyAsync(&bad, B.bad, 4, hipMemcpyDeviceToHost, stream));
  SAH_HIP(hipStreamSynchronize(stream));
  if (bad) return RZ_ERR_INVALID_ARGUMENT;
  if (n <= 2 * gran) {
    // No split position exists, so the reference's recursion is undefined
    // here (bestIndex stays -1); the whole set is one batch in input order.
    sah_iota_kernel<<<grid_of(n), 256, 0, stream>>>(d_indices, n);
    SAH_HIP(hipGetLastError());
    SAH_HIP(hipStreamSynchronize(stream));
    batchOffsets[1] = n;
    *nBatches = 1;
    return RZ_OK;
  }
  // At most n / gran + 1 segments: every closed piece holds gran items or more.
  const size_t maxSeg = size_t(n) / gran + 2;
  const uint32_t nChunks = (n + gran - 1) / gran;
  if (alloc(&B.idx[0], n) != hipSuccess || alloc(&B.idx[1], n) != hipSuccess || alloc(&B.keys[0], n) != hipSuccess ||
      alloc(&B.keys[1], n) != hipSuccess || alloc(&B.segOf, n) != hipSuccess ||
      alloc(&B.chunkSeg, nChunks) != hipSuccess || alloc(&B.revSeg, nChunks) != hipSuccess ||
      alloc(&B.segStart, maxSeg) != hipSuccess || alloc(&B.segLen, maxSeg) != hipSuccess ||
      alloc(&B.segOpen, maxSeg) != hipSuccess || alloc(&B.segBest, maxSeg) != hipSuccess ||
      alloc(&B.fwd, nChunks) != hipSuccess || alloc(&B.fwdScan, nChunks) != hipSuccess ||
      alloc(&B.rev, nChunks) != hipSuccess || alloc(&B.revScan, nChunks) != hipSuccess)
    return RZ_ERR_OUT_OF_MEMORY;
  size_t sortBytes = 0, scanBytes = 0;
  SAH_HIP(rocprim::radix_sort_pairs(nullptr, sortBytes, B.keys[0], B.keys[1], B.idx[0], B.idx[1], n, 0u, 64u,
                                    stream));
  SAH_HIP(rocprim::inclusive_scan_by_key(nullptr, scanBytes, B.chunkSeg, B.fwd, B.fwdScan, size_t(nChunks), MergeExt(),
                                         rocprim::equal_to<uint32_t>(), stream));
  size_t tempBytes = std::max(sortBytes, scanBytes);
  if (alloc(reinterpret_cast<uint8_t**>(&B.temp), tempBytes) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;

  sah_iota_kernel<<<grid_of(n), 256, 0, stream>>>(B.idx[0], n);
  SAH_HIP(hipGetLastError());
  int cur = 0;
  // The root is split whatever its size (generateBatches calls the
  // recursion on it directly); a child recurses when it holds target items
  // or more (SurfaceAreaHeuristic.cpp:81-92).
  std::vector<uint32_t> start{0}, len{n}, open{1};
  std::vector<uint64_t> best;
  while (true) {
    const uint32_t S = uint32_t(start.size());
    uint32_t bits = 0;
    while ((1ull << bits) < S) ++bits;
    SAH_HIP(hipMemcpyAsync(B.segStart, start.data(), S * 4, hipMemcpyHostToDevice, stream));
    SAH_HIP(hipMemcpyAsync(B.segLen, len.data(), S * 4, hipMemcpyHostToDevice, stream));
    SAH_HIP(hipMemcpyAsync(B.segOpen, open.data(), S * 4, hipMemcpyHostToDevice, stream));
    SAH_HIP(hipMemsetAsync(B.segBest, 0xFF, S * 8, stream));
    sah_segof_kernel<<<grid_of(n), 256, 0, stream>>>(B.segStart, S, n, B.segOf);
    SAH_HIP(hipGetLastError());
    for (int axis = 0; axis <= 3; ++axis) {
      // Axes 0-2 score the splits; the fourth pass re-sorts along the winner.
      sah_keys_kernel<<<grid_of(n), 256, 0, stream>>>(d_boxes, B.idx[cur], B.segOf, B.segOpen, B.segBest,
                                                      axis < 3 ? axis : -1, n, B.keys[0]);
      SAH_HIP(hipGetLastError());
      size_t bytes = tempBytes;
      SAH_HIP(rocprim::radix_sort_pairs(B.temp, bytes, B.keys[0], B.keys[1], B.idx[cur], B.idx[cur ^ 1], n, 0u,
                                        32u + bits, stream));
      cur ^= 1;
      if (axis == 3) break;
      sah_chunks_kernel<<<grid_of(nChunks), 256, 0, stream>>>(d_boxes, B.idx[cur], B.segOf, n, gran, nChunks,
                                                              B.fwd, B.rev, B.chunkSeg, B.revSeg);
      SAH_HIP(hipGetLastError());
      bytes = tempBytes;
      SAH_HIP(rocprim::inclusive_scan_by_key(B.temp, bytes, B.chunkSeg, B.fwd, B.fwdScan, size_t(nChunks), MergeExt(),
                                             rocprim::equal_to<uint32_t>(), stream));
      bytes = tempBytes;
      SAH_HIP(rocprim::inclusive_scan_by_key(B.temp, bytes, B.revSeg, B.rev, B.revScan, size_t(nChunks), MergeExt(),
                                             rocprim::equal_to<uint32_t>(), stream));
      sah_cost_kernel<<<grid_of(nChunks), 256, 0, stream>>>(B.fwdScan, B.revScan, B.chunkSeg, B.segStart, B.segLen,
                                                            B.segOpen, nChunks, gran, uint32_t(axis), B.segBest);
      SAH_HIP(hipGetLastError());
    }
    best.resize(S);
    SAH_HIP(hipMemcpyAsync(best.data(), B.segBest, S * 8, hipMemcpyDeviceToHost, stream));
    SAH_HIP(hipStreamSynchronize(stream));
    std::vector<uint32_t> nStart, nLen, nOpen;
    bool any = false;
    for (uint32_t s = 0; s < S; ++s) {
      if (!open[s]) {
        nStart.push_back(start[s]);
        nLen.push_back(len[s]);
        nOpen.push_back(0);
        continue;
      }
      // Every open segment has a split position: the root holds more than
      // 2 gran items and a child stays open only with target > 2 gran items.
      if (best[s] == ~0ull) return RZ_ERR_DEVICE;
      const uint32_t k = uint32_t(best[s] & 0x3FFFFFFFu);
      const uint32_t parts[2][2] = {{start[s], k}, {start[s] + k, len[s] - k}};
      for (const auto& part : parts) {
        const uint32_t o = part[1] >= target ? 1u : 0u;
        nStart.push_back(part[0]);
        nLen.push_back(part[1]);
        nOpen.push_back(o);
        any = any || o;
      }
    }
    start.swap(nStart);
    len.swap(nLen);
    open.swap(nOpen);
    if (!any) break;
  }
  SAH_HIP(hipMemcpyAsync(d_indices, B.idx[cur], size_t(n) * 4, hipMemcpyDeviceToDevice, stream));
  SAH_HIP(hipStreamSynchronize(stream));
  for (size_t s = 0; s < start.size(); ++s) batchOffsets[s + 1] = start[s] + len[s];
  *nBatches = uint32_t(start.size());
  return RZ_OK;
}

// ---------------------------------------------------------------------------
// Scene preparation around the SAH pass (Main.cpp:86-128): padded quad
// boxes, the scene box, and the vertices gathered into batch order.
// ---------------------------------------------------------------------------
namespace {

// Ascending-order key of a float (for exact, order-independent min and max
// with atomics; -0 sorts below +0).
__device__ __forceinline__ uint32_t ordered_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Quad j < nPad: padded quads (j >= nQuads) are four copies of vertex 0, as
// Main.cpp:90-94 pads the index buffer with indices[0].
__device__ __forceinline__ float4 padded_vertex(const float4* __restrict__ v, uint32_t nQuads, uint32_t j, uint32_t k) {
  return j < nQuads ? v[4 * size_t(j) + k] : v[0];
}

__global__ void __launch_bounds__(256) scene_boxes_kernel(const float4* __restrict__ vertices, uint32_t nQuads,
                                                          uint32_t nPad, float4* __restrict__ boxes,
                                                          uint32_t* __restrict__ sceneKeys) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  float4 mn = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
  float4 mx = make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
  uint32_t kmn[4] = {~0u, ~0u, ~0u, ~0u}, kmx[4] = {0u, 0u, 0u, 0u};
  bool bad = false;
  if (j < nPad) {
    // Main.cpp:96-105: the quad's Aabb over its four vertices (all lanes).
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const float4 v = padded_vertex(vertices, nQuads, j, k);
      mn = make_float4(fminf(mn.x, v.x), fminf(mn.y, v.y), fminf(mn.z, v.z), fminf(mn.w, v.w));
      mx = make_float4(fmaxf(mx.x, v.x), fmaxf(mx.y, v.y), fmaxf(mx.z, v.z), fmaxf(mx.w, v.w));
      const float c[4] = {v.x, v.y, v.z, v.w};
      // The SAH costs need finite boxes of bounded size (as rz_sah_batches
      // checks); NaN would also escape the fminf/fmaxf boxes.
      bad = bad || !(fabsf(v.x) <= 0x1p40f) || !(fabsf(v.y) <= 0x1p40f) || !(fabsf(v.z) <= 0x1p40f);
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        kmn[l] = min(kmn[l], ordered_key(c[l]));
        kmx[l] = max(kmx[l], ordered_key(c[l]));
      }
    }
    boxes[2 * size_t(j)] = mn;
    boxes[2 * size_t(j) + 1] = mx;
  }
  if (__ballot(bad) != 0ull && (threadIdx.x & 63u) == 0u) atomicOr(&sceneKeys[8], 1u);
  // Main.cpp:109-113: the scene box over every vertex, as ordered keys.
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const uint32_t a = __ockl_wfred_min_u32(kmn[l]), b = __ockl_wfred_max_u32(kmx[l]);
    if ((threadIdx.x & 63u) == 0u) {
      atomicMin(&sceneKeys[l], a);
      atomicMax(&sceneKeys[4 + l], b);
    }
  }
}

// Main.cpp:109-113 over a whole vertex buffer (which can hold vertices no
// quad references): the scene box as ordered keys, and sceneKeys[8] set when
// a vertex's x, y or z is not finite or exceeds 2^40.
__global__ void __launch_bounds__(256) vertex_keys_kernel(const float4* __restrict__ vertices, uint32_t n,
                                                          uint32_t* __restrict__ sceneKeys) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t kmn[4] = {~0u, ~0u, ~0u, ~0u}, kmx[4] = {0u, 0u, 0u, 0u};
  bool bad = false;
  if (j < n) {
    const float4 v = vertices[j];
    const float c[4] = {v.x, v.y, v.z, v.w};
    bad = !(fabsf(v.x) <= 0x1p40f) || !(fabsf(v.y) <= 0x1p40f) || !(fabsf(v.z) <= 0x1p40f);
#pragma unroll
    for (int l = 0; l < 4; ++l) kmn[l] = kmx[l] = ordered_key(c[l]);
  }
  if (__ballot(bad) != 0ull && (threadIdx.x & 63u) == 0u) atomicOr(&sceneKeys[8], 1u);
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const uint32_t a = __ockl_wfred_min_u32(kmn[l]), b = __ockl_wfred_max_u32(kmx[l]);
    if ((threadIdx.x & 63u) == 0u) {
      atomicMin(&sceneKeys[l], a);
      atomicMax(&sceneKeys[4 + l], b);
    }
  }
}

__global__ void __launch_bounds__(256) scene_gather_kernel(const float4* __restrict__ vertices, uint32_t nQuads,
                                                           uint32_t nPad, const uint32_t* __restrict__ order,
                                                           float4* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * nPad) return;
  out[t] = padded_vertex(vertices, nQuads, order[t >> 2], t & 3u);
}

}  // namespace

hipError_t launch_scene_boxes(const float* vertices, uint32_t nQuads, uint32_t nPad, float* boxes,
                              uint32_t* sceneKeys, hipStream_t stream) {
  scene_boxes_kernel<<<grid_of(nPad), 256, 0, stream>>>(reinterpret_cast<const float4*>(vertices), nQuads, nPad,
                                                        reinterpret_cast<float4*>(boxes), sceneKeys);
  return hipGetLastError();
}

hipError_t launch_vertex_keys(const float* vertices, uint32_t n, uint32_t* sceneKeys, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  vertex_keys_kernel<<<grid_of(n), 256, 0, stream>>>(reinterpret_cast<const float4*>(vertices), n, sceneKeys);
  return hipGetLastError();
}

hipError_t launch_scene_gather(const float* vertices, uint32_t nQuads, uint32_t nPad, const uint32_t* order,
                               float* out, hipStream_t stream) {
  scene_gather_kernel<<<grid_of(4 * nPad), 256, 0, stream>>>(reinterpret_cast<const float4*>(vertices), nQuads, nPad,
                                                             order, reinterpret_cast<float4*>(out));
  return hipGetLastError();
}

}  // namespace rzd
