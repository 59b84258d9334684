// QuadDecomposition::decompose (QuadDecomposition.cpp:311-445) for meshes
// that live in HBM.  Its data-parallel half runs on the device:
//   - the half-edges (QuadDecomposition.cpp:355-365): one 64-bit key
//     (from << 32 | to) per triangle edge, radix-sorted stably with the
//     insertion number 3t + k as the value, so each directed edge's list is
//     contiguous and in the order the reference appends to it;
//   - per half-edge, the start of its reverse edge's list (a binary search);
//   - the candidate graph (:366-384): per half-edge of triangle t, the
//     triangles already listed on the reverse edge (those inserted up to t's
//     last edge), each through the 0.5-unit plane test (:326-343) with the
//     host CPU's RSQRTPS (from its table) and DPPS's summation order.  A count
//     pass and a scan give every link its position in the reference's
//     global link order; a fill pass writes the links, two adjacency entries
//     each, and a stable radix sort by triangle turns them into flat
//     adjacency lists in that order;
//   - the quads (:392-441), once the matching is known: one thread per
//     triangle, positions from a scan of which triangles write.
// The matching itself (:28-309) is a sequential augmenting-path search with
// blossom shrinking; it runs on the host (rz_quad.cpp) over the adjacency
// lists copied back, and the partners are copied to the device.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cstdint>
#include <new>
#include <vector>

#include "rasterizer_amd.h"
#include "rz_kernels.h"
#include "rz_quad.h"
#include "rz_x86.h"

namespace rzd {

namespace {

constexpr uint32_t kQdBlock = 256;
uint32_t qd_grid(size_t n) { return uint32_t((n + kQdBlock - 1) / kQdBlock); }

struct QdParams {
  const uint32_t* idx;  // 3 per triangle
  const float4* vertices;
  uint32_t nTri, nHalf;
  const uint64_t* keys;  // sorted half-edge keys
  const uint32_t* seqs;  // their insertion numbers
  const uint32_t* revFirst;  // per insertion number: first entry of the reverse edge's list
  const uint32_t* rsq;       // 8192-entry RSQRTPS table
  uint32_t* linkCount;       // count pass: links per half-edge
  const uint32_t* linkOff;   // fill pass: their first position in link order
  uint32_t* adjKey;          // fill pass: 2 entries per link (triangle, neighbour)
  uint32_t* adjVal;
  uint32_t* degree;
  uint32_t* bad;
};

__device__ __forceinline__ uint32_t next3(uint32_t k) { return k == 2u ? 0u : k + 1u; }
__device__ __forceinline__ uint32_t prev3(uint32_t k) { return k == 0u ? 2u : k - 1u; }

// Both triangles (v0, v1, v2) and (v2, v3, v0) keep v1 - v3 within 0.5 of
// their planes (QuadDecomposition.cpp:326-343): normals from cross products
// scaled by RSQRTPS of their squared length, distances as DPPS; a NaN
// distance does not exceed the bound (COMISS compares false).
__device__ bool mergeable(float4 v0, float4 v1, float4 v2, float4 v3, const uint32_t* __restrict__ rsq) {
  const V3 n0 = approx_normalize(cross3(sub3(v1, v0), sub3(v2, v0)), rsq);
  const V3 n2 = approx_normalize(cross3(sub3(v3, v2), sub3(v0, v2)), rsq);
  const V3 d = sub3(v1, v3);
  const float a = __builtin_fabsf(dp3(n0, d)), b = __builtin_fabsf(dp3(n2, d));
  return !(a > 0.5f || b > 0.5f);
}

// Half-edge h = 3t + k runs from i[k] to i[(k + 1) % 3].  Every index is
// the origin of one half-edge, so this is where indices are checked.
__global__ void __launch_bounds__(kQdBlock) qd_half_edges_kernel(const uint32_t* __restrict__ idx, uint32_t nTri,
                                                                 uint32_t nVertices, uint64_t* __restrict__ keys,
                                                                 uint32_t* __restrict__ seqs, uint32_t* __restrict__ bad) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= 3u * nTri) return;
  const uint32_t t = h / 3u, k = h - 3u * t;
  const uint32_t a = idx[h], b = idx[3u * t + next3(k)];
  if (a >= nVertices) atomicOr(bad, 1u);
  keys[h] = (uint64_t(a) << 32) | b;
  seqs[h] = h;
}

__global__ void __launch_bounds__(kQdBlock) qd_reverse_kernel(const uint64_t* __restrict__ keys,
                                                              const uint32_t* __restrict__ seqs, uint32_t nHalf,
                                                              uint32_t* __restrict__ revFirst) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nHalf) return;
  const uint64_t key = keys[p];
  const uint64_t rev = (key << 32) | (key >> 32);
  uint32_t lo = 0, hi = nHalf;  // first entry >= rev
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] < rev) lo = mid + 1u;
    else hi = mid;
  }
  revFirst[seqs[p]] = lo;
}

// The candidate links of half-edge h = 3t + k (QuadDecomposition.cpp:366-
// 384): the reverse edge's list entries inserted up to t's last edge (t
// itself included), in list order.
template <bool FILL>
__global__ void __launch_bounds__(kQdBlock) qd_links_kernel(QdParams p) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= p.nHalf) return;
  const uint32_t t = h / 3u, k = h - 3u * t;
  const uint32_t* i = p.idx + 3u * size_t(t);
  const uint32_t a = i[k], b = i[next3(k)], c = i[prev3(k)];
  const uint64_t rev = (uint64_t(b) << 32) | a;
  const float4 v0 = p.vertices[a], v2 = p.vertices[b], v3 = p.vertices[c];
  uint32_t n = 0;
  const uint32_t base = FILL ? p.linkOff[h] : 0u;
  for (uint32_t j = p.revFirst[h]; j < p.nHalf && p.keys[j] == rev; ++j) {
    const uint32_t s = p.seqs[j];
    if (s > 3u * t + 2u) break;  // listed after t
    const uint32_t nb = s / 3u, kk = s - 3u * nb;
    const uint32_t apex = p.idx[3u * size_t(nb) + prev3(kk)];
    if (!mergeable(v0, p.vertices[apex], v2, v3, p.rsq)) continue;
    if (FILL) {
      const size_t q = size_t(base) + n;
      p.adjKey[2 * q] = t;
      p.adjVal[2 * q] = nb;
      p.adjKey[2 * q + 1] = nb;
      p.adjVal[2 * q + 1] = t;
      atomicAdd(&p.degree[t], 1u);
      atomicAdd(&p.degree[nb], 1u);
    }
    ++n;
  }
  if (!FILL) p.linkCount[h] = n;
}

// Which triangles write a quad: a single one (its degenerate quad) and the
// lower triangle of a matched pair; one matched to itself writes nothing.
__global__ void __launch_bounds__(kQdBlock) qd_writers_kernel(const int* __restrict__ mate, uint32_t nTri,
                                                              uint32_t* __restrict__ writes) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nTri) return;
  if (t == nTri) {
    writes[t] = 0u;  // the scan's total
    return;
  }
  const int nb = mate[t];
  writes[t] = (nb == -1 || uint32_t(nb) > t) ? 1u : 0u;
}

// The quads in triangle order (QuadDecomposition.cpp:392-441): a single
// triangle becomes (i0, i2, i1, i0); a matched pair is written from the
// lower triangle's first edge whose reverse list holds the partner, as
// (i[k], i[k+2], i[k+1], the partner's apex on that edge).
__global__ void __launch_bounds__(kQdBlock) qd_emit_kernel(QdParams p, const int* __restrict__ mate,
                                                           const uint32_t* __restrict__ pos, uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.nTri) return;
  const int nb = mate[t];
  if (!(nb == -1 || uint32_t(nb) > t)) return;
  const uint32_t* i = p.idx + 3u * size_t(t);
  uint32_t* o = out + 4u * size_t(pos[t]);
  if (nb == -1) {
    o[0] = i[0];
    o[1] = i[2];
    o[2] = i[1];
    o[3] = i[0];
    return;
  }
  for (uint32_t k = 0; k < 3u; ++k) {
    const uint64_t rev = (uint64_t(i[next3(k)]) << 32) | i[k];
    for (uint32_t j = p.revFirst[3u * t + k]; j < p.nHalf && p.keys[j] == rev; ++j) {
      const uint32_t s = p.seqs[j];
      if (s / 3u != uint32_t(nb)) continue;
      o[0] = i[k];
      o[1] = i[prev3(k)];
      o[2] = i[next3(k)];
      o[3] = p.idx[s - s % 3u + prev3(s % 3u)];
      return;
    }
  }
  atomicOr(p.bad, 2u);  // a matched pair shares no edge: not reachable from the links
}

__global__ void __launch_bounds__(kQdBlock) qd_gather_kernel(const float4* __restrict__ vertices,
                                                             const uint32_t* __restrict__ quadIdx, uint32_t nWords,
                                                             float4* __restrict__ out) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nWords) out[w] = vertices[quadIdx[w]];
}

template <typename T>
hipError_t qd_alloc(T** p, size_t count) {
  *p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(p), (count ? count : 1) * sizeof(T));
}

struct QdBuffers {
  uint64_t* keys[2] = {nullptr, nullptr};
  uint32_t* seqs[2] = {nullptr, nullptr};
  uint32_t *revFirst = nullptr, *linkCount = nullptr, *linkOff = nullptr, *degree = nullptr, *start = nullptr,
           *bad = nullptr, *writes = nullptr, *pos = nullptr;
  uint32_t *adjKey[2] = {nullptr, nullptr}, *adjVal[2] = {nullptr, nullptr};
  int* mate = nullptr;
  void* temp = nullptr;
  ~QdBuffers() {
    for (void* q : {(void*)keys[0], (void*)keys[1], (void*)seqs[0], (void*)seqs[1], (void*)revFirst,
                    (void*)linkCount, (void*)linkOff, (void*)degree, (void*)start, (void*)bad, (void*)writes,
                    (void*)pos, (void*)adjKey[0], (void*)adjKey[1], (void*)adjVal[0], (void*)adjVal[1], (void*)mate,
                    temp})
      if (q) (void)hipFree(q);
  }
  // Temporary storage of at least `bytes` for one rocPRIM call.
  size_t tempBytes = 0;
  hipError_t reserve(size_t bytes) {
    if (bytes <= tempBytes) return hipSuccess;
    if (temp) (void)hipFree(temp);
    temp = nullptr;
    tempBytes = 0;
    const hipError_t e = hipMalloc(&temp, bytes ? bytes : 1);
    if (e == hipSuccess) tempBytes = bytes;
    return e;
  }
};

}  // namespace

#define QD_HIP(call)                                \
  do {                                              \
    if ((call) != hipSuccess) return RZ_ERR_DEVICE; \
  } while (0)

hipError_t launch_quad_gather(const float* vertices, const uint32_t* quadIdx, uint32_t nWords, float* out,
                              hipStream_t stream) {
  if (nWords == 0) return hipSuccess;
  qd_gather_kernel<<<qd_grid(nWords), kQdBlock, 0, stream>>>(reinterpret_cast<const float4*>(vertices), quadIdx,
                                                             nWords, reinterpret_cast<float4*>(out));
  return hipGetLastError();
}

int quad_decompose_run(hipStream_t stream, uint32_t nTri, const uint32_t* d_idx, const float* d_vertices,
                       uint32_t nVertices, const uint32_t* hostRsqrt, uint32_t* d_out, uint32_t* nOut) {
  *nOut = 0;
  if (nTri == 0) return RZ_OK;
  if (nTri > 0x7FFFFFFFu / 3u) return RZ_ERR_TOO_LARGE;
  const uint32_t nHalf = 3u * nTri;
  QdBuffers B;
  uint32_t* d_rsq = nullptr;
  if (qd_alloc(&B.keys[0], nHalf) != hipSuccess || qd_alloc(&B.keys[1], nHalf) != hipSuccess ||
      qd_alloc(&B.seqs[0], nHalf) != hipSuccess || qd_alloc(&B.seqs[1], nHalf) != hipSuccess ||
      qd_alloc(&B.revFirst, nHalf) != hipSuccess || qd_alloc(&B.linkCount, size_t(nHalf) + 1) != hipSuccess ||
      qd_alloc(&B.linkOff, size_t(nHalf) + 1) != hipSuccess || qd_alloc(&B.degree, size_t(nTri) + 1) != hipSuccess ||
      qd_alloc(&B.start, size_t(nTri) + 1) != hipSuccess || qd_alloc(&B.bad, 1) != hipSuccess ||
      qd_alloc(&B.writes, size_t(nTri) + 1) != hipSuccess || qd_alloc(&B.pos, size_t(nTri) + 1) != hipSuccess ||
      qd_alloc(&B.mate, nTri) != hipSuccess || qd_alloc(&d_rsq, 8192) != hipSuccess) {
    if (d_rsq) (void)hipFree(d_rsq);
    return RZ_ERR_OUT_OF_MEMORY;
  }
  struct RsqGuard {
    uint32_t* p;
    ~RsqGuard() { (void)hipFree(p); }
  } rsqGuard{d_rsq};
  QD_HIP(hipMemcpyAsync(d_rsq, hostRsqrt, 8192 * 4, hipMemcpyHostToDevice, stream));
  QD_HIP(hipMemsetAsync(B.bad, 0, 4, stream));

  // Half-edges, sorted by (key, insertion number).
  qd_half_edges_kernel<<<qd_grid(nHalf), kQdBlock, 0, stream>>>(d_idx, nTri, nVertices, B.keys[0], B.seqs[0], B.bad);
  QD_HIP(hipGetLastError());
  uint32_t bad = 0;
  QD_HIP(hipMemcpyAsync(&bad, B.bad, 4, hipMemcpyDeviceToHost, stream));
  QD_HIP(hipStreamSynchronize(stream));
  if (bad) return RZ_ERR_INVALID_ARGUMENT;
  size_t bytes = 0;
  QD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, B.keys[0], B.keys[1], B.seqs[0], B.seqs[1], nHalf, 0u, 64u,
                                   stream));
  if (B.reserve(bytes) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
  QD_HIP(rocprim::radix_sort_pairs(B.temp, bytes, B.keys[0], B.keys[1], B.seqs[0], B.seqs[1], nHalf, 0u, 64u,
                                   stream));
  qd_reverse_kernel<<<qd_grid(nHalf), kQdBlock, 0, stream>>>(B.keys[1], B.seqs[1], nHalf, B.revFirst);
  QD_HIP(hipGetLastError());

  // Candidate links: count, scan, fill.
  QdParams P{};
  P.idx = d_idx;
  P.vertices = reinterpret_cast<const float4*>(d_vertices);
  P.nTri = nTri;
  P.nHalf = nHalf;
  P.keys = B.keys[1];
  P.seqs = B.seqs[1];
  P.revFirst = B.revFirst;
  P.rsq = d_rsq;
  P.linkCount = B.linkCount;
  P.linkOff = B.linkOff;
  P.degree = B.degree;
  P.bad = B.bad;
  QD_HIP(hipMemsetAsync(B.linkCount + nHalf, 0, 4, stream));
  qd_links_kernel<false><<<qd_grid(nHalf), kQdBlock, 0, stream>>>(P);
  QD_HIP(hipGetLastError());
  bytes = 0;
  QD_HIP(rocprim::exclusive_scan(nullptr, bytes, B.linkCount, B.linkOff, 0u, size_t(nHalf) + 1,
                                 rocprim::plus<uint32_t>(), stream));
  if (B.reserve(bytes) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
  QD_HIP(rocprim::exclusive_scan(B.temp, bytes, B.linkCount, B.linkOff, 0u, size_t(nHalf) + 1,
                                 rocprim::plus<uint32_t>(), stream));
  uint32_t nLinks = 0;
  QD_HIP(hipMemcpyAsync(&nLinks, B.linkOff + nHalf, 4, hipMemcpyDeviceToHost, stream));
  QD_HIP(hipStreamSynchronize(stream));
  if (nLinks > 0x7FFFFFFFu / 2u) return RZ_ERR_TOO_LARGE;
  const uint32_t nAdj = 2u * nLinks;
  if (qd_alloc(&B.adjKey[0], nAdj) != hipSuccess || qd_alloc(&B.adjKey[1], nAdj) != hipSuccess ||
      qd_alloc(&B.adjVal[0], nAdj) != hipSuccess || qd_alloc(&B.adjVal[1], nAdj) != hipSuccess)
    return RZ_ERR_OUT_OF_MEMORY;
  P.adjKey = B.adjKey[0];
  P.adjVal = B.adjVal[0];
  QD_HIP(hipMemsetAsync(B.degree, 0, (size_t(nTri) + 1) * 4, stream));
  qd_links_kernel<true><<<qd_grid(nHalf), kQdBlock, 0, stream>>>(P);
  QD_HIP(hipGetLastError());
  // Adjacency lists: entries stably sorted by triangle keep link order.
  if (nAdj > 0) {
    uint32_t bits = 1;
    while (bits < 32 && (uint64_t(1) << bits) < nTri) ++bits;
    bytes = 0;
    QD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, B.adjKey[0], B.adjKey[1], B.adjVal[0], B.adjVal[1], nAdj, 0u,
                                     bits, stream));
    if (B.reserve(bytes) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
    QD_HIP(rocprim::radix_sort_pairs(B.temp, bytes, B.adjKey[0], B.adjKey[1], B.adjVal[0], B.adjVal[1], nAdj, 0u,
                                     bits, stream));
  }
  bytes = 0;
  QD_HIP(rocprim::exclusive_scan(nullptr, bytes, B.degree, B.start, 0u, size_t(nTri) + 1, rocprim::plus<uint32_t>(),
                                 stream));
  if (B.reserve(bytes) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
  QD_HIP(rocprim::exclusive_scan(B.temp, bytes, B.degree, B.start, 0u, size_t(nTri) + 1, rocprim::plus<uint32_t>(),
                                 stream));

  // The matching, on the host.
  std::vector<int> mate;
  try {
    std::vector<uint32_t> start(size_t(nTri) + 1);
    std::vector<int> adj(nAdj);
    QD_HIP(hipMemcpyAsync(start.data(), B.start, start.size() * 4, hipMemcpyDeviceToHost, stream));
    if (nAdj) QD_HIP(hipMemcpyAsync(adj.data(), B.adjVal[1], size_t(nAdj) * 4, hipMemcpyDeviceToHost, stream));
    QD_HIP(hipStreamSynchronize(stream));
    if (start[nTri] != nAdj) return RZ_ERR_DEVICE;
    mate.resize(nTri);
    rzq::match(nTri, start.data(), adj.data(), mate.data());
  } catch (const std::bad_alloc&) {
    return RZ_ERR_OUT_OF_MEMORY;
  }
  QD_HIP(hipMemcpyAsync(B.mate, mate.data(), size_t(nTri) * 4, hipMemcpyHostToDevice, stream));

  // The quads.
  qd_writers_kernel<<<qd_grid(size_t(nTri) + 1), kQdBlock, 0, stream>>>(B.mate, nTri, B.writes);
  QD_HIP(hipGetLastError());
  bytes = 0;
  QD_HIP(rocprim::exclusive_scan(nullptr, bytes, B.writes, B.pos, 0u, size_t(nTri) + 1, rocprim::plus<uint32_t>(),
                                 stream));
  if (B.reserve(bytes) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
  QD_HIP(rocprim::exclusive_scan(B.temp, bytes, B.writes, B.pos, 0u, size_t(nTri) + 1, rocprim::plus<uint32_t>(),
                                 stream));
  qd_emit_kernel<<<qd_grid(nTri), kQdBlock, 0, stream>>>(P, B.mate, B.pos, d_out);
  QD_HIP(hipGetLastError());
  uint32_t quads = 0;
  QD_HIP(hipMemcpyAsync(&quads, B.pos + nTri, 4, hipMemcpyDeviceToHost, stream));
  QD_HIP(hipMemcpyAsync(&bad, B.bad, 4, hipMemcpyDeviceToHost, stream));
  QD_HIP(hipStreamSynchronize(stream));
  if (bad) return RZ_ERR_DEVICE;
  *nOut = 4u * quads;
  return RZ_OK;
}

}  // namespace rzd
