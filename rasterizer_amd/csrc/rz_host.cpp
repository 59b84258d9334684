// Host-side parts of the rasterizer library: the coverage table built by the
// reference's constructor, the RCPPS table of the host CPU, the prebaked MVP
// and Occluder::bake.  Compiled for x86-64 with -mavx2 -mfma
// -ffp-contract=off; the x86 approximations (RCPPS, RSQRTPS) are taken from
// the CPU this library runs on, which is the CPU the reference would run on.
#include "rz_host.h"

#include <immintrin.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace rzh {
namespace {

inline uint32_t as_u32(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float as_f32(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

struct V4 {
  float v[4];
};
inline V4 load4(const float* p) { return V4{{p[0], p[1], p[2], p[3]}}; }
inline __m128 to_m(const V4& a) { return _mm_loadu_ps(a.v); }
inline V4 from_m(__m128 m) {
  V4 r;
  _mm_storeu_ps(r.v, m);
  return r;
}

// Pixel (x, y) sampled at bit 8x + y is stored at bit 8x + 4 + y/2 (even y) or
// 8x + (y-1)/2 (odd y): the row-pair interleave the block loop's mask expansion
// reads (Rasterizer.cpp:527-545 and :1259-1267).
uint64_t interleave_rows(uint64_t sampled) {
  uint64_t out = 0;
  for (int bit = 0; bit < 64; ++bit) {
    if (!((sampled >> bit) & 1)) continue;
    int x = bit >> 3, y = bit & 7;
    int dst = 8 * x + ((y & 1) ? (y >> 1) : 4 + (y >> 1));
    out |= uint64_t(1) << dst;
  }
  return out;
}

}  // namespace

// Rasterizer::precomputeRasterizationTable (Rasterizer.cpp:547-604): for each
// of 2000 edge angles in [-0.1, 6.3] and 2000 offsets in [-0.6, 0.6], the set
// of block pixels with (x-3.5)/8*nx + (y-3.5)/8*ny + offset <= 0 is OR-ed into
// the entry (slope bucket, offset bucket).  Returns false when the
// reference's own invariant (bucket 0 full, bucket 63 empty) fails.
bool build_coverage_table(uint64_t* table) {
  constexpr float kMinEdge = -0.45f;
  constexpr float kMaxEdge = 0.45f;
  const float offsetScale = 63 / (kMaxEdge - kMinEdge);
  const float offsetBias = 0.5f - kMinEdge * offsetScale;
  std::memset(table, 0, 4096 * sizeof(uint64_t));
  bool ok = true;
  for (uint32_t a = 0; a < 2000; ++a) {
    const float angle = -0.1f + 6.4f * float(a) / float(1999);
    float nx = std::cos(angle);
    float ny = std::sin(angle);
    const float norm = 1.0f / (std::fabs(nx) + std::fabs(ny));
    nx *= norm;
    ny *= norm;
    // 128-bit quantizeSlopeLookup (:470-480): ny < 0 selects the odd bucket.
    const int32_t q = _mm_cvttss_si32(_mm_set_ss(std::fma(nx, 15.5f, 16.0f)));
    const uint32_t slope = ((uint32_t(q) << 1) + (ny < 0.0f ? 1u : 0u)) << 6;
    for (uint32_t o = 0; o < 2000; ++o) {
      const float offset = -0.6f + 1.2f * float(o) / float(1999);
      int32_t bucket = _mm_cvttss_si32(_mm_set_ss(offset * offsetScale + offsetBias));
      bucket = bucket < 0 ? 0 : (bucket > 63 ? 63 : bucket);
      uint64_t sampled = 0;
      for (int x = 0; x < 8; ++x) {
        const float fx = (float(x) - 3.5f) / 8.0f * nx;
        for (int y = 0; y < 8; ++y) {
          const float d = offset + fx + (float(y) - 3.5f) / 8.0f * ny;
          if (d <= 0.0f) sampled |= uint64_t(1) << (8 * x + y);
        }
      }
      table[slope | uint32_t(bucket)] |= interleave_rows(sampled);
    }
    if (table[slope] != ~uint64_t(0) || table[slope + 63] != 0) ok = false;
  }
  return ok;
}

// RCPPS of this CPU.  Each probed x86 implementation returns a value that
// depends only on the sign, the exponent and the top 11 (Intel) or 12 (AMD
// Zen) mantissa bits, scales exactly with the exponent, returns +-inf for
// zero and denormal inputs, +-0 for infinities and for results below the
// normal range, and the quieted input for NaN.  The table holds RCPPS(1.m)
// for the 4096 values of the top 12 mantissa bits.  Returns 0 when this CPU
// follows that structure over all 2^23 mantissas and a sweep of exponents,
// -1 otherwise.
int build_rcp_table(uint32_t* table) {
  auto rcp = [](float x) { return as_u32(_mm_cvtss_f32(_mm_rcp_ps(_mm_set1_ps(x)))); };
  for (uint32_t i = 0; i < 4096; ++i) table[i] = rcp(as_f32(0x3F800000u | (i << 11)));
  for (uint32_t m = 0; m < (1u << 23); ++m) {
    if (rcp(as_f32(0x3F800000u | m)) != table[m >> 11]) return -1;
  }
  auto predict = [&](uint32_t u) -> uint32_t {
    uint32_t s = u & 0x80000000u, e = (u >> 23) & 0xFFu, m = u & 0x7FFFFFu;
    if (e == 0xFF) return m ? (u | 0x400000u) : s;
    if (e == 0) return s | 0x7F800000u;
    uint32_t t = table[m >> 11];
    int re = int((t >> 23) & 0xFF) + 127 - int(e);
    if (re <= 0) return s;
    return s | (uint32_t(re) << 23) | (t & 0x7FFFFFu);
  };
  for (uint32_t e = 0; e < 256; ++e)
    for (uint32_t m = 0; m < (1u << 23); m += 4093) {
      for (uint32_t s = 0; s < 2; ++s) {
        uint32_t u = (s << 31) | (e << 23) | m;
        if (e == 0xFF && m) u |= 0x400000u;  // probe quiet NaNs only
        if (rcp(as_f32(u)) != predict(u)) return -1;
      }
    }
  return 0;
}

int build_rsqrt_table(uint32_t* table) {
  auto rsq = [](float x) { return as_u32(_mm_cvtss_f32(_mm_rsqrt_ps(_mm_set1_ps(x)))); };
  for (uint32_t p = 0; p < 2; ++p)
    for (uint32_t i = 0; i < 4096; ++i) table[p * 4096 + i] = rsq(as_f32(((127u + p) << 23) | (i << 11)));
  for (uint32_t p = 0; p < 2; ++p)
    for (uint32_t m = 0; m < (1u << 23); ++m)
      if (rsq(as_f32(((127u + p) << 23) | m)) != table[p * 4096 + (m >> 11)]) return -1;
  auto predict = [&](uint32_t u) -> uint32_t {
    const uint32_t s = u & 0x80000000u, e = (u >> 23) & 0xFFu, m = u & 0x7FFFFFu;
    if (e == 0xFF) return m ? (u | 0x400000u) : (s ? 0xFFC00000u : 0u);
    if (e == 0) return s | 0x7F800000u;
    if (s) return 0xFFC00000u;
    const int E = int(e) - 127, p = E & 1, k = (E - p) / 2;
    const uint32_t t = table[uint32_t(p) * 4096 + (m >> 11)];
    return (uint32_t(int((t >> 23) & 0xFF) - k) << 23) | (t & 0x7FFFFFu);
  };
  for (uint32_t e = 0; e < 256; ++e)
    for (uint32_t m = 0; m < (1u << 23); m += 4093) {
      for (uint32_t s = 0; s < 2; ++s) {
        uint32_t u = (s << 31) | (e << 23) | m;
        if (e == 0xFF && m) u |= 0x400000u;  // probe quiet NaNs only
        if (rsq(as_f32(u)) != predict(u)) return -1;
      }
    }
  return 0;
}

// setModelViewProjection (Rasterizer.cpp:76-105): raw rows are the transpose
// of the row-vector input; the prebaked rows fold the viewport (x, y mapped
// to [0, W-8] x [0, H-8]) and the depth mapping ([-1,1] -> [bias, 0]); both
// are stored as the transpose of those rows.
void prebake_mvp(uint32_t width, uint32_t height, const float* matrix, float* prebaked, float* raw) {
  float row[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) row[r][c] = matrix[4 * c + r];
  std::memcpy(raw, row, sizeof row);
  const float bias = 2.5237386e-29f;
  const float kx = float(width) * 0.5f - 4.0f;
  const float ky = float(height) * 0.5f - 4.0f;
  const float kz = 0.5f * bias;
  for (int c = 0; c < 4; ++c) {
    const float x = (row[0][c] + row[3][c]) * kx;
    const float y = (row[1][c] + row[3][c]) * ky;
    const float z = (row[3][c] - row[2][c]) * kz;
    prebaked[4 * c + 0] = x;
    prebaked[4 * c + 1] = y;
    prebaked[4 * c + 2] = z;
    prebaked[4 * c + 3] = row[3][c];
  }
}

// Occluder::bake (Occluder.cpp:7-181) and the VectorMath.h helpers it uses.
namespace {

// cross / normal (VectorMath.h:6-18): component-wise mul and sub, no fusion.
__m128 cross3(__m128 a, __m128 b) {
  __m128 a1 = _mm_shuffle_ps(a, a, _MM_SHUFFLE(3, 0, 2, 1));
  __m128 b1 = _mm_shuffle_ps(b, b, _MM_SHUFFLE(3, 0, 2, 1));
  __m128 c = _mm_sub_ps(_mm_mul_ps(a, b1), _mm_mul_ps(a1, b));
  return _mm_shuffle_ps(c, c, _MM_SHUFFLE(3, 0, 2, 1));
}
__m128 tri_normal(__m128 p0, __m128 p1, __m128 p2) { return cross3(_mm_sub_ps(p1, p0), _mm_sub_ps(p2, p0)); }
// normalize (VectorMath.h:20-23): scale by the approximate RSQRTPS of |v|^2.
__m128 approx_normalize(__m128 v) { return _mm_mul_ps(v, _mm_rsqrt_ps(_mm_dp_ps(v, v, 0x7F))); }

}  // namespace

int bake(const float* vertices, uint32_t n, const float* refMin, const float* refMax, uint32_t* packets,
         float* boundsMin, float* boundsMax, float* center) {
  if (n == 0 || (n % 32) != 0) return -1;
  const uint32_t quads = n / 4;
  const __m128* vtx = nullptr;
  std::vector<__m128> verts(n);
  for (uint32_t i = 0; i < n; ++i) verts[i] = _mm_loadu_ps(vertices + 4 * i);
  vtx = verts.data();

  // Per-quad normal: sum of the two triangle normals, approximately
  // normalized (Occluder.cpp:11-21).
  std::vector<__m128> nrm(quads);
  for (uint32_t q = 0; q < quads; ++q) {
    const __m128* p = vtx + 4 * q;
    nrm[q] = approx_normalize(_mm_add_ps(tri_normal(p[0], p[1], p[2]), tri_normal(p[0], p[2], p[3])));
  }

  // K-means over six axis-aligned seeds, at most 10 rounds, ties go to the
  // later centroid (Occluder.cpp:23-78).
  __m128 seeds[6] = {_mm_setr_ps(1, 0, 0, 0),  _mm_setr_ps(0, 1, 0, 0),  _mm_setr_ps(0, 0, 1, 0),
                     _mm_setr_ps(0, -1, 0, 0), _mm_setr_ps(0, 0, -1, 0), _mm_setr_ps(-1, 0, 0, 0)};
  std::vector<uint32_t> cluster(quads, 0);
  bool changed = true;
  for (int round = 0; round < 10 && changed; ++round) {
    changed = false;
    for (uint32_t q = 0; q < quads; ++q) {
      __m128 best = _mm_set1_ps(-INFINITY);
      uint32_t pick = 0;
      for (uint32_t k = 0; k < 6; ++k) {
        __m128 d = _mm_dp_ps(seeds[k], nrm[q], 0x7F);
        if (_mm_comige_ss(d, best)) {
          best = d;
          pick = k;
        }
      }
      if (cluster[q] != pick) {
        cluster[q] = pick;
        changed = true;
      }
    }
    for (auto& s : seeds) s = _mm_setzero_ps();
    for (uint32_t q = 0; q < quads; ++q) seeds[cluster[q]] = _mm_add_ps(seeds[cluster[q]], nrm[q]);
    for (auto& s : seeds) s = approx_normalize(s);
  }

  // Stable regroup by cluster index (Occluder.cpp:80-93).
  std::vector<uint32_t> order;
  order.reserve(quads);
  for (uint32_t k = 0; k < 6; ++k)
    for (uint32_t q = 0; q < quads; ++q)
      if (cluster[q] == k) order.push_back(q);

  // 11/11/10-bit quantization against [refMin, refMax] and the 8-quad packet
  // layout: word (4b + k) * 8 + lane = vertex k of quad 8b + lane
  // (Occluder.cpp:95-157).
  const __m128 lo = _mm_loadu_ps(refMin);
  const __m128 inv = _mm_div_ps(_mm_set1_ps(1.0f), _mm_sub_ps(_mm_loadu_ps(refMax), lo));
  for (uint32_t b = 0; b < quads / 8; ++b) {
    for (uint32_t k = 0; k < 4; ++k) {
      for (uint32_t lane = 0; lane < 8; ++lane) {
        V4 t = from_m(_mm_mul_ps(_mm_sub_ps(vtx[4 * order[8 * b + lane] + k], lo), inv));
        const float sx = std::fma(t.v[0], 2047.0f, 0.5f);
        const float sy = std::fma(t.v[1], 2047.0f, 0.5f);
        const float sz = std::fma(t.v[2], 1023.0f, 0.5f);
        const uint32_t qx = uint32_t(_mm_cvttss_si32(_mm_set_ss(sx)) - 1024);
        const uint32_t qy = uint32_t(_mm_cvttss_si32(_mm_set_ss(sy)));
        const uint32_t qz = uint32_t(_mm_cvttss_si32(_mm_set_ss(sz)));
        packets[(4 * b + k) * 8 + lane] = (qx << 21) | (qy << 10) | qz;
      }
    }
  }

  // Bounds over the input vertices with w forced to 1, and their centre
  // (Occluder.cpp:159-178).
  __m128 mn = _mm_set1_ps(INFINITY), mx = _mm_set1_ps(-INFINITY);
  for (uint32_t i = 0; i < n; ++i) {
    mn = _mm_min_ps(vtx[i], mn);
    mx = _mm_max_ps(vtx[i], mx);
  }
  mn = _mm_blend_ps(mn, _mm_set1_ps(1.0f), 0x8);
  mx = _mm_blend_ps(mx, _mm_set1_ps(1.0f), 0x8);
  _mm_storeu_ps(boundsMin, mn);
  _mm_storeu_ps(boundsMax, mx);
  _mm_storeu_ps(center, _mm_mul_ps(_mm_add_ps(mx, mn), _mm_set1_ps(0.5f)));
  return 0;
}

}  // namespace rzh
