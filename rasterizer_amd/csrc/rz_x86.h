// Device-side restatements of the x86 SIMD semantics the reference's
// arithmetic depends on.  gfx950 instructions differ from SSE/AVX in exactly
// these places, so each is written out:
//  - RCPPS is an approximation whose bits are CPU-specific: emulated from a
//    4096-entry table of the host CPU's RCPPS(1.m) (built and checked by the
//    host library), with exact exponent scaling and the special cases RCPPS
//    has (inf for 0/denormals, 0 for inf and for underflowing results).
//  - CVTTPS2DQ returns 0x80000000 for NaN and out-of-range inputs, while
//    v_cvt_i32_f32 saturates (and gives 0 for NaN).
//  - MINPS/MAXPS return the second operand when either input is NaN (and on
//    equal inputs, which matters for +-0), unlike v_min_f32 / v_max_f32.
//  - DPPS sums its products as (p0 + p1) + (p2 + p3) with each product and sum
//    rounded to float.
//  - PACKUSDW saturates the signed 32-bit input to [0, 65535]; PAVGW rounds up.
//  - NaNs: an invalid operation (0/0, inf/inf, 0 * inf, inf - inf) gives the
//    x86 default NaN 0xFFC00000 (negative) where gfx950 gives 0x7FC00000, and
//    a NaN operand propagates quieted with its sign (FNMADD does not negate
//    it).  Packed depth reads the sign (a negative NaN packs to 0, a positive
//    one to 65535), so the divide and FMAs whose results become depth are
//    written out (div_x86, fmadd_x86, fnmadd_x86).  NaN operands are taken in
//    argument order.
// The file is compiled with -ffp-contract=off: a * b + c is two roundings,
// and fmaf() is used exactly where the reference calls an FMA intrinsic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rzd {

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ float xor_sign(float f, uint32_t s) { return bitsf(fbits(f) ^ s); }

__device__ __forceinline__ float rcp_x86(float x, const uint32_t* __restrict__ tab) {
  const uint32_t u = fbits(x);
  const uint32_t s = u & 0x80000000u;
  const uint32_t e = (u >> 23) & 0xFFu;
  const uint32_t m = u & 0x7FFFFFu;
  if (e == 0xFFu) return bitsf(m ? (u | 0x400000u) : s);
  if (e == 0u) return bitsf(s | 0x7F800000u);
  const uint32_t t = tab[m >> 11];
  const int re = int((t >> 23) & 0xFFu) + 127 - int(e);
  if (re <= 0) return bitsf(s);
  return bitsf(s | (uint32_t(re) << 23) | (t & 0x7FFFFFu));
}

// RSQRTPS from the host's table (8192 entries: exponent parity p, the top 12
// mantissa bits; see rzh::build_rsqrt_table).
__device__ __forceinline__ float rsqrt_x86(float x, const uint32_t* __restrict__ tab) {
  const uint32_t u = fbits(x);
  const uint32_t s = u & 0x80000000u;
  const uint32_t e = (u >> 23) & 0xFFu;
  const uint32_t m = u & 0x7FFFFFu;
  if (e == 0xFFu) return bitsf(m ? (u | 0x400000u) : (s ? 0xFFC00000u : 0u));
  if (e == 0u) return bitsf(s | 0x7F800000u);
  if (s) return bitsf(0xFFC00000u);
  const int E = int(e) - 127, p = E & 1, k = (E - p) / 2;
  const uint32_t t = tab[uint32_t(p) * 4096u + (m >> 11)];
  return bitsf((uint32_t(int((t >> 23) & 0xFFu) - k) << 23) | (t & 0x7FFFFFu));
}

// The xyz lanes of an SSE vector (the w lane never reaches a result below:
// DPPS with mask 0x7F leaves it out).
struct V3 {
  float x, y, z;
};

__device__ __forceinline__ V3 sub3(float4 a, float4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

// cross3 of the reference's shuffle form: (a * b.yzx - a.yzx * b).yzx.
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// _mm_dp_ps(a, b, 0x7F): (x + y) + (z + the masked-out lane's +0).
__device__ __forceinline__ float dp3(V3 a, V3 b) {
  const float p0 = a.x * b.x, p1 = a.y * b.y, p2 = a.z * b.z;
  return (p0 + p1) + (p2 + 0.0f);
}

// normalize (VectorMath.h:20-23): v * RSQRTPS(v . v).
__device__ __forceinline__ V3 approx_normalize(V3 v, const uint32_t* __restrict__ rsq) {
  const float r = rsqrt_x86(dp3(v, v), rsq);
  return {v.x * r, v.y * r, v.z * r};
}

__device__ __forceinline__ int32_t cvtt_x86(float x) {
  return (__builtin_fabsf(x) < 2147483648.0f) ? int32_t(x) : int32_t(0x80000000u);
}

__device__ __forceinline__ float min_x86(float a, float b) { return (a < b) ? a : b; }
__device__ __forceinline__ float max_x86(float a, float b) { return (a > b) ? a : b; }

__device__ __forceinline__ float dp4_x86(float a0, float a1, float a2, float a3, float b0, float b1, float b2,
                                         float b3) {
  const float p0 = a0 * b0, p1 = a1 * b1, p2 = a2 * b2, p3 = a3 * b3;
  return (p0 + p1) + (p2 + p3);
}

constexpr uint32_t kX86DefaultNaN = 0xFFC00000u;
__device__ __forceinline__ bool nan_x86(float x) { return (fbits(x) & 0x7FFFFFFFu) > 0x7F800000u; }
__device__ __forceinline__ float quiet_x86(float x) { return bitsf(fbits(x) | 0x400000u); }

// DIVSS.
__device__ __forceinline__ float div_x86(float a, float b) {
  if (nan_x86(a)) return quiet_x86(a);
  if (nan_x86(b)) return quiet_x86(b);
  const float r = a / b;
  return nan_x86(r) ? bitsf(kX86DefaultNaN) : r;
}

// VFMADD: a * b + c, one rounding.
__device__ __forceinline__ float fmadd_x86(float a, float b, float c) {
  if (nan_x86(a)) return quiet_x86(a);
  if (nan_x86(b)) return quiet_x86(b);
  if (nan_x86(c)) return quiet_x86(c);
  const float r = fmaf(a, b, c);
  return nan_x86(r) ? bitsf(kX86DefaultNaN) : r;
}

// VFNMADD: -(a * b) + c, one rounding.
__device__ __forceinline__ float fnmadd_x86(float a, float b, float c) {
  if (nan_x86(a)) return quiet_x86(a);
  if (nan_x86(b)) return quiet_x86(b);
  if (nan_x86(c)) return quiet_x86(c);
  const float r = fmaf(-a, b, c);
  return nan_x86(r) ? bitsf(kX86DefaultNaN) : r;
}

__device__ __forceinline__ uint32_t pack_depth(float d) {
  const int32_t v = int32_t(fbits(d)) >> 12;
  return v < 0 ? 0u : (v > 65535 ? 65535u : uint32_t(v));
}

__device__ __forceinline__ uint32_t avg_u16(uint32_t a, uint32_t b) { return (a + b + 1u) >> 1; }

}  // namespace rzd
