// Host-side helpers of the rasterizer library (x86-64 code, see rz_host.cpp).
#pragma once

#include <cstdint>

namespace rzh {

// 4096-entry coverage table of the reference constructor; false if the
// table's own invariant fails.
bool build_coverage_table(uint64_t* table);

// 4096-entry RCPPS(1.m) table of this CPU; 0 on success, -1 if this CPU's
// RCPPS does not have the structure the device emulation relies on.
int build_rcp_table(uint32_t* table);
// 8192-entry RSQRTPS table of this CPU: entry p * 4096 + (m >> 11) is
// RSQRTPS(2^p * 1.m); 0 on success, -1 if this CPU's RSQRTPS does not have
// that structure (a result fixed by the exponent's parity and the top 12
// mantissa bits, the exponent following, zeros and denormals to infinities,
// negatives to the default NaN).
int build_rsqrt_table(uint32_t* table);

// setModelViewProjection's stored state: prebaked columns and raw rows.
void prebake_mvp(uint32_t width, uint32_t height, const float* matrix, float* prebaked, float* raw);

// Occluder::bake; 0 on success, -1 if n is not a positive multiple of 32.
int bake(const float* vertices, uint32_t n, const float* refMin, const float* refMax, uint32_t* packets,
         float* boundsMin, float* boundsMax, float* center);

}  // namespace rzh
