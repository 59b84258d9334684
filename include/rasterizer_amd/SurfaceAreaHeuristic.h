// Drop-in replacement for rawrunprotected/rasterizer's SurfaceAreaHeuristic.h
// (SoftwareRasterizer/SurfaceAreaHeuristic.h:5-11): the same static
// generateBatches, computed on the GPU by rz_sah_batches with the reference's
// exact batches (SurfaceAreaHeuristic.cpp:10-104).  As in the reference, Aabb
// (two __m128: m_min, m_max) only needs to be complete where generateBatches
// is called.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

struct Aabb;

namespace rz_dropin {

// Batches of n boxes laid out as std::vector<Aabb> stores them (8 floats
// each); throws std::runtime_error on a non-zero status.
std::vector<std::vector<uint32_t>> sah_generate_batches(const float* boxes, size_t n, uint32_t targetSize,
                                                        uint32_t splitGranularity);

}  // namespace rz_dropin

// SurfaceAreaHeuristic::generateBatches(aabbs, targetSize, splitGranularity)
// as the reference calls it (Main.cpp:107).  A plain class, so that it can be
// forward-declared as the reference's header allows; generateBatches is a
// member template on the box type, compiled only where it is called, where
// the caller's std::vector<Aabb> has made Aabb complete.
class SurfaceAreaHeuristic {
 public:
  template <typename AabbT = Aabb>
  static std::vector<std::vector<uint32_t>> generateBatches(const std::vector<AabbT>& aabbs, uint32_t targetSize,
                                                            uint32_t splitGranularity) {
    static_assert(sizeof(AabbT) == 8 * sizeof(float), "Aabb is two __m128 (m_min, m_max)");
    return rz_dropin::sah_generate_batches(reinterpret_cast<const float*>(aabbs.data()), aabbs.size(), targetSize,
                                           splitGranularity);
  }
};
