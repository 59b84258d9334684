// Drop-in replacement for rawrunprotected/rasterizer's QuadDecomposition.h
// (SoftwareRasterizer/QuadDecomposition.h:5-11): the same static decompose,
// computed by the library's rz_quad_decompose with the reference's exact
// quads (QuadDecomposition.cpp:311-445).
#pragma once

#include <immintrin.h>

#include <cstdint>
#include <vector>

class QuadDecomposition {
 public:
  // Main.cpp:86.  Throws std::runtime_error on a non-zero status (an index
  // outside the vertex array, which the reference leaves undefined).
  static std::vector<uint32_t> decompose(const std::vector<uint32_t>& indices, const std::vector<__m128>& vertices);
};
