// Drop-in replacement for rawrunprotected/rasterizer's Occluder.h
// (SoftwareRasterizer/Occluder.h:7-21): same static bake and public fields.
// The baked packets also get a device copy, created on first use.
#pragma once

#include <immintrin.h>

#include <cstdint>
#include <memory>
#include <vector>

struct rz_occluder;
struct rz_context;

struct Occluder {
  // Occluder.h:9 / Occluder.cpp:7-181
  static std::unique_ptr<Occluder> bake(const std::vector<__m128>& vertices, __m128 refMin, __m128 refMax);

  Occluder() = default;
  ~Occluder();
  Occluder(const Occluder&) = delete;
  Occluder& operator=(const Occluder&) = delete;

  __m128 m_center;

  __m128 m_refMin;
  __m128 m_refMax;

  __m128 m_boundsMin;
  __m128 m_boundsMax;

  __m256i* m_vertexData = nullptr;
  uint32_t m_packetCount = 0;

  // Device copy of m_vertexData (created by the first Rasterizer that draws it).
  rz_occluder* device(rz_context* ctx) const;

 private:
  mutable rz_occluder* m_device = nullptr;
};
