// Drop-in replacement for rawrunprotected/rasterizer's Rasterizer.h
// (SoftwareRasterizer/Rasterizer.h:10-61): same class name and public
// signatures, backed by the MI355X kernels through the C ABI in
// rasterizer_amd.h.  Failures of the device surface as std::runtime_error
// (the reference has no error channel).
#pragma once

#include <immintrin.h>

#include <cstdint>
#include <memory>
#include <vector>

struct Occluder;
struct rz_context;

class Rasterizer {
 public:
  // Rasterizer.h:13
  Rasterizer(uint32_t width, uint32_t height);
  ~Rasterizer();
  Rasterizer(const Rasterizer&) = delete;
  Rasterizer& operator=(const Rasterizer&) = delete;

  // Rasterizer.h:15
  void setModelViewProjection(const float* matrix);

  // Rasterizer.h:17
  void clear();

  // Rasterizer.h:19-20 (instantiated for true and false)
  template <bool possiblyNearClipped>
  void rasterize(const Occluder& occluder);

  // Rasterizer.h:22
  bool queryVisibility(__m128 boundsMin, __m128 boundsMax, bool& needsClipping);

  // Rasterizer.h:24
  bool query2D(uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY, uint32_t maxZ) const;

  // Rasterizer.h:26
  void readBackDepth(void* target) const;

  // Not in the reference: Main.cpp:192-206's query-gated loop over
  // occluders in draw order, run on the device without a host round trip per
  // occluder.  Same buffer as calling queryVisibility/rasterize one by one.
  void drawOccluders(const std::vector<const Occluder*>& occluders);

  rz_context* context() const { return m_ctx; }

 private:
  rz_context* m_ctx = nullptr;
};
