// Test program for the SurfaceAreaHeuristic drop-in header.  Reads n boxes
// (8 floats each: min xyzw, max xyzw) from argv[1], calls
// SurfaceAreaHeuristic::generateBatches(boxes, argv[3], argv[4]) as
// Main.cpp:107 does, and writes u32 words to argv[2]: the batch count, then
// each batch's size followed by its indices.  Aabb is defined here with the
// reference's layout (two __m128), after the drop-in header, as Main.cpp
// includes VectorMath.h after SurfaceAreaHeuristic.h.
#include <immintrin.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

// A forward declaration, as code written against the reference's header may
// carry: the drop-in's SurfaceAreaHeuristic is a class, so it must agree.
class SurfaceAreaHeuristic;

#include "SurfaceAreaHeuristic.h"

struct Aabb {
  __m128 m_min, m_max;
};

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s boxes.f32 out.u32 targetSize splitGranularity\n", argv[0]);
    return 2;
  }
  std::FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<float> raw;
  float buf[8];
  while (std::fread(buf, sizeof(float), 8, f) == 8) raw.insert(raw.end(), buf, buf + 8);
  std::fclose(f);
  std::vector<Aabb> boxes(raw.size() / 8);
  for (size_t i = 0; i < boxes.size(); ++i) {
    boxes[i].m_min = _mm_loadu_ps(&raw[8 * i]);
    boxes[i].m_max = _mm_loadu_ps(&raw[8 * i + 4]);
  }
  std::vector<std::vector<uint32_t>> batches;
  try {
    batches = SurfaceAreaHeuristic::generateBatches(boxes, uint32_t(std::atoi(argv[3])), uint32_t(std::atoi(argv[4])));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  std::FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  const uint32_t nb = uint32_t(batches.size());
  std::fwrite(&nb, 4, 1, o);
  for (const auto& b : batches) {
    const uint32_t sz = uint32_t(b.size());
    std::fwrite(&sz, 4, 1, o);
    std::fwrite(b.data(), 4, b.size(), o);
  }
  std::fclose(o);
  return 0;
}
