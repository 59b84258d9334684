// TEST INFRASTRUCTURE ONLY: a C-ABI caller of the reference's own
// SurfaceAreaHeuristic::generateBatches (SurfaceAreaHeuristic.cpp:96-104),
// compiled from the reference's unmodified source into oracle/_ref/ by
// oracle/Makefile's `ref` target.  Used by tests/test_sah.py to check the
// restatement against the reference itself; nothing in the product loads it.
#include <cstdint>
#include <limits>
#include <vector>

#include "SurfaceAreaHeuristic.h"
#include "VectorMath.h"

// boxes: n Aabbs (min xyzw, max xyzw).  Writes the batches' indices to
// indices (n words) and the boundaries to offsets (nBatches + 1 words; room
// for n + 1).  Returns -1 where the reference's behaviour is undefined: an
// empty set (it takes &indices[0]) or one with no split position.
extern "C" int sah_ref_batches(uint32_t n, const float* boxes, uint32_t target, uint32_t gran, uint32_t* indices,
                               uint32_t* offsets, uint32_t* nBatches) {
  if (n == 0 || gran == 0 || n <= 2 * gran) return -1;
  std::vector<Aabb> aabbs(n);
  for (uint32_t i = 0; i < n; ++i) {
    aabbs[i].m_min = _mm_loadu_ps(boxes + 8 * size_t(i));
    aabbs[i].m_max = _mm_loadu_ps(boxes + 8 * size_t(i) + 4);
  }
  const std::vector<std::vector<uint32_t>> batches = SurfaceAreaHeuristic::generateBatches(aabbs, target, gran);
  uint32_t at = 0;
  offsets[0] = 0;
  for (size_t b = 0; b < batches.size(); ++b) {
    for (uint32_t q : batches[b]) indices[at++] = q;
    offsets[b + 1] = at;
  }
  *nBatches = uint32_t(batches.size());
  return 0;
}
