// TEST INFRASTRUCTURE ONLY: a C-ABI caller of the reference's own
// QuadDecomposition::decompose (QuadDecomposition.cpp:346-445), compiled from
// the reference's unmodified source into oracle/_ref/ by oracle/Makefile's
// `ref` target.  Used by tests/test_quad.py to check the library's
// rz_quad_decompose against the reference itself, and by tests/cpp's check
// of Main.cpp's sequence; nothing in the product loads it.
#include <immintrin.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "QuadDecomposition.h"

// indices: nIndices words (triangles); vertices: nVertices x 4 floats.
// Writes the quads' indices to out (room for 4 * (nIndices / 3) words) and
// their count to *nOut.  Returns -1 for an index outside the vertex array
// (undefined in the reference).
extern "C" int quad_ref_decompose(const uint32_t* indices, size_t nIndices, const float* vertices, size_t nVertices,
                                  uint32_t* out, size_t* nOut) {
  for (size_t i = 0; i < nIndices; ++i)
    if (indices[i] >= nVertices) return -1;
  std::vector<uint32_t> idx(indices, indices + nIndices);
  std::vector<__m128> verts(nVertices);
  for (size_t v = 0; v < nVertices; ++v) verts[v] = _mm_loadu_ps(vertices + 4 * v);
  const std::vector<uint32_t> quads = QuadDecomposition::decompose(idx, verts);
  if (!quads.empty()) std::memcpy(out, quads.data(), quads.size() * sizeof(uint32_t));
  *nOut = quads.size();
  return 0;
}
