// Test program for the C++ drop-in headers: the reference demo's scene
// preparation (Main.cpp:54-128) and frame (Main.cpp:172-206, 229), headless,
// written against include/rasterizer_amd/{Rasterizer,Occluder,
// QuadDecomposition,SurfaceAreaHeuristic}.h exactly as Main.cpp calls them.
// Aabb is defined here with the reference's layout and member functions
// (VectorMath.h:25-54), as Main.cpp gets it from VectorMath.h.
//
// usage: main_sequence_check dir width height nFrames
//   dir/vertices.f32  n x 4 floats (the vertex buffer)
//   dir/indices.u32   triangle indices (the index buffer)
//   dir/frames.f32    nFrames x (16 matrix floats + 4 camera position floats)
// writes
//   dir/occluders.bin for each occluder: u32 packetCount, then 16 floats
//                     (boundsMin, boundsMax, refMin, refMax), then the packets
//   dir/frame<k>.bin  u32 order[nOccluders] (the sorted draw order as
//                     original occluder numbers), readBackDepth's
//                     width*height*4 bytes, then raw depth and hiZ
#include <immintrin.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Occluder.h"
#include "QuadDecomposition.h"
#include "Rasterizer.h"
#include "SurfaceAreaHeuristic.h"
#include "rasterizer_amd.h"

struct Aabb {
  Aabb() {
    m_min = _mm_set1_ps(+std::numeric_limits<float>::infinity());
    m_max = _mm_set1_ps(-std::numeric_limits<float>::infinity());
  }
  __m128 m_min, m_max;
  void include(__m128 point) {
    m_min = _mm_min_ps(m_min, point);
    m_max = _mm_max_ps(m_max, point);
  }
};

namespace {

template <typename T>
std::vector<T> read_all(const std::string& path) {
  std::vector<T> out;
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  T v;
  while (std::fread(&v, sizeof v, 1, f) == 1) out.push_back(v);
  std::fclose(f);
  return out;
}

void write_all(std::FILE* f, const void* p, size_t n) {
  if (n && std::fwrite(p, 1, n, f) != n) throw std::runtime_error("short write");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s dir width height nFrames\n", argv[0]);
    return 2;
  }
  try {
    const std::string dir = argv[1];
    const uint32_t W = uint32_t(std::atoi(argv[2])), H = uint32_t(std::atoi(argv[3]));
    const int nFrames = std::atoi(argv[4]);
    const std::vector<float> vflat = read_all<float>(dir + "/vertices.f32");
    std::vector<__m128> vertices(vflat.size() / 4);
    for (size_t i = 0; i < vertices.size(); ++i) vertices[i] = _mm_loadu_ps(&vflat[4 * i]);
    std::vector<uint32_t> indices = read_all<uint32_t>(dir + "/indices.u32");
    const std::vector<float> frames = read_all<float>(dir + "/frames.f32");

    // Main.cpp:86-128.
    indices = QuadDecomposition::decompose(indices, vertices);

    auto rasterizer = std::make_unique<Rasterizer>(W, H);

    while (indices.size() % 32 != 0) indices.push_back(indices[0]);

    std::vector<Aabb> quadAabbs;
    for (size_t quadIndex = 0; quadIndex < indices.size() / 4; ++quadIndex) {
      Aabb aabb;
      aabb.include(vertices[indices[4 * quadIndex + 0]]);
      aabb.include(vertices[indices[4 * quadIndex + 1]]);
      aabb.include(vertices[indices[4 * quadIndex + 2]]);
      aabb.include(vertices[indices[4 * quadIndex + 3]]);
      quadAabbs.push_back(aabb);
    }

    auto batchAssignment = SurfaceAreaHeuristic::generateBatches(quadAabbs, 512, 8);

    Aabb refAabb;
    for (auto v : vertices) refAabb.include(v);

    std::vector<std::unique_ptr<Occluder>> occluders;
    for (const auto& batch : batchAssignment) {
      std::vector<__m128> batchVertices;
      for (auto quadIndex : batch) {
        batchVertices.push_back(vertices[indices[quadIndex * 4 + 0]]);
        batchVertices.push_back(vertices[indices[quadIndex * 4 + 1]]);
        batchVertices.push_back(vertices[indices[quadIndex * 4 + 2]]);
        batchVertices.push_back(vertices[indices[quadIndex * 4 + 3]]);
      }
      occluders.push_back(Occluder::bake(batchVertices, refAabb.m_min, refAabb.m_max));
    }

    {
      std::FILE* f = std::fopen((dir + "/occluders.bin").c_str(), "wb");
      if (!f) throw std::runtime_error("cannot write occluders.bin");
      for (const auto& o : occluders) {
        float box[16];
        _mm_storeu_ps(box + 0, o->m_boundsMin);
        _mm_storeu_ps(box + 4, o->m_boundsMax);
        _mm_storeu_ps(box + 8, o->m_refMin);
        _mm_storeu_ps(box + 12, o->m_refMax);
        write_all(f, &o->m_packetCount, 4);
        write_all(f, box, sizeof box);
        write_all(f, o->m_vertexData, size_t(o->m_packetCount) * 32);
      }
      std::fclose(f);
    }
    // Each occluder's original number, carried through the sort.
    std::vector<std::pair<const Occluder*, uint32_t>> numbered;
    for (uint32_t i = 0; i < occluders.size(); ++i) numbered.emplace_back(occluders[i].get(), i);

    std::vector<char> rawData(size_t(W) * H * 4);
    for (int k = 0; k < nFrames; ++k) {
      float mvp[16];
      std::copy(frames.begin() + 20 * k, frames.begin() + 20 * k + 16, mvp);
      const __m128 cameraPosition = _mm_loadu_ps(&frames[20 * size_t(k) + 16]);

      // Main.cpp:180-206.
      rasterizer->clear();
      rasterizer->setModelViewProjection(mvp);

      std::sort(begin(numbered), end(numbered), [&](const auto& o1, const auto& o2) {
        __m128 dist1 = _mm_sub_ps(o1.first->m_center, cameraPosition);
        __m128 dist2 = _mm_sub_ps(o2.first->m_center, cameraPosition);
        return _mm_comilt_ss(_mm_dp_ps(dist1, dist1, 0x7f), _mm_dp_ps(dist2, dist2, 0x7f));
      });

      for (const auto& entry : numbered) {
        const Occluder* occluder = entry.first;
        bool needsClipping;
        if (rasterizer->queryVisibility(occluder->m_boundsMin, occluder->m_boundsMax, needsClipping)) {
          if (needsClipping) {
            rasterizer->rasterize<true>(*occluder);
          } else {
            rasterizer->rasterize<false>(*occluder);
          }
        }
      }

      // Main.cpp:229.
      rasterizer->readBackDepth(&*rawData.begin());

      const size_t nb = size_t(W / 8) * (H / 8);
      std::vector<uint16_t> depth(nb * 64), hiz(nb + 8);
      if (rz_readback_raw(rasterizer->context(), depth.data(), hiz.data()) != RZ_OK)
        throw std::runtime_error("rz_readback_raw");
      std::FILE* f = std::fopen((dir + "/frame" + std::to_string(k) + ".bin").c_str(), "wb");
      if (!f) throw std::runtime_error("cannot write frame");
      for (const auto& entry : numbered) write_all(f, &entry.second, 4);
      write_all(f, rawData.data(), rawData.size());
      write_all(f, depth.data(), depth.size() * 2);
      write_all(f, hiz.data(), hiz.size() * 2);
      std::fclose(f);
    }
    std::printf("occluders %zu frames %d\n", occluders.size(), nFrames);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "main_sequence_check: %s\n", e.what());
    return 1;
  }
  return 0;
}
