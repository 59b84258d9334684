// Reference-compatible Rasterizer / Occluder classes over the C ABI
// (include/rasterizer_amd/Rasterizer.h, Occluder.h).
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "rasterizer_amd.h"
#include "rasterizer_amd/Occluder.h"
#include "rasterizer_amd/QuadDecomposition.h"
#include "rasterizer_amd/Rasterizer.h"
#include "rasterizer_amd/SurfaceAreaHeuristic.h"

namespace {

void check(int status, const char* what) {
  if (status != RZ_OK) throw std::runtime_error(std::string(what) + ": " + rz_status_string(status));
}

void store4(__m128 v, float* out) { _mm_storeu_ps(out, v); }

}  // namespace

// ----------------------------------------------------------------- Occluder

std::unique_ptr<Occluder> Occluder::bake(const std::vector<__m128>& vertices, __m128 refMin, __m128 refMax) {
  const uint32_t n = uint32_t(vertices.size());
  if (n == 0 || n % 32 != 0) throw std::invalid_argument("Occluder::bake: vertex count must be a multiple of 32");
  std::vector<float> flat(size_t(n) * 4);
  for (uint32_t i = 0; i < n; ++i) store4(vertices[i], &flat[4 * size_t(i)]);
  float lo[4], hi[4], bmin[4], bmax[4], center[4];
  store4(refMin, lo);
  store4(refMax, hi);
  auto occ = std::make_unique<Occluder>();
  occ->m_vertexData = static_cast<__m256i*>(std::aligned_alloc(32, size_t(n) * 4));
  if (!occ->m_vertexData) throw std::bad_alloc();
  check(rz_bake(flat.data(), n, lo, hi, reinterpret_cast<uint32_t*>(occ->m_vertexData), bmin, bmax, center),
        "Occluder::bake");
  occ->m_packetCount = n / 8;
  occ->m_refMin = refMin;
  occ->m_refMax = refMax;
  occ->m_boundsMin = _mm_loadu_ps(bmin);
  occ->m_boundsMax = _mm_loadu_ps(bmax);
  occ->m_center = _mm_loadu_ps(center);
  return occ;
}

Occluder::~Occluder() {
  if (m_device) rz_occluder_destroy(m_device);
  std::free(m_vertexData);
}

rz_occluder* Occluder::device(rz_context* ctx) const {
  if (!m_device) {
    float lo[4], hi[4], bmin[4], bmax[4];
    store4(m_refMin, lo);
    store4(m_refMax, hi);
    store4(m_boundsMin, bmin);
    store4(m_boundsMax, bmax);
    check(rz_occluder_create(ctx, reinterpret_cast<const uint32_t*>(m_vertexData), m_packetCount, lo, hi, bmin, bmax,
                             &m_device),
          "Occluder upload");
  }
  return m_device;
}

// --------------------------------------------------------------- Rasterizer

Rasterizer::Rasterizer(uint32_t width, uint32_t height) { check(rz_create(width, height, &m_ctx), "Rasterizer"); }

Rasterizer::~Rasterizer() {
  if (m_ctx) rz_destroy(m_ctx);
}

void Rasterizer::setModelViewProjection(const float* matrix) {
  check(rz_set_mvp(m_ctx, matrix), "setModelViewProjection");
}

void Rasterizer::clear() { check(rz_clear(m_ctx), "clear"); }

template <bool possiblyNearClipped>
void Rasterizer::rasterize(const Occluder& occluder) {
  check(rz_rasterize(m_ctx, occluder.device(m_ctx), possiblyNearClipped ? 1 : 0), "rasterize");
}

template void Rasterizer::rasterize<true>(const Occluder& occluder);
template void Rasterizer::rasterize<false>(const Occluder& occluder);

bool Rasterizer::queryVisibility(__m128 boundsMin, __m128 boundsMax, bool& needsClipping) {
  float lo[4], hi[4];
  store4(boundsMin, lo);
  store4(boundsMax, hi);
  // needsClipping is only ever written (Rasterizer.cpp:164-167 leaves it
  // unwritten for a frustum-culled box), so the caller's value is not read:
  // rz_query replaces the sentinel only when it writes the flag.
  int visible = 0, clip = -1;
  check(rz_query(m_ctx, lo, hi, &visible, &clip), "queryVisibility");
  if (clip != -1) needsClipping = clip != 0;
  return visible != 0;
}

bool Rasterizer::query2D(uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY, uint32_t maxZ) const {
  int visible = 0;
  check(rz_query2d(m_ctx, minX, maxX, minY, maxY, maxZ, &visible), "query2D");
  return visible != 0;
}

void Rasterizer::readBackDepth(void* target) const { check(rz_readback_depth(m_ctx, target), "readBackDepth"); }

void Rasterizer::drawOccluders(const std::vector<const Occluder*>& occluders) {
  std::vector<const rz_occluder*> handles;
  handles.reserve(occluders.size());
  for (const Occluder* o : occluders) handles.push_back(o->device(m_ctx));
  check(rz_draw_frame(m_ctx, uint32_t(handles.size()), handles.data(), nullptr), "drawOccluders");
}

// ------------------------------------------------------- QuadDecomposition

std::vector<uint32_t> QuadDecomposition::decompose(const std::vector<uint32_t>& indices,
                                                   const std::vector<__m128>& vertices) {
  std::vector<float> flat(vertices.size() * 4);
  for (size_t i = 0; i < vertices.size(); ++i) store4(vertices[i], &flat[4 * i]);
  std::vector<uint32_t> out(4 * (indices.size() / 3));
  size_t n = 0;
  check(rz_quad_decompose(indices.data(), indices.size(), flat.data(), vertices.size(), out.data(), &n),
        "QuadDecomposition::decompose");
  out.resize(n);
  return out;
}

// SurfaceAreaHeuristic::generateBatches (SurfaceAreaHeuristic.cpp:96-104) on
// the GPU: a context of its own for the stream, the boxes uploaded once.
std::vector<std::vector<uint32_t>> rz_dropin::sah_generate_batches(const float* boxes, size_t n, uint32_t targetSize,
                                                                   uint32_t splitGranularity) {
  if (n > 0xFFFFFFFFull) throw std::runtime_error("SurfaceAreaHeuristic::generateBatches: too many boxes");
  std::vector<std::vector<uint32_t>> out;
  rz_context* ctx = nullptr;
  check(rz_create(8, 8, &ctx), "rz_create");
  void* d_boxes = nullptr;
  void* d_indices = nullptr;
  int status = RZ_OK;
  std::vector<uint32_t> offsets(n + 2, 0u), indices(n);
  uint32_t nBatches = 0;
  if (n) {
    status = rz_device_alloc(ctx, n * 32, boxes, &d_boxes);
    if (status == RZ_OK) status = rz_device_alloc(ctx, n * 4, nullptr, &d_indices);
  }
  if (status == RZ_OK)
    status = rz_sah_batches(ctx, uint32_t(n), static_cast<const float*>(d_boxes), targetSize, splitGranularity,
                            static_cast<uint32_t*>(d_indices), offsets.data(), &nBatches);
  if (status == RZ_OK && n) status = rz_device_read(ctx, d_indices, n * 4, indices.data());
  if (d_boxes) rz_device_free(d_boxes);
  if (d_indices) rz_device_free(d_indices);
  rz_destroy(ctx);
  check(status, "SurfaceAreaHeuristic::generateBatches");
  out.reserve(nBatches);
  for (uint32_t b = 0; b < nBatches; ++b)
    out.emplace_back(indices.begin() + offsets[b], indices.begin() + offsets[b + 1]);
  return out;
}
