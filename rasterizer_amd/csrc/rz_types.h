// Layouts shared by the host library and the gfx950 kernels.
#pragma once

#include <stdint.h>

namespace rzd {

// What a draw item asks for.
enum ItemKind : uint32_t {
  kItemQuery = 0,        // Main.cpp:195-204: queryVisibility, then rasterize<needsClipping> if visible
  kItemForceNoClip = 1,  // rasterize<false>
  kItemForceClip = 2,    // rasterize<true>
  kItemClassify = 3,     // depth-independent part of queryVisibility only, then rasterize<needsClipping>
};

// What the prep kernel decided for an item.
enum ItemStatus : uint32_t {
  kStatusSkip = 0,    // not drawn (frustum-culled or empty screen rectangle)
  kStatusNoClip = 1,  // drawn with rasterize<false>
  kStatusClip = 2,    // drawn with rasterize<true> (box near the camera plane)
  kStatusTest = 3,    // drawn with rasterize<false> if query2D over qRect says visible
};

// One draw item, written by the host per frame (88 bytes).
struct ItemDesc {
  const uint32_t* packets;  // device copy of Occluder::m_vertexData
  uint32_t quadBase;        // index of the item's first quad in the frame's quad arrays
  uint32_t quadCount;       // 8 per batch of 4 packets
  uint32_t kind;            // ItemKind
  uint32_t view;            // index into the view table of a multi-view launch (0 otherwise)
  float refMin[4], refMax[4];
  float boundsMin[4], boundsMax[4];
};

// Per-item state computed on the device (128 bytes).
struct ItemState {
  float T[4][4];   // rasterize prologue: (X, Y, Z, W) from (Xf, Yf, Zf, 1) (Rasterizer.cpp:616-642)
  float c0, c1;    // z = c0 + c1 / w (Rasterizer.cpp:645-655)
  uint32_t status; // ItemStatus
  uint32_t qMinX, qMaxX, qMinY, qMaxY, qMaxZ;  // query2D arguments (pixels, inclusive)
  uint32_t expected;  // number of tiles that overlap the query rectangle
  uint32_t fp[4];     // footprint of the valid quads in blocks: minX, maxX, minY, maxY (inclusive)
  uint32_t pad;
};

// Setup record of one quad: everything the block loop of
// Rasterizer.cpp:1095-1292 reads for one primitive (80 bytes).
struct QuadRec {
  float nx[4], ny[4];  // normalized, quantization-scaled edge normals
  float off[4];        // edge offsets at the rectangle origin (offset-bucket units)
  float dp0, dDx, dDy; // depth plane at the rectangle origin, per-block gradients
  uint32_t fb;         // first block index, with the reference's 16-bit multiply
  uint32_t range;      // rangeX | rangeY << 16 (blocks)
  uint32_t modeZ;      // primitive mode | packed max depth << 16
  uint32_t slope01;    // slope lookups of edges 0 and 1 (already << 6)
  uint32_t slope23;
};

// MVP state as setModelViewProjection stores it (passed to kernels by value).
struct Mvp {
  float prebaked[16];  // m_modelViewProjection
  float raw[16];       // m_modelViewProjectionRaw
};

// One view of a multi-view launch (256 bytes): its buffer, its matrix, its
// slice of the launch's items and queries, and its range of (view, tile)
// work units.
struct ViewDesc {
  uint32_t width, height, bX, bY;
  uint32_t tilesX, tilesY, pyrX, pyrY;
  uint32_t itemBase, nItems, unitBase, nUnits;
  uint32_t queryBase, nQueries;
  uint32_t batchBase, nBatches;  // the view's 8-quad batches in the launch's batch numbering
  uint16_t* depth;
  uint16_t* hiz;
  uint16_t* tileMin;
  const float* queryBounds;  // nQueries x 8 floats
  uint8_t* queryVisible;     // nQueries
  uint8_t* queryClip;        // nQueries
  uint32_t* winner;          // per pixel, for draw-order labels (null when labels are off)
  uint32_t listBase;         // first entry of the view's per-tile item lists
  uint32_t tileB;            // raster tile edge in 8x8 blocks (8 or 16); tilesX/tilesY count these tiles
  Mvp mvp;
};
static_assert(sizeof(ViewDesc) == 256, "ViewDesc layout");

enum ErrorBits : uint32_t {
  kErrTimeout = 1u,  // a frame-kernel wait did not finish in time
};

}  // namespace rzd
