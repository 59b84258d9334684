// Kernel parameter blocks and launchers of the gfx950 rasterizer kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rz_types.h"

namespace rzd {

struct PrepParams {
  const ItemDesc* items;
  ItemState* state;
  uint32_t nItems;
  uint32_t width, height;
  uint32_t tileW, tileH;  // frame-kernel tile size in blocks
  const uint32_t* rcp;
  uint32_t* done;
  uint32_t* visible;
  Mvp mvp;
};

struct SetupParams {
  const ItemDesc* items;
  ItemState* state;
  uint32_t nItems, nQuads;
  uint32_t bX, bY;
  const uint32_t* rcp;
  QuadRec* recs;
  uint64_t* bbox;
  uint64_t* groupBox;  // box of each aligned group of 64 quads (union of valid quads)
};

struct FrameParams {
  const ItemDesc* items;
  const ItemState* state;
  uint32_t nItems;
  uint32_t bX, bY, tilesX, tilesY;
  const QuadRec* recs;
  const uint64_t* bbox;
  const uint64_t* groupBox;
  uint32_t* done;
  uint32_t* visible;
  uint16_t* depth;
  uint16_t* hiz;
  const uint64_t* cov;
  uint32_t* err;
  uint64_t timeoutTicks;  // s_memrealtime ticks (100 MHz)
  uint16_t* tileMin;      // min hiZ per 8x8-block tile (pyrX per row)
  uint32_t pyrX;
  uint64_t* stats;        // optional per-tile counters (kFrameStatWords per tile), null when off
};

// Per-tile counters of the frame kernel (observability; 100 MHz ticks).
enum FrameStat : int {
  kStatTotalTicks = 0,   // kernel entry to exit
  kStatWaitTicks = 1,    // spinning on another tile's query contribution
  kStatWaits = 2,
  kStatQueryTicks = 3,   // local query2D contributions
  kStatQueries = 4,
  kStatRasterTicks = 5,  // quad list build + block loop
  kStatQuads = 6,        // quads drawn into this tile
  kStatItems = 7,        // items touching this tile
  kStatScanTicks = 8,    // overlap tests and quad compaction
  kStatPasses = 9,       // 256-quad scan passes run
  kStatLoadTicks = 10,   // table, depth and hiZ staging at entry
  kStatLoopTicks = 11,   // the whole item loop
  kStatDrawTicks = 12,   // staged-quad drawing (inside raster ticks)
  kStatSpare0 = 13,
  kStatSpare1 = 14,
  kStatSpare2 = 15,
  kFrameStatWords = 16,
};

struct QueryParams {
  const float* bounds;  // n x 8 floats
  uint32_t n;
  uint32_t width, height, bX;
  const uint16_t* depth;
  const uint16_t* hiz;
  const uint16_t* tileMin;  // min hiZ per 8x8-block tile
  uint32_t bY, pyrX;
  const uint32_t* rcp;
  uint8_t* visible;
  uint8_t* needsClipping;  // may be null
  uint8_t* code;           // may be null: visible | clip << 1 | clipWritten << 2
  Mvp mvp;
};

// Launchers (defined with the kernels); all enqueue on `stream`.
hipError_t launch_prep(const PrepParams& p, hipStream_t stream);
hipError_t launch_setup(const SetupParams& p, hipStream_t stream);
// tileW x tileH must be 8x8 or 16x16; the grid is tilesX * tilesY
// cooperative work-groups.
hipError_t launch_frame(const FrameParams& p, int tileW, hipStream_t stream);
hipError_t frame_max_groups(int tileW, int* maxGroups);
hipError_t launch_status(const ItemState* state, const uint32_t* visible, uint32_t n, uint8_t* out,
                         hipStream_t stream);
hipError_t launch_query(const QueryParams& p, hipStream_t stream);
hipError_t launch_query2d(uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY, uint32_t maxZ,
                          const QueryParams& p, uint8_t* out, hipStream_t stream);
// Recomputes the min-hiZ of every 8x8-block tile from hiz.
hipError_t launch_tile_min(const uint16_t* hiz, uint32_t bX, uint32_t bY, uint16_t* tileMin, uint32_t pyrX,
                           uint32_t pyrY, hipStream_t stream);
hipError_t launch_readback(const uint16_t* depth, const uint16_t* hiz, uint32_t width, uint32_t height, uint32_t bX,
                           uint32_t* rgba, hipStream_t stream);
hipError_t launch_fill_u16(uint16_t* p, uint32_t n, uint16_t v, hipStream_t stream);

}  // namespace rzd
