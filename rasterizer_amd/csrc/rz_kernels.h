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
};

struct FrameParams {
  const ItemDesc* items;
  const ItemState* state;
  uint32_t nItems;
  uint32_t bX, bY, tilesX, tilesY;
  const QuadRec* recs;
  const uint64_t* bbox;
  uint32_t* done;
  uint32_t* visible;
  uint16_t* depth;
  uint16_t* hiz;
  const uint64_t* cov;
  uint32_t* err;
  uint64_t timeoutTicks;  // s_memrealtime ticks (100 MHz)
};

struct QueryParams {
  const float* bounds;  // n x 8 floats
  uint32_t n;
  uint32_t width, height, bX;
  const uint16_t* depth;
  const uint16_t* hiz;
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
                          const uint16_t* depth, const uint16_t* hiz, uint32_t bX, uint8_t* out, hipStream_t stream);
hipError_t launch_readback(const uint16_t* depth, const uint16_t* hiz, uint32_t width, uint32_t height, uint32_t bX,
                           uint32_t* rgba, hipStream_t stream);
hipError_t launch_fill_u16(uint16_t* p, uint32_t n, uint16_t v, hipStream_t stream);

}  // namespace rzd
