// gfx950 kernels of the occlusion-culling rasterizer.
//
// Pipeline of one draw (a frame of items, or a single rasterize call):
//   rz_prep_kernel   one thread per item: rasterize prologue (matrix, c0, c1)
//                    and the depth-independent half of queryVisibility.
//   rz_setup_kernel  one thread per quad: transform, cull, bounding rectangle,
//                    depth plane, edge normalization (Rasterizer.cpp:657-1081).
//   rz_frame_kernel  one work-group per screen tile, all tiles co-resident:
//                    the tile's depth and hiZ live in LDS; items are walked in
//                    draw order; a query2D is reduced across the tiles that
//                    overlap its rectangle with device atomics, so each item
//                    is drawn exactly when the reference would draw it.  One
//                    wave64 works on one 8x8 block at a time with lane = pixel
//                    (Rasterizer.cpp:1095-1292).
#include "rz_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>

#include "rz_x86.h"

extern "C" __device__ uint32_t __ockl_wfred_min_u32(uint32_t);
extern "C" __device__ uint32_t __ockl_wfred_max_u32(uint32_t);

namespace rzd {

constexpr uint32_t kModeTriangle0 = 1, kModeTriangle1 = 2, kModeConcaveRight = 3, kModeConcaveLeft = 4,
                   kModeConcaveCenter = 5, kModeConvex = 6;

// modeTable of Rasterizer.cpp:30-64: primitive mode for each 8-bit
// configuration (bits 0-3: triangle area <= 0, bits 4-7: w < 0).
__constant__ uint8_t c_modeTable[256] = {
    6, 2, 4, 2, 1, 0, 1, 0, 3, 2, 0, 2, 1, 0, 1, 0, 6, 2, 4, 2, 1, 0, 1, 0, 0, 2, 5, 2, 1, 0, 1, 0,
    6, 0, 4, 2, 1, 0, 1, 0, 3, 2, 5, 2, 1, 0, 0, 0, 6, 2, 4, 0, 1, 0, 1, 0, 3, 2, 5, 2, 0, 0, 1, 0,
    6, 2, 0, 2, 1, 0, 1, 0, 3, 2, 5, 2, 1, 0, 1, 0, 0, 2, 5, 2, 1, 0, 1, 0, 5, 2, 5, 2, 1, 0, 1, 0,
    6, 2, 4, 2, 1, 0, 0, 0, 3, 0, 5, 2, 1, 0, 1, 0, 2, 2, 2, 2, 0, 0, 0, 0, 2, 2, 2, 0, 0, 0, 0, 0,
    6, 2, 4, 2, 0, 0, 1, 0, 3, 2, 5, 0, 1, 0, 1, 0, 6, 2, 4, 2, 1, 0, 0, 0, 3, 0, 5, 2, 1, 0, 1, 0,
    0, 2, 4, 2, 1, 0, 1, 0, 3, 2, 5, 2, 1, 0, 1, 0, 3, 2, 0, 2, 1, 0, 1, 0, 3, 2, 5, 2, 1, 0, 1, 0,
    6, 2, 4, 0, 1, 0, 1, 0, 3, 2, 5, 2, 0, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 0, 0,
    4, 2, 4, 2, 1, 0, 1, 0, 0, 2, 5, 2, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
};

constexpr uint64_t kEmptyBox = 0x000000000000FFFFull;  // min column 0xFFFF > max column 0

// ---------------------------------------------------------------------------
// Depth-independent half of queryVisibility (Rasterizer.cpp:125-273).
// ---------------------------------------------------------------------------
struct QueryPre {
  uint32_t kind;  // 0 frustum-culled, 1 near the camera plane, 2 empty rectangle, 3 rectangle test
  uint32_t minX, maxX, minY, maxY, maxZ;
};

__device__ QueryPre query_pre(const Mvp& M, const float* bmin, const float* bmax, uint32_t width, uint32_t height,
                              const uint32_t* __restrict__ rcpTab) {
  QueryPre out = {0, 0, 0, 0, 0, 0};
  float ext[4], cen[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    ext[j] = bmax[j] - bmin[j];
    cen[j] = bmax[j] + bmin[j];
  }
  // Six frustum planes from the raw rows; the box corner is chosen by the
  // plane's sign bits (:135-167).  A set sign bit on any distance culls.
  const float* R = M.raw;
  uint32_t signs = 0;
#pragma unroll
  for (int pl = 0; pl < 6; ++pl) {
    const int rr = pl >> 1;
    float plane[4], off[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      plane[j] = (pl & 1) ? (R[12 + j] - R[4 * rr + j]) : (R[12 + j] + R[4 * rr + j]);
      off[j] = cen[j] + xor_sign(ext[j], fbits(plane[j]) & 0x80000000u);
    }
    signs |= fbits(dp4_x86(plane[0], plane[1], plane[2], plane[3], off[0], off[1], off[2], off[3]));
  }
  if (signs & 0x80000000u) return out;

  // Eight corners with the prebaked columns (:169-198).
  const float* C = M.prebaked;
  float c[8][4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float e0 = C[0 + j] * ext[0];
    const float e1 = C[4 + j] * ext[1];
    const float e2 = C[8 + j] * ext[2];
    c[0][j] = fmaf(C[0 + j], bmin[0], fmaf(C[4 + j], bmin[1], fmaf(C[8 + j], bmin[2], C[12 + j])));
    c[1][j] = c[0][j] + e0;
    c[2][j] = c[0][j] + e1;
    c[4][j] = c[0][j] + e2;
    c[3][j] = c[1][j] + e1;
    c[5][j] = c[4][j] + e0;
    c[6][j] = c[2][j] + e2;
    c[7][j] = c[6][j] + e0;
  }
  // Per-lane near-plane epsilon (:204-213): lane j compares corners j and j+4.
  float me1[4], me2[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) me1[j] = max_x86(ext[j], ext[j ^ 2]);
#pragma unroll
  for (int j = 0; j < 4; ++j) me2[j] = max_x86(me1[j], me1[j ^ 1]);
  bool close = false;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float eps = me2[j] * 0.001f;
    close |= (c[j][3] < eps) || (c[4 + j][3] < eps);
  }
  if (close) {
    out.kind = 1;
    return out;
  }
  // Projection, screen rectangle (:217-264).
  float px[8], py[8], pz[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float rw = rcp_x86(c[k][3], rcpTab);
    px[k] = c[k][0] * rw;
    py[k] = c[k][1] * rw;
    pz[k] = c[k][2] * rw;
  }
  float minsX[4], maxsX[4], minsY[4], maxsY[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    minsX[j] = min_x86(px[j], px[4 + j]);
    maxsX[j] = max_x86(px[j], px[4 + j]);
    minsY[j] = min_x86(py[j], py[4 + j]);
    maxsY[j] = max_x86(py[j], py[4 + j]);
  }
  float mXY[4] = {min_x86(minsX[0], minsX[2]), min_x86(minsY[0], minsY[2]), min_x86(minsX[1], minsX[3]),
                  min_x86(minsY[1], minsY[3])};
  float MXY[4] = {max_x86(maxsX[0], maxsX[2]), max_x86(maxsY[0], maxsY[2]), max_x86(maxsX[1], maxsX[3]),
                  max_x86(maxsY[1], maxsY[3])};
  const float lim[4] = {float(width - 1), float(height - 1), float(width - 1), float(height - 1)};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    mXY[j] = max_x86(mXY[j], 0.0f);
    MXY[j] = -min_x86(MXY[j], lim[j]);
  }
  const float bf[4] = {min_x86(mXY[0], mXY[2]), min_x86(MXY[0], MXY[2]), min_x86(mXY[1], mXY[3]),
                       min_x86(MXY[1], MXY[3])};
  int32_t bi[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) bi[j] = cvtt_x86(floorf(bf[j]));
  bi[1] = int32_t(0u - uint32_t(bi[1]));
  bi[3] = int32_t(0u - uint32_t(bi[3]));
  if (bi[0] >= bi[1] || bi[2] >= bi[3]) {
    out.kind = 2;
    return out;
  }
  uint32_t maxZ = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) maxZ = max(maxZ, pack_depth(pz[k]));
  out.kind = 3;
  out.minX = uint32_t(bi[0]);
  out.maxX = uint32_t(bi[1]);
  out.minY = uint32_t(bi[2]);
  out.maxY = uint32_t(bi[3]);
  out.maxZ = maxZ;
  return out;
}

// query2D (Rasterizer.cpp:283-349) by one wave on the buffers in device
// memory.  The result is the OR over the rectangle's blocks of "maxZ > hiZ
// and (block interior to the rectangle, or a pixel in range nearer than
// maxZ)", so blocks may be visited in any order: lanes take 8x8-block tiles
// whose minimum hiZ is below maxZ (other tiles hold no candidate block), then
// one block per lane inside each candidate tile.
__device__ bool query2d_wave(uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY, uint32_t maxZ,
                             const uint16_t* __restrict__ depth, const uint16_t* __restrict__ hiz,
                             const uint16_t* __restrict__ tileMin, uint32_t bX, uint32_t pyrX, uint32_t lane) {
  const uint32_t bx0 = minX >> 3, bx1 = maxX >> 3, by0 = minY >> 3, by1 = maxY >> 3;
  const uint32_t tx0 = bx0 >> 3, tx1 = bx1 >> 3, ty0 = by0 >> 3, ty1 = by1 >> 3;
  const uint32_t ntx = tx1 - tx0 + 1, nt = ntx * (ty1 - ty0 + 1);
  for (uint32_t tb = 0; tb < nt; tb += 64) {
    const uint32_t t = tb + lane;
    const bool cand = t < nt && uint32_t(tileMin[(ty0 + t / ntx) * pyrX + tx0 + t % ntx]) < maxZ;
    uint64_t cm = __ballot(cand);
    while (cm) {
      const uint32_t k = uint32_t(__builtin_ctzll(cm));
      cm &= cm - 1;
      const uint32_t tt = tb + k;
      const uint32_t bx = (tx0 + tt % ntx) * 8 + (lane & 7u), by = (ty0 + tt / ntx) * 8 + (lane >> 3);
      bool v = false;
      if (bx >= bx0 && bx <= bx1 && by >= by0 && by <= by1) {
        const uint32_t b = by * bX + bx;
        if (maxZ > uint32_t(hiz[b])) {
          const int32_t sy = int32_t(minY - 8 * by), ey = int32_t(maxY - 8 * by);
          const int32_t sx = int32_t(minX - 8 * bx), ex = int32_t(maxX - 8 * bx);
          const uint32_t startY = uint32_t(sy > 0 ? sy : 0), endY = uint32_t(ey < 7 ? ey : 7);
          const uint32_t startX = uint32_t(sx > 0 ? sx : 0), endX = uint32_t(ex < 7 ? ex : 7);
          if (startY == 0 && endY == 7 && startX == 0 && endX == 7) {
            v = true;
          } else {
            // The rectangle's rows of the block, four rows' loads issued
            // before their tests (a row-by-row loop with an early exit would
            // wait for each load in turn).
            const uint4* rows = reinterpret_cast<const uint4*>(depth + size_t(b) * 64);
            for (uint32_t h = startY & ~3u; h <= endY && !v; h += 4) {
              uint4 r[4];
#pragma unroll
              for (uint32_t j = 0; j < 4; ++j) {
                r[j] = make_uint4(0, 0, 0, 0);
                if (h + j >= startY && h + j <= endY) r[j] = rows[h + j];
              }
#pragma unroll
              for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t y = h + j;
                const uint32_t w[4] = {r[j].x, r[j].y, r[j].z, r[j].w};
#pragma unroll
                for (uint32_t x = 0; x < 8; ++x)
                  v |= y >= startY && y <= endY && x >= startX && x <= endX &&
                       ((w[x >> 1] >> (16 * (x & 1))) & 0xFFFFu) < maxZ;
              }
            }
          }
        }
      }
      if (__ballot(v) != 0ull) return true;
    }
  }
  return false;
}

// ---------------------------------------------------------------------------
// Prep: rasterize prologue per item (Rasterizer.cpp:616-655) and the item's
// status for the draw loop.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) rz_prep_kernel(PrepParams p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.nItems) return;
  ItemDesc D;
  if (p.hostItems) {
    D = p.hostItems[i];
    p.itemsOut[i] = D;
  } else {
    D = p.items[i];
  }
  if (p.groupItem && D.quadCount != 0u) {
    // Items cover the quads 0..nQuads-1 in order, so each group's first
    // quad lies in exactly one non-empty item.
    const uint32_t last = (D.quadBase + D.quadCount - 1u) >> 6;
    for (uint32_t k = (D.quadBase + 63u) >> 6; k <= last; ++k) p.groupItem[k] = i;
  }
  ItemState& S = p.state[i];
  const Mvp* M = &p.mvp;
  uint32_t width = p.width, height = p.height, tileW = p.tileW, tileH = p.tileH;
  if (p.views) {
    const ViewDesc& V = p.views[D.view];
    M = &V.mvp;
    width = V.width;
    height = V.height;
    tileW = tileH = V.tileB;
  }
  if (i == 0 && p.queue) *p.queue = 0u;
  const float* P = M->prebaked;
  float ext[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) ext[j] = D.refMax[j] - D.refMin[j];
  const float k0 = 1.0f / float(2047ull << 21);
  const float k1 = 1.0f / float(2047 << 10);
  const float k2 = 1.0f / float(1023);
  float m0[4], m1[4], m2[4], m3[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    m3[j] = fmaf(P[0 + j], D.refMin[0], fmaf(P[4 + j], D.refMin[1], fmaf(P[8 + j], D.refMin[2], P[12 + j])));
    m0[j] = P[0 + j] * (ext[0] * k0);
    m1[j] = P[4 + j] * (ext[1] * k1);
    m2[j] = P[8 + j] * (ext[2] * k2);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    m3[j] = fmaf(m0[j], 2147483648.0f, m3[j]);
    m1[j] = m1[j] - m0[j];
    m2[j] = m2[j] - m0[j];
  }
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    S.T[o][0] = m0[o];
    S.T[o][1] = m1[o];
    S.T[o][2] = m2[o];
    S.T[o][3] = m3[o];
  }
  const float Za = m3[2], Zb = dp4_x86(m0[2], m1[2], m2[2], m3[2], 2097152.0f, 1024.0f, 1.0f, 1.0f);
  const float Wa = m3[3], Wb = dp4_x86(m0[3], m1[3], m2[3], m3[3], 2097152.0f, 1024.0f, 1.0f, 1.0f);
  // x86 NaN rules: a tiny reference box can make both differences 0.
  const float q = div_x86(Za - Zb, Wa - Wb);
  S.c0 = q;
  S.c1 = fnmadd_x86(q, Wa, Za);

  S.fp[0] = 0xFFFFFFFFu;
  S.fp[1] = 0u;
  S.fp[2] = 0xFFFFFFFFu;
  S.fp[3] = 0u;
  S.expected = 0;
  S.qMinX = S.qMaxX = S.qMinY = S.qMaxY = S.qMaxZ = 0;
  p.done[i] = 0;
  p.visible[i] = 0;
  uint32_t status;
  if (D.kind == kItemForceNoClip) {
    status = kStatusNoClip;
  } else if (D.kind == kItemForceClip) {
    status = kStatusClip;
  } else if (D.kind == kItemClassify) {
    // Frustum cull, near-plane test and empty rectangle only; no query2D.
    const QueryPre Q = query_pre(*M, D.boundsMin, D.boundsMax, width, height, p.rcp);
    status = Q.kind == 1 ? kStatusClip : (Q.kind == 3 ? kStatusNoClip : kStatusSkip);
  } else {
    const QueryPre Q = query_pre(*M, D.boundsMin, D.boundsMax, width, height, p.rcp);
    if (Q.kind == 1) {
      status = kStatusClip;
    } else if (Q.kind == 3) {
      status = kStatusTest;
      S.qMinX = Q.minX;
      S.qMaxX = Q.maxX;
      S.qMinY = Q.minY;
      S.qMaxY = Q.maxY;
      S.qMaxZ = Q.maxZ;
      const uint32_t tx0 = (Q.minX >> 3) / tileW, tx1 = (Q.maxX >> 3) / tileW;
      const uint32_t ty0 = (Q.minY >> 3) / tileH, ty1 = (Q.maxY >> 3) / tileH;
      S.expected = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
    } else {
      status = kStatusSkip;
    }
  }
  S.status = status;
}

// ---------------------------------------------------------------------------
// Setup: one quad per thread (Rasterizer.cpp:660-1063).
// ---------------------------------------------------------------------------
struct SetupOut {
  bool valid;
  QuadRec rec;
};

__device__ void setup_quad(const ItemState& S, const uint32_t I[4], bool clip, uint32_t bX, uint32_t bY,
                           const uint32_t* __restrict__ rcpTab, SetupOut& out) {
  const float maxInvW = 1.84467430e+19f;  // sqrt(FLT_MAX) rounded to float (Rasterizer.cpp:11)
  float x[4], y[4], invW[4];
  uint32_t ws[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float Xf = float(int32_t(I[k]));
    const float Yf = float(int32_t(I[k] & (2047u << 10)));
    const float Zf = float(int32_t(I[k] & 1023u));
    const float X = fmaf(Xf, S.T[0][0], fmaf(Yf, S.T[0][1], fmaf(Zf, S.T[0][2], S.T[0][3])));
    const float Y = fmaf(Xf, S.T[1][0], fmaf(Yf, S.T[1][1], fmaf(Zf, S.T[1][2], S.T[1][3])));
    const float W = fmaf(Xf, S.T[3][0], fmaf(Yf, S.T[3][1], fmaf(Zf, S.T[3][2], S.T[3][3])));
    const float r = rcp_x86(W, rcpTab);
    invW[k] = clip ? min_x86(maxInvW, max_x86(-maxInvW, r)) : r;
    x[k] = rintf(X * invW[k]) * 0.125f;
    y[k] = rintf(Y * invW[k]) * 0.125f;
    ws[k] = clip ? (fbits(invW[k]) & 0x80000000u) : 0u;
  }
  float eX[5], eY[5];
  eX[0] = y[1] - y[0];
  eX[1] = y[2] - y[1];
  eX[2] = y[3] - y[2];
  eX[3] = y[0] - y[3];
  eY[0] = x[0] - x[1];
  eY[1] = x[1] - x[2];
  eY[2] = x[2] - x[3];
  eY[3] = x[3] - x[0];
  const float area0 = fmaf(eX[0], eY[1], -(eX[1] * eY[0]));
  const float area1 = fmaf(eX[1], eY[2], -(eX[2] * eY[1]));
  const float area2 = fmaf(eX[2], eY[3], -(eX[3] * eY[2]));
  const float area3 = (area0 + area2) - area1;
  // Area signs with zero counted as negative; flipped per w < 0 vertex when
  // clipping (:775-803).
  const uint32_t s0 = xor_sign(xor_sign(area0, ws[0]), ws[1] ^ ws[2]) <= 0.0f;
  const uint32_t s1 = xor_sign(xor_sign(area1, ws[1]), ws[2] ^ ws[3]) <= 0.0f;
  const uint32_t s2 = xor_sign(xor_sign(area2, ws[0]), ws[2] ^ ws[3]) <= 0.0f;
  const uint32_t s3 = xor_sign(xor_sign(area3, ws[1]), ws[0] ^ ws[3]) <= 0.0f;
  const uint32_t config = s0 | (s1 << 1) | (s2 << 2) | (s3 << 3) | (ws[0] >> 27) | (ws[1] >> 26) | (ws[2] >> 25) |
                          (ws[3] >> 24);
  const uint32_t mode = c_modeTable[config];
  out.valid = false;
  if (mode == 0) return;

  float minFx, minFy, maxFx, maxFy;
  if (clip) {
    // Clipless bounding box from the intervals of w > 0 and w < 0 vertices
    // (:820-909).
    const float infP = 10000.0f, infN = -10000.0f;
    float minPx[4], minPy[4], maxPx[4], maxPy[4], minNx[4], minNy[4], maxNx[4], maxNy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool neg = ws[k] != 0u;
      minPx[k] = neg ? infP : x[k];
      minPy[k] = neg ? infP : y[k];
      maxPx[k] = xor_sign(minPx[k], ws[k]);
      maxPy[k] = xor_sign(minPy[k], ws[k]);
      minNx[k] = neg ? x[k] : infP;
      minNy[k] = neg ? y[k] : infP;
      maxNx[k] = neg ? x[k] : infN;
      maxNy[k] = neg ? y[k] : infN;
    }
    // Reductions in the reference's pairing: op(op(v0, v1), op(v2, v3)).
    const float mPx = min_x86(min_x86(minPx[0], minPx[1]), min_x86(minPx[2], minPx[3]));
    const float mPy = min_x86(min_x86(minPy[0], minPy[1]), min_x86(minPy[2], minPy[3]));
    const float MPx = max_x86(max_x86(maxPx[0], maxPx[1]), max_x86(maxPx[2], maxPx[3]));
    const float MPy = max_x86(max_x86(maxPy[0], maxPy[1]), max_x86(maxPy[2], maxPy[3]));
    const float mNx = min_x86(min_x86(minNx[0], minNx[1]), min_x86(minNx[2], minNx[3]));
    const float mNy = min_x86(min_x86(minNy[0], minNy[1]), min_x86(minNy[2], minNy[3]));
    const float MNx = max_x86(max_x86(maxNx[0], maxNx[1]), max_x86(maxNx[2], maxNx[3]));
    const float MNy = max_x86(max_x86(maxNy[0], maxNy[1]), max_x86(maxNy[2], maxNy[3]));
    const float incAx = (MNx > mPx) ? infN : mPx;
    const float incAy = (MNy > mPy) ? infN : mPy;
    const float incBx = (MPx > mNx) ? infP : MPx;
    const float incBy = (MPy > mNy) ? infP : MPy;
    minFx = min_x86(incAx, incBx);
    minFy = min_x86(incAy, incBy);
    maxFx = max_x86(incAx, incBx);
    maxFy = max_x86(incAy, incBy);
  } else {
    minFx = min_x86(min_x86(x[0], x[1]), min_x86(x[2], x[3]));
    minFy = min_x86(min_x86(y[0], y[1]), min_x86(y[2], y[3]));
    maxFx = max_x86(max_x86(x[0], x[1]), max_x86(x[2], x[3]));
    maxFy = max_x86(max_x86(y[0], y[1]), max_x86(y[2], y[3]));
  }
  int32_t minX = cvtt_x86(minFx + 0.6249875f);
  int32_t minY = cvtt_x86(minFy + 0.6249875f);
  int32_t maxX = cvtt_x86(maxFx + 1.375f);
  int32_t maxY = cvtt_x86(maxFy + 1.375f);
  minX = max(minX, 0);
  minY = max(minY, 0);
  maxX = min(maxX, int32_t(bX));
  maxY = min(maxY, int32_t(bY));
  if (!(maxX > minX && maxY > minY)) return;

  float z[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) z[k] = fmadd_x86(invW[k], S.c1, S.c0);
  float maxZ = max_x86(max_x86(z[0], z[1]), max_x86(z[2], z[3]));
  if ((ws[0] | ws[1] | ws[2] | ws[3]) != 0u) maxZ = 1.0f;

  // Depth plane from the larger triangle (:963-999).
  const bool tri0 = mode == kModeTriangle0, tri1 = mode == kModeTriangle1;
  const bool greater = !tri0 && (tri1 || (__builtin_fabsf(area0) < __builtin_fabsf(area2)));
  const float selArea = greater ? area2 : area0;
  const float invArea = clip ? 1.0f / selArea : rcp_x86(selArea, rcpTab);
  const float z12 = z[1] - z[2], z20 = z[2] - z[0], z30 = z[3] - z[0];
  eX[4] = y[0] - y[2];
  eY[4] = x[2] - x[0];
  const float dpx = greater ? fmaf(-z20, eX[3], z30 * eX[4]) : fmaf(z20, eX[1], -(z12 * eX[4]));
  const float dpy = greater ? fmaf(-z20, eY[3], z30 * eY[4]) : fmaf(z20, eY[1], -(z12 * eY[4]));
  const float dp1 = invArea * dpx;
  const float dp2 = invArea * dpy;
  const float x0 = x[0] - float(minX);
  const float y0 = y[0] - float(minY);
  const float dp0 = fmaf(-x0, dp1, fmaf(-y0, dp2, z[0]));

  // Triangle modes swap in the diagonal edge (:1001-1005); edges facing w < 0
  // are flipped (:1007-1022); normalization to 70 / (|nx| + |ny|) (:452-468).
  if (tri0) {
    eX[2] = eX[4];
    eY[2] = eY[4];
  }
  if (tri1) {
    eX[0] = -eX[4];
    eY[0] = -eY[4];
  }
  uint32_t flip[4] = {0u, 0u, 0u, 0u};
  if (clip) {
    flip[0] = ws[0] ^ (tri1 ? ws[2] : ws[1]);
    flip[1] = ws[1] ^ ws[2];
    flip[2] = ws[2] ^ (tri0 ? ws[0] : ws[3]);
    flip[3] = ws[0] ^ ws[3];
  }
  const float edgeScale = 63.0f / (0.45f - -0.45f);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    float invLen = rcp_x86(__builtin_fabsf(eX[e]) + __builtin_fabsf(eY[e]), rcpTab);
    const float mul = clip ? xor_sign(edgeScale, flip[e]) : edgeScale;
    invLen = mul * invLen;
    eX[e] = eX[e] * invLen;
    eY[e] = eY[e] * invLen;
  }
  // Edge offsets at the rectangle origin (:1030-1045).
  const float add = 0.5f - -0.45f * 63.0f / (0.45f - -0.45f);
  const float fMinX = float(minX), fMinY = float(minY);
  float off0 = fmaf(-x0, eX[0], fmaf(-y0, eY[0], add));
  float off1 = fmaf(-x[1], eX[1], fmaf(-y[1], eY[1], add));
  float off2 = fmaf(-x[2], eX[2], fmaf(-y[2], eY[2], add));
  float off3 = fmaf(-x[3], eX[3], fmaf(-y[3], eY[3], add));
  off1 = fmaf(fMinX, eX[1], off1);
  off2 = fmaf(fMinX, eX[2], off2);
  off3 = fmaf(fMinX, eX[3], off3);
  off1 = fmaf(fMinY, eY[1], off1);
  off2 = fmaf(fMinY, eY[2], off2);
  off3 = fmaf(fMinY, eY[3], off3);
  // Slope buckets with ny <= 0 selecting the odd bucket (:482-493).
  const float slopeMul = 15.5f / edgeScale;
  uint32_t slope[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int32_t q = cvtt_x86(fmaf(eX[e], slopeMul, 16.0f));
    slope[e] = ((uint32_t(q) << 1) + (eY[e] <= 0.0f ? 1u : 0u)) << 6;
  }

  QuadRec& R = out.rec;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    R.nx[e] = eX[e];
    R.ny[e] = eY[e];
  }
  R.off[0] = off0;
  R.off[1] = off1;
  R.off[2] = off2;
  R.off[3] = off3;
  R.dp0 = dp0;
  R.dDx = dp1;
  R.dDy = dp2;
  // 16-bit lane multiply of the first block index (:1054).
  R.fb = ((uint32_t(minY) * bX) & 0xFFFFu) + uint32_t(minX);
  R.range = uint32_t(maxX - minX) | (uint32_t(maxY - minY) << 16);
  R.modeZ = mode | (pack_depth(maxZ) << 16);
  R.slope01 = (slope[0] & 0xFFFFu) | (slope[1] << 16);
  R.slope23 = (slope[2] & 0xFFFFu) | (slope[3] << 16);
  out.valid = true;
}

// Conservative screen-block box of a quad's blocks fb + r*bX + c,
// 0 <= r < rangeY, 0 <= c < rangeX (exact unless the 16-bit multiply
// wrapped the row index, when it spans whole rows).
__device__ uint64_t quad_box(uint32_t fb, uint32_t rangeX, uint32_t rangeY, uint32_t bX, uint32_t bY) {
  const uint32_t col = fb % bX, row = fb / bX;
  uint32_t c0, c1, r0 = row, r1;
  if (col + rangeX <= bX) {
    c0 = col;
    c1 = col + rangeX - 1;
    r1 = row + rangeY - 1;
  } else {
    c0 = 0;
    c1 = bX - 1;
    r1 = row + rangeY;
  }
  r1 = min(r1, bY - 1);
  return uint64_t(c0) | (uint64_t(c1) << 16) | (uint64_t(r0) << 32) | (uint64_t(r1) << 48);
}

// The item that holds quad g: the last item whose quadBase <= g (items are
// in quad order; empty items share their successor's quadBase).  With a
// group table (groupItem[k] = the item of quad 64k) the search starts there
// and steps forward over the items that begin inside the 64-quad group.
__device__ __forceinline__ uint32_t item_of_quad(const ItemDesc* __restrict__ items, uint32_t nItems, uint32_t g,
                                                 const uint32_t* __restrict__ groupItem = nullptr) {
  if (groupItem) {
    uint32_t i = min(groupItem[g >> 6], nItems - 1u);
    while (i + 1 < nItems && items[i + 1].quadBase <= g) ++i;
    return i;
  }
  uint32_t lo = 0, hi = nItems - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (items[mid].quadBase <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) rz_setup_kernel(SetupParams p) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const bool inRange = g < p.nQuads;
  uint32_t item = 0;
  bool valid = false;
  uint64_t box = kEmptyBox;
  uint32_t fz = 0u;
  if (inRange) {
    item = item_of_quad(p.items, p.nItems, g, p.groupItem);
    const ItemDesc& D = p.items[item];
    const ItemState& S = p.state[item];
    const uint32_t st = S.status;
    if (st != kStatusSkip) {
      uint32_t bX = p.bX, bY = p.bY;
      if (p.views) {
        bX = p.views[D.view].bX;
        bY = p.views[D.view].bY;
      }
      const uint32_t q = g - D.quadBase;
      const uint32_t* pk = D.packets + (q >> 3) * 32 + (q & 7);
      const uint32_t I[4] = {pk[0], pk[8], pk[16], pk[24]};
      SetupOut o;
      setup_quad(S, I, st == kStatusClip, bX, bY, p.rcp, o);
      if (o.valid) {
        valid = true;
        p.recs[g] = o.rec;
        box = quad_box(o.rec.fb, o.rec.range & 0xFFFFu, o.rec.range >> 16, bX, bY);
        if ((o.rec.range & 0xFFFFu) != 0u && (o.rec.range >> 16) != 0u) fz = o.rec.modeZ | kFeatValid;
      }
    }
    // Records and boxes of items that are not drawn are never read (tiles,
    // binning and features test the item's status first), so only drawn
    // items write them; an empty rectangle marks a quad as not drawn.
    if (st != kStatusSkip) {
      if (!valid) p.recs[g].range = 0u;
      p.bbox[g] = box;
      if (p.featZ) p.featZ[g] = fz;
    }
  }
  // Footprint of each item's valid quads: one atomic per wave when the whole
  // wave works on one item, else one per valid lane.
  const uint32_t c0 = valid ? uint32_t(box & 0xFFFF) : 0xFFFFFFFFu;
  const uint32_t c1 = valid ? uint32_t((box >> 16) & 0xFFFF) : 0u;
  const uint32_t r0 = valid ? uint32_t((box >> 32) & 0xFFFF) : 0xFFFFFFFFu;
  const uint32_t r1 = valid ? uint32_t(box >> 48) : 0u;
  const uint32_t firstItem = __builtin_amdgcn_readfirstlane(item);
  const bool sameItem = __ballot(inRange && item != firstItem) == 0ull;
  const uint32_t a = __ockl_wfred_min_u32(c0), b = __ockl_wfred_max_u32(c1);
  const uint32_t c = __ockl_wfred_min_u32(r0), d = __ockl_wfred_max_u32(r1);
  // Box of this aligned group of 64 quads (the wave's), for the frame
  // kernel's coarse overlap test.
  if ((threadIdx.x & 63u) == 0u && inRange)
    p.groupBox[g >> 6] = a != 0xFFFFFFFFu
                             ? uint64_t(a) | (uint64_t(b) << 16) | (uint64_t(c) << 32) | (uint64_t(d) << 48)
                             : kEmptyBox;
  if (sameItem) {
    if ((threadIdx.x & 63u) == 0u && a != 0xFFFFFFFFu) {
      uint32_t* fp = p.state[firstItem].fp;
      atomicMin(&fp[0], a);
      atomicMax(&fp[1], b);
      atomicMin(&fp[2], c);
      atomicMax(&fp[3], d);
    }
  } else if (valid) {
    uint32_t* fp = p.state[item].fp;
    atomicMin(&fp[0], c0);
    atomicMax(&fp[1], c1);
    atomicMin(&fp[2], r0);
    atomicMax(&fp[3], r1);
  }
}

// ---------------------------------------------------------------------------
// Frame kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ uint32_t readlane_u(uint32_t v, uint32_t l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, uint32_t l) {
  return uint64_t(readlane_u(uint32_t(v), l)) | uint64_t(readlane_u(uint32_t(v >> 32), l)) << 32;
}

__device__ __forceinline__ int div_bx(int d, int bX, float invBX) {
  int q = int(float(d) * invBX);
  if (q * bX > d) --q;
  else if ((q + 1) * bX <= d) ++q;
  return q;
}

// What a tile needs of one item, gathered once per item window.
struct ItemLocal {
  uint32_t index;
  uint32_t flags;  // bit 0: query rectangle overlaps the tile, bit 1: footprint does, bits 2-3: status
  uint32_t qMinX, qMaxX, qMinY, qMaxY, qMaxZ;
  uint32_t expected;
  uint32_t quadBase, quadCount;
  uint64_t groups;  // bit k: aligned 64-quad group (quadBase >> 6) + k may touch the tile
};

constexpr uint32_t kItemWindow = 128;
constexpr uint32_t kQuadChunk = 64;   // records staged in LDS at a time
constexpr uint32_t kNoWinner = 0xFFFFFFFFu;  // pixel winner of labelled draws: none
// A scan pass tests one quad per thread: 64 * NW quads.

template <int TBX, int TBY>
struct TileSmem {
  uint64_t cov[4096];              // coverage table (Rasterizer.cpp:547-604)
  uint16_t depth[TBX * TBY * 64];  // the tile's depth, block-linear
  uint16_t hiz[TBX * TBY];
  ItemLocal items[kItemWindow];    // this window's items that touch the tile
  QuadRec recs[kQuadChunk];        // setup records of the quads being drawn
  uint32_t qid[kQuadChunk];        // their frame quad numbers (labelled draws)
  // Block-major draw (8x8-block tiles): per staged quad its rectangle in the
  // tile and origin, and per tile row the folded edge offsets and depth
  // samples at the rectangle's first column in the tile.
  uint32_t qinfo[kQuadChunk][4];
  float entry[(TBX == 8 && TBY == 8) ? kQuadChunk * TBY * 12 : 1];
  uint32_t wcnt[8];
  uint32_t part[2][8];
  // Early query answers (query-gated launches): per window item, a bit when
  // this tile has already contributed its partial, and the tile's draw count
  // at which an early evaluation last failed; per-wave results; wait state.
  uint32_t spec[kItemWindow / 32];
  uint32_t tried[kItemWindow];
  uint32_t spart[2][8];
  uint32_t wstate[2];
  uint32_t wnext;          // candidate position where the next item window starts
  uint32_t blockStats[4];  // observability builds: visited, rejected, written, fold steps
};

// Wave that owns tile-local block (lbx, lby) while a chunk of quads is drawn:
// diagonals, so that any horizontal or vertical run of blocks is shared by all
// four waves.
template <int NW>
__device__ __forceinline__ uint32_t block_owner(int lbx, int lby) { return uint32_t(lbx + lby) & uint32_t(NW - 1); }

// Adds `step` to `s` n times (uniform n), each add rounded: the reference's
// per-row and per-block folds (Rasterizer.cpp:1130-1131, 1145-1146).  IEEE
// addition commutes, so step + s and s + step are the same operation.
__device__ __forceinline__ float fold_n(float s, float step, int n) {
  for (; n >= 8; n -= 8) {
    s = s + step; s = s + step; s = s + step; s = s + step;
    s = s + step; s = s + step; s = s + step; s = s + step;
  }
  for (; n > 0; --n) s = s + step;
  return s;
}

struct QuadConst {
  uint32_t mode, maxZ;
  uint32_t sl[4];
  float dDx, dDy;
};

// One block of one quad (Rasterizer.cpp:1148-1290): hiZ reject, coverage
// mask from the four edge lookups (Convex: any offset >= 63 rejects the block;
// other modes: an empty mask does), depth rows 0/1 and 8/9 from the plane
// with rows 2-7 as rounding-up averages, max-merge unless the block is
// fast-cleared (hiZ == 1), hiZ = min.  `s` holds the folded edge offsets and
// depth samples in lanes 12g .. 12g+11 (offsets first).  The rejections come
// first and each only waits for the LDS reads it needs: most visited blocks
// end at the hiZ test.
// Per-wave counters of the block loop (observability builds only): blocks
// visited, rejected by hiZ or by the edge tests, and written.
struct BlockCounters {
  uint32_t visited = 0, rejected = 0, written = 0, foldSteps = 0;
};

template <int TBX, int TBY, bool STATS>
__device__ __forceinline__ void raster_block(TileSmem<TBX, TBY>& sm, int lb, float s, int g, uint32_t lane,
                                             int kLane, const QuadConst& Q, BlockCounters& bc) {
  if (STATS) bc.visited++;
  const uint32_t hz = sm.hiz[lb];
  if (hz >= Q.maxZ) {
    if (STATS) bc.rejected++;
    return;
  }
  // Lane L looks up edge (L & 3): its folded offset from lane 12g + (L & 3),
  // one coverage read, and lanes 0-3's words back to the wave as scalars.
  const uint32_t e = lane & 3u;
  const float oe = __shfl(s, 12 * g + int(e));
  const bool convex = Q.mode == kModeConvex;
  if (convex && (__ballot(oe >= 63.0f) & 0xFull) != 0ull) {
    if (STATS) bc.rejected++;
    return;
  }
  const uint32_t sl = e == 0u ? Q.sl[0] : e == 1u ? Q.sl[1] : e == 2u ? Q.sl[2] : Q.sl[3];
  const int32_t t = cvtt_x86(oe);
  const uint32_t tc = convex ? uint32_t(max(t, 0)) : uint32_t(min(max(t, 0), 63));
  const uint64_t word = sm.cov[(sl | tc) & 4095u];
  const uint64_t A = readlane_u64(word, 0), B = readlane_u64(word, 1), C = readlane_u64(word, 2),
                 D = readlane_u64(word, 3);
  uint64_t mask;
  if (Q.mode == kModeConvex) mask = (A & B) & (C & D);
  else if (Q.mode == kModeTriangle0) mask = A & B & C;
  else if (Q.mode == kModeTriangle1) mask = A & C & D;
  else if (Q.mode == kModeConcaveRight) mask = (A | D) & (B & C);
  else if (Q.mode == kModeConcaveCenter) mask = (A & B) | (C & D);
  else mask = (A & D) & (B | C);
  if (Q.mode != kModeConvex && mask == 0) {
    if (STATS) bc.rejected++;
    return;
  }
  if (STATS) bc.written++;

  const uint32_t old = sm.depth[lb * 64 + int(lane)];
  float v = __shfl(s, 12 * g + 4 + kLane);
  const uint32_t px = lane & 7u, py = lane >> 3;
  if (px >= 4u) v = fmaf(Q.dDx, 0.5f, v);
  const uint32_t lo = pack_depth(v);
  const uint32_t hi = pack_depth(Q.dDy + v);
  const uint32_t mid = avg_u16(lo, hi);
  const uint32_t q1 = avg_u16(lo, mid), q3 = avg_u16(mid, hi);
  const uint32_t j = py >> 1;
  const uint32_t a01 = (j & 1u) ? q1 : lo, a23 = (j & 1u) ? q3 : mid;
  uint32_t d = (j & 2u) ? a23 : a01;
  const uint32_t maskBit = 8u * px + ((py & 1u) ? (py >> 1) : 4u + (py >> 1));
  d = ((mask >> maskBit) & 1ull) ? d : 0u;
  if (hz != 1u) d = max(d, old);
  sm.depth[lb * 64 + int(lane)] = uint16_t(d);
  const uint32_t mn = __ockl_wfred_min_u32(d);
  if (lane == 0) sm.hiz[lb] = uint16_t(mn);
}

// Draws the blocks of one quad that this wave owns in the tile
// (Rasterizer.cpp:1095-1292).  Blocks are independent, so the order among a
// quad's blocks does not matter; the order among quads is the caller's.  The
// per-block values are the reference's folds: the edge offsets and the eight
// depth samples of rows 0/1 start at the rectangle origin, add the row step
// once per block row, then the column step once per block column, each add
// rounded.  Lane 12g + e carries component e (4 offsets, 8 depth samples) for
// the g-th of up to five rows handled together.
template <int TBX, int TBY, int NW, bool STATS>
__device__ __forceinline__ void raster_quad_wave(const QuadRec& R, TileSmem<TBX, TBY>& sm, int tbx0, int tbx1,
                                                 int tby0, int tby1, int bX, float invBX, uint32_t wave,
                                                 uint32_t lane, BlockCounters& bc) {
  const int fb = int(R.fb);
  const int rangeX = int(R.range & 0xFFFFu), rangeY = int(R.range >> 16);
  QuadConst Q;
  Q.mode = R.modeZ & 0xFFu;
  Q.maxZ = R.modeZ >> 16;
  Q.sl[0] = R.slope01 & 0xFFFFu;
  Q.sl[1] = R.slope01 >> 16;
  Q.sl[2] = R.slope23 & 0xFFFFu;
  Q.sl[3] = R.slope23 >> 16;
  Q.dDx = R.dDx;
  Q.dDy = R.dDy;
  const int minY = div_bx(fb, bX, invBX), minX = fb - minY * bX;
  const int kLane = int(((lane & 7u) & 3u) + 4u * ((lane >> 3) & 1u));

  const uint32_t g = lane / 12u, e = lane - 12u * g;
  float base, stepY, stepX;
  if (e < 4) {
    base = e == 0 ? R.off[0] : e == 1 ? R.off[1] : e == 2 ? R.off[2] : R.off[3];
    stepY = e == 0 ? R.ny[0] : e == 1 ? R.ny[1] : e == 2 ? R.ny[2] : R.ny[3];
    stepX = e == 0 ? R.nx[0] : e == 1 ? R.nx[1] : e == 2 ? R.nx[2] : R.nx[3];
  } else {
    const uint32_t k = e - 4;
    const float spx = -0.4375f + 0.125f * float(k & 3u);
    const float spy = -0.4375f + 0.125f * float(k >> 2);
    base = fmaf(Q.dDx, spx, fmaf(Q.dDy, spy, R.dp0));
    stepY = Q.dDy;
    stepX = Q.dDx;
  }

  if (minX + rangeX > bX) {
    // The 16-bit multiply of the first block index wrapped (Rasterizer.cpp:1054):
    // the quad's blocks fb + r*bX + c no longer form a screen rectangle, so
    // each owned block finds its (r, c) and folds from the origin.
    for (int by = tby0; by <= tby1; ++by)
      for (int bx = tbx0; bx <= tbx1; ++bx) {
        if (block_owner<NW>(bx - tbx0, by - tby0) != wave) continue;
        const int d = by * bX + bx - fb;
        if (d < 0) continue;
        const int r = div_bx(d, bX, invBX), c = d - r * bX;
        if (r >= rangeY || c >= rangeX) continue;
        const float s = fold_n(fold_n(base, stepY, r), stepX, c);
        raster_block<TBX, TBY, STATS>(sm, (by - tby0) * TBX + (bx - tbx0), s, 0, lane, kLane, Q, bc);
      }
    return;
  }
  const int x0 = max(tbx0, minX), x1 = min(tbx1, minX + rangeX - 1);
  const int y0 = max(tby0, minY), y1 = min(tby1, minY + rangeY - 1);
  if (x0 > x1 || y0 > y1) return;
  // Skip quads none of whose blocks in the tile belong to this wave.
  if ((x1 - x0) + (y1 - y0) < 3) {
    bool mine = false;
    for (int by = y0; by <= y1; ++by)
      for (int bx = x0; bx <= x1; ++bx) mine |= block_owner<NW>(bx - tbx0, by - tby0) == wave;
    if (!mine) return;
  }
  const int cLo = x0 - minX;
  // Lane group g starts at quad row (y0 - minY) + g.
  float line = fold_n(base, stepY, y0 - minY);
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i)
    if (i < g) line = line + stepY;
  for (int py = y0; py <= y1; py += 5) {
    float s = fold_n(line, stepX, cLo);
    for (int bx = x0; bx <= x1; ++bx, s = stepX + s) {
      for (int gg = 0; gg < 5 && py + gg <= y1; ++gg) {
        const int lbx = bx - tbx0, lby = py + gg - tby0;
        if (block_owner<NW>(lbx, lby) != wave) continue;
        raster_block<TBX, TBY, STATS>(sm, lby * TBX + lbx, s, gg, lane, kLane, Q, bc);
      }
    }
    line = fold_n(line, stepY, 5);
  }
}

// One block of one quad in the block-major draw, the part that depends on
// nothing the previous quads wrote (Rasterizer.cpp:1148-1290): the four edge
// lookups and the coverage mask (Convex blocks with an offset >= 63 and other
// modes' empty masks set `skip`, the same in every lane), depth rows 0/1 and
// 8/9 from the plane, packed averages between, 0 outside the mask.  `o` holds
// the block's four folded edge offsets (the same values in every lane) and
// `v` the lane's row-0/1 depth sample (sample kLane of the eight), as
// raster_block reads them from its lane groups.  The hiZ reject and the merge
// are the caller's, in submission order.
__device__ __forceinline__ void sample_block(const uint64_t* __restrict__ cov, const float (&o)[4], float v,
                                             uint32_t lane, const QuadConst& Q, uint32_t& dOut, bool& skip) {
  const bool convex = Q.mode == kModeConvex;
  bool outside = false;
  uint32_t idx[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int32_t t = cvtt_x86(o[e]);
    outside |= o[e] >= 63.0f;
    const uint32_t tc = convex ? uint32_t(max(t, 0)) : uint32_t(min(max(t, 0), 63));
    idx[e] = (Q.sl[e] | tc) & 4095u;
  }
  const uint64_t A = cov[idx[0]], B = cov[idx[1]], C = cov[idx[2]], D = cov[idx[3]];
  const uint64_t mConvex = (A & B) & (C & D);
  const uint64_t mTri0 = A & B & C, mTri1 = A & C & D;
  const uint64_t mRight = (A | D) & (B & C), mCenter = (A & B) | (C & D), mLeft = (A & D) & (B | C);
  const uint64_t mask = convex ? mConvex
                        : Q.mode == kModeTriangle0 ? mTri0
                        : Q.mode == kModeTriangle1 ? mTri1
                        : Q.mode == kModeConcaveRight ? mRight
                        : Q.mode == kModeConcaveCenter ? mCenter
                                                       : mLeft;
  // Every lane computes the same test from the same values; the ballot
  // makes it a wave-uniform value for the caller's branch.
  skip = __ballot(convex ? outside : mask == 0ull) != 0ull;
  const uint32_t px = lane & 7u, py = lane >> 3;
  if (px >= 4u) v = fmaf(Q.dDx, 0.5f, v);
  const uint32_t lo = pack_depth(v);
  const uint32_t hi = pack_depth(Q.dDy + v);
  const uint32_t mid = avg_u16(lo, hi);
  const uint32_t q1 = avg_u16(lo, mid), q3 = avg_u16(mid, hi);
  const uint32_t j = py >> 1;
  const uint32_t a01 = (j & 1u) ? q1 : lo, a23 = (j & 1u) ? q3 : mid;
  const uint32_t d = (j & 2u) ? a23 : a01;
  const uint32_t maskBit = 8u * px + ((py & 1u) ? (py >> 1) : 4u + (py >> 1));
  dOut = ((mask >> maskBit) & 1ull) ? d : 0u;
}


// sample_block with the four edge lookups spread over the lanes: lane L
// holds edge (L & 3)'s folded offset in `oe` and looks up its coverage word;
// lanes 0-3's words come back to the wave as scalars, where the mode's mask
// logic runs.  `sl01` / `sl23` are the quad's packed slope lookups.  The
// same values as sample_block, with one LDS lookup instead of four and the
// per-edge index arithmetic done once for all four edges.
__device__ __forceinline__ void sample_block_lanes(const uint64_t* __restrict__ cov, float oe, float v, uint32_t lane,
                                                   const QuadConst& Q, uint32_t sl01, uint32_t sl23, uint32_t& dOut,
                                                   bool& skip) {
  const bool convex = Q.mode == kModeConvex;
  const uint32_t e = lane & 3u;
  const uint32_t slw = (e & 2u) ? sl23 : sl01;
  const uint32_t sl = (e & 1u) ? (slw >> 16) : (slw & 0xFFFFu);
  const int32_t t = cvtt_x86(oe);
  const uint32_t tc = convex ? uint32_t(max(t, 0)) : uint32_t(min(max(t, 0), 63));
  const uint64_t word = cov[(sl | tc) & 4095u];
  const uint64_t A = readlane_u64(word, 0), B = readlane_u64(word, 1), C = readlane_u64(word, 2),
                 D = readlane_u64(word, 3);
  const bool outside = (__ballot(oe >= 63.0f) & 0xFull) != 0ull;
  uint64_t mask;
  if (convex) mask = (A & B) & (C & D);
  else if (Q.mode == kModeTriangle0) mask = A & B & C;
  else if (Q.mode == kModeTriangle1) mask = A & C & D;
  else if (Q.mode == kModeConcaveRight) mask = (A | D) & (B & C);
  else if (Q.mode == kModeConcaveCenter) mask = (A & B) | (C & D);
  else mask = (A & D) & (B | C);
  skip = convex ? outside : mask == 0ull;
  const uint32_t px = lane & 7u, py = lane >> 3;
  if (px >= 4u) v = fmaf(Q.dDx, 0.5f, v);
  const uint32_t lo = pack_depth(v);
  const uint32_t hi = pack_depth(Q.dDy + v);
  const uint32_t mid = avg_u16(lo, hi);
  const uint32_t q1 = avg_u16(lo, mid), q3 = avg_u16(mid, hi);
  const uint32_t j = py >> 1;
  const uint32_t a01 = (j & 1u) ? q1 : lo, a23 = (j & 1u) ? q3 : mid;
  const uint32_t d = (j & 2u) ? a23 : a01;
  const uint32_t maskBit = 8u * px + ((py & 1u) ? (py >> 1) : 4u + (py >> 1));
  dOut = ((mask >> maskBit) & 1ull) ? d : 0u;
}

// s + step, n times (0 <= n <= 7), each add rounded, without branches.
__device__ __forceinline__ float fold_upto7(float s, float step, int n) {
#pragma unroll
  for (int k = 0; k < 7; ++k) s = k < n ? s + step : s;
  return s;
}

__device__ __forceinline__ QuadConst quad_const(const QuadRec& R) {
  QuadConst Q;
  Q.mode = R.modeZ & 0xFFu;
  Q.maxZ = R.modeZ >> 16;
  Q.sl[0] = R.slope01 & 0xFFFFu;
  Q.sl[1] = R.slope01 >> 16;
  Q.sl[2] = R.slope23 & 0xFFFFu;
  Q.sl[3] = R.slope23 >> 16;
  Q.dDx = R.dDx;
  Q.dDy = R.dDy;
  return Q;
}

// Component e of a quad's fold at its rectangle origin and its row and column
// steps: e < 4 edge offset e, else depth sample e - 4 (rows 0/1 of a block).
__device__ __forceinline__ void quad_component(const QuadRec& R, uint32_t e, float& base, float& stepY,
                                               float& stepX) {
  if (e < 4) {
    base = e == 0 ? R.off[0] : e == 1 ? R.off[1] : e == 2 ? R.off[2] : R.off[3];
    stepY = e == 0 ? R.ny[0] : e == 1 ? R.ny[1] : e == 2 ? R.ny[2] : R.ny[3];
    stepX = e == 0 ? R.nx[0] : e == 1 ? R.nx[1] : e == 2 ? R.nx[2] : R.nx[3];
  } else {
    const uint32_t k = e - 4;
    const float spx = -0.4375f + 0.125f * float(k & 3u);
    const float spy = -0.4375f + 0.125f * float(k >> 2);
    base = fmaf(R.dDx, spx, fmaf(R.dDy, spy, R.dp0));
    stepY = R.dDy;
    stepX = R.dDx;
  }
}

// Block-major draw of a staged chunk of n quads in an 8x8-block tile.
// Phase A: each quad's rectangle in the tile, and for each of its rows in the
// tile the twelve folded components at the rectangle's first column (row
// steps from the origin, then column steps, each add rounded, as the block
// loop folds them); quads are split over the waves.  Phase B: each wave walks
// its blocks; for a block, lane q tests staged quad q against the block
// (inside its rectangle, and not below the block's hiZ unless the block is
// fast-cleared, since hiZ only grows once written), and the survivors are
// drawn in submission order with the block's hiZ and depth in registers,
// each finishing its fold with the column steps left.
//
// LABELS: each block's pixel winners (`winner`, the view's block-linear
// array) are loaded with its depth unless the block is cleared (its first
// write sets them all) and stored with it; quad q's batch is qid[q] / 8.
//
// Returns (observability builds) the s_memrealtime stamp at which phase B
// started.
template <int TBX, int TBY, int NW, bool STATS, bool LABELS>
__device__ uint64_t draw_chunk_block_major(TileSmem<TBX, TBY>& sm, uint32_t n, int tbx0, int tbx1, int tby0, int tby1,
                                           int bX, float invBX, uint32_t wave, uint32_t lane, BlockCounters& bc,
                                           uint32_t* __restrict__ winner) {
  // The wave index as a scalar, so that block indices and fold counts
  // derived from it are wave-uniform.
  wave = __builtin_amdgcn_readfirstlane(wave);
  const uint32_t g = lane / 12u, e = lane - 12u * g;
  for (uint32_t q = wave; q < n; q += NW) {
    const QuadRec& R = sm.recs[q];
    const int fb = int(R.fb);
    const int rangeX = int(R.range & 0xFFFFu), rangeY = int(R.range >> 16);
    const int minY = div_bx(fb, bX, invBX), minX = fb - minY * bX;
    const bool wrap = minX + rangeX > bX;
    int x0 = 0, x1 = TBX - 1, y0 = 0, y1 = TBY - 1;
    if (!wrap) {
      x0 = max(tbx0, minX) - tbx0;
      x1 = min(tbx1, minX + rangeX - 1) - tbx0;
      y0 = max(tby0, minY) - tby0;
      y1 = min(tby1, minY + rangeY - 1) - tby0;
    }
    const bool empty = x0 > x1 || y0 > y1;
    if (lane == 0) {
      sm.qinfo[q][0] = empty ? 0xFFu : uint32_t(x0) | uint32_t(x1) << 8 | uint32_t(y0) << 16 | uint32_t(y1) << 24;
      sm.qinfo[q][1] = uint32_t(minX);
      sm.qinfo[q][2] = uint32_t(minY);
      sm.qinfo[q][3] = wrap ? 1u : 0u;
    }
    if (wrap || empty) continue;
    float base, stepY, stepX;
    quad_component(R, e, base, stepY, stepX);
    const int cLo = x0 + tbx0 - minX;
    if (STATS && lane == 0) bc.foldSteps += uint32_t(y0 + tby0 - minY + ((y1 - y0) / 5 + 1) * (cLo + 5));
    float line = fold_n(base, stepY, y0 + tby0 - minY);
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i)
      if (i < g) line = line + stepY;
    for (int r0 = y0; r0 <= y1; r0 += 5) {
      const float sv = fold_n(line, stepX, cLo);
      const int row = r0 + int(g);
      if (g < 5u && row <= y1) sm.entry[(q * TBY + uint32_t(row)) * 12u + e] = sv;
      line = fold_n(line, stepY, 5);
    }
  }
  __syncthreads();
  const uint64_t tB = STATS ? __builtin_amdgcn_s_memrealtime() : 0;

  const int kLane = int(((lane & 7u) & 3u) + 4u * ((lane >> 3) & 1u));
  const bool valid = lane < n;
  // Lane q holds staged quad q's constants for phase B (read with readlane):
  // its rectangle in the tile, wrap flag, mode and packed maximum depth,
  // slope lookups and depth gradients.
  const uint32_t rect = valid ? sm.qinfo[lane][0] : 0xFFu;
  const uint32_t wrapq = valid ? sm.qinfo[lane][3] : 0u;
  const uint32_t modeZq = valid ? sm.recs[lane].modeZ : 0u;
  const uint32_t maxZq = modeZq >> 16;
  const uint32_t sl01q = valid ? sm.recs[lane].slope01 : 0u, sl23q = valid ? sm.recs[lane].slope23 : 0u;
  const float dDxq = valid ? sm.recs[lane].dDx : 0.0f, dDyq = valid ? sm.recs[lane].dDy : 0.0f;
  const int qx0 = int(rect & 0xFFu), qx1 = int((rect >> 8) & 0xFFu);
  const int qy0 = int((rect >> 16) & 0xFFu), qy1 = int(rect >> 24);
  const uint32_t ce = lane < 12u ? lane : 0u;
  // The wave's blocks: (lbx + lby) & (NW - 1) == wave, TBX / NW per row.
  static_assert(TBX % NW == 0, "diagonal block ownership");
  constexpr int kPerRow = TBX / NW, kOwned = TBX * TBY / NW;
  // Two blocks are drawn at a time, two candidates per step: the samples
  // depend on nothing the merges change, so they are computed together and
  // only the merges run in submission order.
  //
  // A block's state: until its first write in this chunk, the hiZ it was
  // opened with (which the reject test and the fast-clear test read); after
  // it, hiZ is the minimum of the depth in registers, so the reject test is a
  // ballot over the depth (hiZ >= maxZ exactly when no pixel is below maxZ)
  // and the fast-clear test (hiZ == 1) a flag kept by two ballots.  The
  // minimum itself is taken once, when the block is closed.
  struct Slot {
    int lb = 0;  // lbx = lb % TBX, lby = lb / TBX
    uint32_t hz = 0, dval = 0, win = kNoWinner;
    bool dirty = false, one = false;
    uint64_t m = 0;  // candidates left, in submission order
  };
  int nextK = 0;
  // Next owned block with candidates into S (false when none is left).
  auto open = [&](Slot& S) -> bool {
    while (nextK < kOwned) {
      const int k = nextK++;
      const int lby = k / kPerRow;
      const int lbx = int((wave - uint32_t(lby)) & uint32_t(NW - 1)) + NW * (k % kPerRow);
      const int lb = lby * TBX + lbx;
      const uint32_t hz = sm.hiz[lb];
      const bool inRect = valid && lbx >= qx0 && lbx <= qx1 && lby >= qy0 && lby <= qy1;
      const bool cand = inRect && !(hz != 1u && hz >= maxZq);
      const uint64_t m = __ballot(cand);
      if (STATS) {
        const uint32_t pre = uint32_t(__popcll(__ballot(inRect) & ~m));
        bc.visited += pre;
        bc.rejected += pre;
      }
      if (m == 0ull) continue;
      S.lb = lb;
      S.hz = hz;
      S.m = m;
      S.dval = sm.depth[lb * 64 + int(lane)];
      S.dirty = false;
      S.one = false;
      S.win = kNoWinner;
      if (LABELS && hz != 1u) S.win = winner[(size_t(tby0 + lby) * size_t(bX) + size_t(tbx0 + lbx)) * 64u + lane];
      return true;
    }
    return false;
  };
  auto close = [&](const Slot& S) {
    if (!S.dirty) return;
    sm.depth[S.lb * 64 + int(lane)] = uint16_t(S.dval);
    const uint32_t mn = __ockl_wfred_min_u32(S.dval);
    if (lane == 0) sm.hiz[S.lb] = uint16_t(mn);
    if (LABELS)
      winner[(size_t(tby0 + S.lb / TBX) * size_t(bX) + size_t(tbx0 + S.lb % TBX)) * 64u + lane] = S.win;
  };
  auto qconst = [&](uint32_t q) -> QuadConst {
    QuadConst Q;
    const uint32_t mz = readlane_u(modeZq, q), s01 = readlane_u(sl01q, q), s23 = readlane_u(sl23q, q);
    Q.mode = mz & 0xFFu;
    Q.maxZ = mz >> 16;
    Q.sl[0] = s01 & 0xFFFFu;
    Q.sl[1] = s01 >> 16;
    Q.sl[2] = s23 & 0xFFFFu;
    Q.sl[3] = s23 >> 16;
    Q.dDx = readlane_f(dDxq, q);
    Q.dDy = readlane_f(dDyq, q);
    return Q;
  };
  // Staged quad q's sample for block lb when its first-block index did not
  // wrap: phase A's components at the rectangle's first column in this row,
  // folded with the column steps left; every lane folds the four edge
  // offsets and its own depth sample.
  auto sample_fast = [&](int lb, uint32_t q, uint32_t& d, bool& skip) {
    const QuadConst Q = qconst(q);
    const int c = lb % TBX - int(readlane_u(rect, q) & 0xFFu);
    const float* E = &sm.entry[(q * TBY + uint32_t(lb / TBX)) * 12u];
    const float step = sm.recs[q].nx[lane & 3u];
    float oe = E[lane & 3u];
    float v = E[4 + kLane];
    // c is wave-uniform (0..7): a scalar loop of two adds a step, the lane's
    // edge offset and its depth sample.
    for (int k = 0; k < c; ++k) {
      oe = oe + step;
      v = v + Q.dDx;
    }
    sample_block_lanes(sm.cov, oe, v, lane, Q, readlane_u(sl01q, q), readlane_u(sl23q, q), d, skip);
  };
  // Any staged quad: false when a quad with a wrapped first-block index does
  // not cover the block; such a quad's block (row, column) in its block
  // numbering is found and its twelve components folded from the origin.
  auto sample_any = [&](int lb, uint32_t q, uint32_t& d, bool& skip) -> bool {
    if (!(readlane_u(wrapq, q) & 1u)) {
      sample_fast(lb, q, d, skip);
      return true;
    }
    const QuadRec& R = sm.recs[q];
    const int dd = (tby0 + lb / TBX) * bX + tbx0 + lb % TBX - int(R.fb);
    const int rangeX = int(R.range & 0xFFFFu), rangeY = int(R.range >> 16);
    if (dd < 0) return false;
    const int r = div_bx(dd, bX, invBX), c = dd - r * bX;
    if (r >= rangeY || c >= rangeX) return false;
    float base, stepY, stepX;
    quad_component(R, ce, base, stepY, stepX);
    const float s = fold_n(fold_n(base, stepY, r), stepX, c);
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = readlane_f(s, k);
    const float v = __shfl(s, 4 + kLane);
    sample_block(sm.cov, o, v, lane, qconst(q), d, skip);
    return true;
  };
  // Candidate q of slot S, in submission order: the hiZ reject, then the
  // merge (max unless the block is fast-cleared), Rasterizer.cpp:1148-1290.
  //
  // LABELS: `win` is the lane's pixel's winner, the batch whose write
  // produced its depth: the first write of a cleared block (hiZ == 1) sets it
  // (none where the written depth is 0), later writes only where they raise
  // the depth.
  auto visit = [&](Slot& S, uint32_t q, uint32_t d, bool skip) {
    const uint32_t maxZ = readlane_u(maxZq, q);
    if (STATS) bc.visited++;
    const bool rejected = S.dirty ? __ballot(S.dval < maxZ) == 0ull : S.hz >= maxZ;
    if (rejected || skip) {
      if (STATS) bc.rejected++;
      return;
    }
    if (STATS) bc.written++;
    const bool cleared = S.dirty ? S.one : S.hz == 1u;
    if (LABELS) {
      const uint32_t batch = sm.qid[q] >> 3;
      S.win = cleared ? (d != 0u ? batch : kNoWinner) : (d > S.dval ? batch : S.win);
    }
    if (!cleared) d = max(d, S.dval);
    S.dval = d;
    // hiZ == 1 after this write: no pixel at 0 and one at 1.
    S.one = __ballot(d == 0u) == 0ull && __ballot(d == 1u) != 0ull;
    S.dirty = true;
  };
  // Two samples per step: the next candidates of both open blocks, or the
  // next two of the only open one (kept in A).
  Slot A, B;
  bool aOn = open(A), bOn = open(B);
  while (aOn || bOn) {
    if (!aOn) {
      A = B;
      aOn = true;
      bOn = false;
    }
    // An open slot has at least one candidate left.
    const uint32_t qa = uint32_t(__builtin_ctzll(A.m));
    A.m &= A.m - 1ull;
    uint32_t qb = 0;
    const bool second = bOn || A.m != 0ull;
    if (bOn) {
      qb = uint32_t(__builtin_ctzll(B.m));
      B.m &= B.m - 1ull;
    } else if (second) {
      qb = uint32_t(__builtin_ctzll(A.m));
      A.m &= A.m - 1ull;
    }
    const int lbb = bOn ? B.lb : A.lb;
    uint32_t da = 0, db = 0;
    bool ka = true, kb = true, ca = true, cb = second;
    if (!((readlane_u(wrapq, qa) | readlane_u(wrapq, qb)) & 1u)) {
      // Without a second candidate, quad 0 is sampled and not used.
      sample_fast(A.lb, qa, da, ka);
      sample_fast(lbb, qb, db, kb);
    } else {
      ca = sample_any(A.lb, qa, da, ka);
      cb = cb && sample_any(lbb, qb, db, kb);
    }
    if (ca) visit(A, qa, da, ka);
    if (cb) {
      if (bOn) visit(B, qb, db, kb);
      else visit(A, qb, db, kb);
    }
    if (A.m == 0ull) {
      close(A);
      aOn = open(A);
    }
    if (bOn && B.m == 0ull) {
      close(B);
      bOn = open(B);
    }
  }
  return tB;
}

// One row of blocks of query2D inside the tile, one lane per block:
// returns 1 if the wave finds a visible pixel in row lr, and with SPEC 2 if
// a block of the rectangle in that row has hiZ 0 or 1; 0 otherwise.  The
// hiZ tests of the row's blocks run together; only blocks that pass them are
// tested pixel by pixel.
template <int TBX, int TBY, bool SPEC>
__device__ __forceinline__ uint32_t query_row(const ItemLocal& S, const TileSmem<TBX, TBY>& sm, int lr, int by,
                                              int tbx0, int tbx1, uint32_t lane) {
  static_assert(TBX <= 64, "one lane per block of a tile row");
  const uint32_t maxZ = S.qMaxZ;
  const int qbx0 = int(S.qMinX >> 3), qbx1 = int(S.qMaxX >> 3);
  const int startY = max(int(S.qMinY) - 8 * by, 0), endY = min(int(S.qMaxY) - 8 * by, 7);
  const int cx0 = max(tbx0, qbx0), cx1 = min(tbx1, qbx1);
  const int n = cx1 - cx0 + 1;
  bool cleared = false, cand = false, full = false;
  if (int(lane) < n) {
    const int bx = cx0 + int(lane);
    const uint32_t hz = sm.hiz[lr * TBX + (bx - tbx0)];
    cleared = hz <= 1u;
    cand = maxZ > hz;
    full = cand && startY == 0 && endY == 7 && int(S.qMinX) <= 8 * bx && int(S.qMaxX) >= 8 * bx + 7;
  }
  if (SPEC && __ballot(cleared) != 0ull) return 2u;
  if (__ballot(full) != 0ull) return 1u;
  uint64_t m = __ballot(cand);
  const int px = int(lane & 7u), py = int(lane >> 3);
  while (m) {
    const int bx = cx0 + int(__builtin_ctzll(m));
    m &= m - 1ull;
    const int lb = lr * TBX + (bx - tbx0);
    const int startX = max(int(S.qMinX) - 8 * bx, 0), endX = min(int(S.qMaxX) - 8 * bx, 7);
    const bool mine = py >= startY && py <= endY && px >= startX && px <= endX &&
                      uint32_t(sm.depth[lb * 64 + int(lane)]) < maxZ;
    if (__ballot(mine) != 0ull) return 1u;
  }
  return 0u;
}

// query2D restricted to this wave's rows of the tile (Rasterizer.cpp:283-349):
// true if a pixel of the rectangle in those rows is visible.  Any order of
// the blocks gives the same answer.
template <int TBX, int TBY, int NW>
__device__ __forceinline__ bool query_wave(const ItemLocal& S, TileSmem<TBX, TBY>& sm, int tbx0, int tbx1,
                                           int tby0, int tby1, uint32_t wave, uint32_t lane) {
  const int qby0 = int(S.qMinY >> 3), qby1 = int(S.qMaxY >> 3);
  for (int lr = int(wave); lr < TBY; lr += NW) {
    const int by = tby0 + lr;
    if (by > tby1) break;
    if (by < qby0 || by > qby1) continue;
    if (query_row<TBX, TBY, false>(S, sm, lr, by, tbx0, tbx1, lane)) return true;
  }
  return false;
}

// query2D of this wave's rows evaluated ahead of the item's turn: bit 0 when
// a pixel is visible, bit 1 when a block of the rectangle has hiZ 0 or 1.
// A block with hiZ >= 2 can never read as cleared (hiZ == 1) again, since its
// minimum only grows: every later write max-merges, so its depth and hiZ only
// grow, and every test of query2D (the hiZ skip and the row test) can only
// move from visible to not visible.  So a result of 0 from every wave is the
// answer the tile gives at the item's turn.  A cleared block's next write
// replaces its stale depth, and a block at hiZ 0 could reach exactly 1, so
// neither proves anything.
template <int TBX, int TBY, int NW>
__device__ __forceinline__ uint32_t spec_query_wave(const ItemLocal& S, TileSmem<TBX, TBY>& sm, int tbx0, int tbx1,
                                                    int tby0, int tby1, uint32_t wave, uint32_t lane) {
  const int qby0 = int(S.qMinY >> 3), qby1 = int(S.qMaxY >> 3);
  for (int lr = int(wave); lr < TBY; lr += NW) {
    const int by = tby0 + lr;
    if (by > tby1) break;
    if (by < qby0 || by > qby1) continue;
    const uint32_t r = query_row<TBX, TBY, true>(S, sm, lr, by, tbx0, tbx1, lane);
    if (r) return r;
  }
  return 0u;
}

__device__ __forceinline__ bool box_touches(uint64_t bb, int tbx0, int tbx1, int tby0, int tby1) {
  const int c0 = int(bb & 0xFFFF), c1 = int((bb >> 16) & 0xFFFF), r0 = int((bb >> 32) & 0xFFFF), r1 = int(bb >> 48);
  return c0 <= c1 && c0 <= tbx1 && c1 >= tbx0 && r0 <= tby1 && r1 >= tby0;
}

// True when no pixel of the tile can be covered by the quad because some edge
// is outside for every block of the tile.  The block loop folds each edge
// offset as off + nY + ... + nY (row steps) + nX + ... + nX (column steps) in
// float; after k steps the fold differs from the exact off + r*nY + c*nX by at
// most k * max|partial sum| * 2^-24 (each rounding is within half an ulp), so
// an exact minimum at least 63 + that bound proves the offset test
// (Rasterizer.cpp:1158) or the clamped lookup of bucket 63, whose table entry
// is empty (:599), rejects every block.  For the non-convex modes the offset
// must also stay below 2^31, where the truncation would wrap to bucket 0.
// Quads whose block index wrapped in the 16-bit multiply are never rejected.
__device__ bool quad_misses_tile(const QuadRec& R, int tbx0, int tbx1, int tby0, int tby1, int bX) {
  const int fb = int(R.fb), rangeX = int(R.range & 0xFFFFu), rangeY = int(R.range >> 16);
  const int minX = fb % bX, minY = fb / bX;
  if (minX + rangeX > bX) return false;
  const int c0 = max(tbx0, minX) - minX, c1 = min(tbx1, minX + rangeX - 1) - minX;
  const int r0 = max(tby0, minY) - minY, r1 = min(tby1, minY + rangeY - 1) - minY;
  if (c0 > c1 || r0 > r1) return true;
  const uint32_t mode = R.modeZ & 0xFFu;
  bool out[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double o = R.off[e], nx = R.nx[e], ny = R.ny[e];
    const double v00 = o + r0 * ny + c0 * nx, v01 = o + r0 * ny + c1 * nx;
    const double v10 = o + r1 * ny + c0 * nx, v11 = o + r1 * ny + c1 * nx;
    const double lo = fmin(fmin(v00, v01), fmin(v10, v11));
    const double hi = fmax(fmax(v00, v01), fmax(v10, v11));
    const double steps = double(r1 + c1);
    const double bound = fabs(o) + r1 * fabs(ny) + c1 * fabs(nx);
    const double err = steps * (bound * 0x1p-23 + 0x1p-126);
    const bool finite = isfinite(o) && isfinite(nx) && isfinite(ny) && isfinite(bound);
    const bool below31 = hi + err < 2147483648.0;
    out[e] = finite && lo - err >= 63.0 && (mode == kModeConvex || below31);
  }
  switch (mode) {
    case kModeConvex: return out[0] || out[1] || out[2] || out[3];
    case kModeTriangle0: return out[0] || out[1] || out[2];
    case kModeTriangle1: return out[0] || out[2] || out[3];
    case kModeConcaveRight: return out[1] || out[2] || (out[0] && out[3]);
    case kModeConcaveLeft: return out[0] || out[3] || (out[1] && out[2]);
    case kModeConcaveCenter: return (out[0] || out[1]) && (out[2] || out[3]);
    default: return false;
  }
}

// Tile-overlap tests shared by the prep kernel's tile count and the frame
// kernel (they must agree exactly).
template <int TBX, int TBY>
__device__ __forceinline__ bool query_overlaps_tile(uint32_t qMinX, uint32_t qMaxX, uint32_t qMinY, uint32_t qMaxY,
                                                   uint32_t tX, uint32_t tY) {
  return tX >= (qMinX >> 3) / TBX && tX <= (qMaxX >> 3) / TBX && tY >= (qMinY >> 3) / TBY &&
         tY <= (qMaxY >> 3) / TBY;
}

// Relevance of one item to a tile: bit 0 the query rectangle overlaps it,
// bit 1 some valid quad's blocks do; status in bits 2-3; 0 when neither or
// when the item is not drawn.
template <int TBX, int TBY>
__device__ __forceinline__ uint32_t tile_relevance(const ItemState& S, uint32_t tX, uint32_t tY, int tbx0, int tbx1,
                                                   int tby0, int tby1) {
  const uint32_t status = S.status;
  if (status == kStatusSkip) return 0u;
  const bool q = status == kStatusTest && query_overlaps_tile<TBX, TBY>(S.qMinX, S.qMaxX, S.qMinY, S.qMaxY, tX, tY);
  const bool f = S.fp[0] <= S.fp[1] && int(S.fp[0]) <= tbx1 && int(S.fp[1]) >= tbx0 && int(S.fp[2]) <= tby1 &&
                 int(S.fp[3]) >= tby0;
  return (q || f) ? ((q ? 1u : 0u) | (f ? 2u : 0u) | (status << 2)) : 0u;
}

// done[i] packs, over the tiles of item i's query rectangle, the number of
// tiles that have contributed (bits 0-15) and how many found a visible pixel
// (bits 16-31).
constexpr uint32_t kDoneCountMask = 0xFFFFu;
constexpr uint32_t kDoneVisibleOne = 0x10000u;

// One screen tile of one view: items [itemBegin, itemEnd) of the launch are
// walked in draw order against the tile's depth and hiZ held in LDS, then the
// tile is written back.  Shared by the single-view frame kernel (one
// work-group per tile, gated query items) and the multi-view persistent
// kernel (work-groups take (view, tile) units from a queue; items are never
// gated, so no tile waits on another).
struct TileJob {
  const ItemDesc* items;
  const ItemState* state;
  uint32_t itemBegin, itemEnd;
  int bX, bY;
  uint32_t tX, tY;
  const QuadRec* recs;
  const uint64_t* groupBox;
  uint16_t* depth;
  uint16_t* hiz;
  uint16_t* tileMin;
  uint32_t pyrX;
  uint32_t* done;
  uint32_t* err;
  uint64_t timeoutTicks;
  uint32_t* winner;  // labelled draws: per-pixel winners (block-linear), and
  uint32_t* labels;  // per-batch counts of written blocks each batch wins
  const uint32_t* list;  // optional: this tile's relevant items, in submission order
  uint32_t listLen;
};

// LABELS (8x8-block tiles): the draw tracks each pixel's winning batch, and
// once the tile is final every written block adds one to the label of each
// distinct batch that wins one of its pixels.
template <int TBX, int TBY, int NW, bool STATS, bool LABELS>
__device__ __forceinline__ void process_tile(const TileJob& J, TileSmem<TBX, TBY>& sm,
                                             const uint64_t* __restrict__ cov, bool& covLoaded,
                                             uint64_t (&ctr)[kFrameStatWords]) {
  static_assert(!LABELS || (TBX == 8 && TBY == 8), "labelled draws use the block-major draw");
  auto stamp = []() -> uint64_t { return STATS ? __builtin_amdgcn_s_memrealtime() : 0; };
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t tX = J.tX, tY = J.tY;
  const int bX = J.bX, bY = J.bY;
  const int tbx0 = int(tX) * TBX, tby0 = int(tY) * TBY;
  const int tbx1 = min(tbx0 + TBX, bX) - 1, tby1 = min(tby0 + TBY, bY) - 1;
  const float invBX = 1.0f / float(bX);

  auto relevance = [&](uint32_t i) -> uint32_t {
    return tile_relevance<TBX, TBY>(J.state[i], tX, tY, tbx0, tbx1, tby0, tby1);
  };
  // The tile's items: its binned list (the items relevant to it, in
  // submission order) when the launch has one, else the whole item range.
  const uint32_t nList = J.list ? J.listLen : J.itemEnd - J.itemBegin;
  auto item_at = [&](uint32_t pos) -> uint32_t { return J.list ? J.list[pos] : J.itemBegin + pos; };

  int involved = 0;
  constexpr uint32_t NT = 64u * NW;
  static_assert(NT >= kItemWindow, "the item window gathers one item per thread");
  if (J.list) involved = nList > 0u;
  else
    for (uint32_t i = J.itemBegin + tid; i < J.itemEnd && !involved; i += NT) involved = relevance(i) != 0u;
  if (!__syncthreads_or(involved)) return;

  const uint64_t tLoad = stamp();
  if (!covLoaded) {
    for (uint32_t i = tid; i < 4096; i += NT) sm.cov[i] = cov[i];
    covLoaded = true;
  }
  constexpr int NB = TBX * TBY;
  for (int rrow = int(tid); rrow < NB * 8; rrow += int(NT)) {
    const int lb = rrow >> 3, y = rrow & 7;
    const int gx = tbx0 + lb % TBX, gy = tby0 + lb / TBX;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (gx <= tbx1 && gy <= tby1)
      v = *reinterpret_cast<const uint4*>(J.depth + (size_t(gy) * bX + gx) * 64 + y * 8);
    *reinterpret_cast<uint4*>(&sm.depth[lb * 64 + y * 8]) = v;
  }
  for (int lb = int(tid); lb < NB; lb += int(NT)) {
    const int gx = tbx0 + lb % TBX, gy = tby0 + lb / TBX;
    sm.hiz[lb] = (gx <= tbx1 && gy <= tby1) ? J.hiz[size_t(gy) * bX + gx] : uint16_t(0xFFFF);
  }

  uint32_t qPar = 0;
  BlockCounters bc;
  // Early "not visible" answers for later query items of the window, tried
  // while the tile waits (see spec_query_wave), in query-gated launches only.
  const bool gated = J.done != nullptr;
  uint32_t drawVersion = 1, sPar = 0, wPar = 0, nRelW = 0;
  auto answered_early = [&](uint32_t k) -> bool { return (sm.spec[k >> 5] >> (k & 31u)) & 1u; };
  auto next_early = [&](uint32_t from) -> uint32_t {
    for (uint32_t c = from; c < nRelW; ++c)
      if ((sm.items[c].flags & 1u) && !answered_early(c) && sm.tried[c] != drawVersion) return c;
    return nRelW;
  };
  auto try_early = [&](uint32_t c) {
    const ItemLocal& L = sm.items[c];
    const uint32_t r = spec_query_wave<TBX, TBY, NW>(L, sm, tbx0, tbx1, tby0, tby1, wave, lane);
    if (lane == 0) sm.spart[sPar][wave] = r;
    __syncthreads();
    if (tid == 0) {
      uint32_t any = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) any |= sm.spart[sPar][w];
      if (any == 0u) {
        __hip_atomic_fetch_add(&J.done[L.index], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        sm.spec[c >> 5] |= 1u << (c & 31u);
      } else {
        sm.tried[c] = drawVersion;
      }
      if (STATS) {
        ctr[kStatEarlyTried]++;
        if (any == 0u) ctr[kStatEarlyAnswered]++;
      }
    }
    sPar ^= 1u;
  };
  if (STATS) {
    if (tid < 4) sm.blockStats[tid] = 0;
    __syncthreads();
    if (tid == 0) ctr[kStatLoadTicks] += stamp() - tLoad;
  }
  const uint64_t tLoop = stamp();
  for (uint32_t wbase = 0; wbase < nList;) {
    // Gather this window's items that touch the tile, in draw order: every
    // thread tests one candidate, the first kItemWindow relevant ones are
    // kept, and the next window starts at the first relevant candidate not
    // kept (or after the last candidate tested).
    const bool inWin = wbase + tid < nList;
    const uint32_t i = inWin ? item_at(wbase + tid) : 0u;
    const uint32_t rel = inWin ? relevance(i) : 0u;
    const uint64_t m = __ballot(rel != 0u);
    const uint32_t pre = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
    if (lane == 0) sm.wcnt[wave] = uint32_t(__popcll(m));
    if (tid == 0) sm.wnext = min(wbase + NT, nList);
    __syncthreads();
    uint32_t off = 0, nRel = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
      off += w < wave ? sm.wcnt[w] : 0u;
      nRel += sm.wcnt[w];
    }
    if (rel && off + pre == kItemWindow) sm.wnext = wbase + tid;
    nRel = min(nRel, kItemWindow);
    if (rel && off + pre < kItemWindow) {
      const ItemState& S = J.state[i];
      ItemLocal& L = sm.items[off + pre];
      L.index = i;
      L.flags = rel;
      L.qMinX = S.qMinX;
      L.qMaxX = S.qMaxX;
      L.qMinY = S.qMinY;
      L.qMaxY = S.qMaxY;
      L.qMaxZ = S.qMaxZ;
      L.expected = S.expected;
      const uint32_t qb = J.items[i].quadBase, qc = J.items[i].quadCount;
      L.quadBase = qb;
      L.quadCount = qc;
      uint64_t groups = ~0ull;
      if ((rel & 2u) && qc > 0) {
        const uint32_t g0 = qb >> 6, g1 = (qb + qc - 1) >> 6;
        if (g1 - g0 < 64) {
          groups = 0;
          for (uint32_t gi = g0; gi <= g1; ++gi)
            if (box_touches(J.groupBox[gi], tbx0, tbx1, tby0, tby1)) groups |= 1ull << (gi - g0);
        }
      }
      L.groups = groups;
    }
    if (gated) {
      if (tid < kItemWindow / 32) sm.spec[tid] = 0u;
      if (tid < kItemWindow) sm.tried[tid] = 0u;
    }
    __syncthreads();
    nRelW = nRel;
    const uint32_t wnext = sm.wnext;
    // Later query items are answered early only while the tile waits: a
    // pass over the window's query items at its start delays the tiles with
    // the longest lists, which other tiles' waits depend on.

    for (uint32_t k = 0; k < nRel; ++k) {
      const ItemLocal& L = sm.items[k];
      const uint32_t item = L.index, flags = L.flags, status = flags >> 2;
      if (STATS && tid == 0) ctr[kStatItems]++;
      if ((flags & 1u) && !(gated && answered_early(k))) {
        // This tile's share of query2D for the item, at its exact turn.
        const uint64_t tq = stamp();
        const bool found = query_wave<TBX, TBY, NW>(L, sm, tbx0, tbx1, tby0, tby1, wave, lane);
        if (lane == 0) sm.part[qPar][wave] = found ? 1u : 0u;
        __syncthreads();
        if (tid == 0) {
          uint32_t part = 0;
#pragma unroll
          for (int w = 0; w < NW; ++w) part |= sm.part[qPar][w];
          __hip_atomic_fetch_add(&J.done[item], 1u + (part ? kDoneVisibleOne : 0u), __ATOMIC_RELEASE,
                                 __HIP_MEMORY_SCOPE_AGENT);
        }
        qPar ^= 1u;
        if (STATS) {
          ctr[kStatQueryTicks] += stamp() - tq;
          ctr[kStatQueries]++;
        }
      }
      if (!(flags & 2u)) continue;
      if (status == kStatusTest) {
        // Wait until every tile of the rectangle has answered, or one of them
        // has already seen a visible pixel; meanwhile answer later items of
        // the window early where the tile's state already proves them.
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        uint32_t c = k + 1;
        bool vis;
        while (true) {
          if (tid == 0) {
            const uint32_t d = __hip_atomic_load(&J.done[item], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t st = 0;
            if ((d & kDoneCountMask) >= L.expected || d >= kDoneVisibleOne) {
              st = d >= kDoneVisibleOne ? 2u : 1u;
            } else if (__builtin_amdgcn_s_memrealtime() - t0 > J.timeoutTicks ||
                       __hip_atomic_load(J.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
              // After one wait of the launch has timed out, the others stop
              // waiting too (the launch's result is reported as an error).
              __hip_atomic_fetch_or(J.err, uint32_t(kErrTimeout), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              st = d >= kDoneVisibleOne ? 2u : 1u;
            }
            sm.wstate[wPar] = st;
          }
          __syncthreads();
          const uint32_t st = sm.wstate[wPar];
          wPar ^= 1u;
          if (st != 0u) {
            vis = st == 2u;
            break;
          }
          c = next_early(c);
          if (c < nRel) {
            try_early(c);
            ++c;
          } else if (tid == 0) {
            __builtin_amdgcn_s_sleep(1);
          }
        }
        if (STATS && tid == 0) {
          ctr[kStatWaitTicks] += stamp() - t0;
          ctr[kStatWaits]++;
        }
        if (!vis) continue;
      }
      // Draw the item's quads that touch the tile, in submission order: each
      // pass tests 64 * NW quads (one per thread; passes and waves whose
      // 64-quad groups miss the tile are skipped); hits are staged into LDS
      // and drawn 64 at a time.
      const uint64_t tr = stamp();
      const uint32_t gBase = L.quadBase >> 6;
      for (uint32_t base = 0; base < L.quadCount; base += NT) {
        // groups overlapping this pass, relative to gBase
        const uint32_t pg0 = ((L.quadBase + base) >> 6) - gBase;
        const uint32_t pg1 = ((L.quadBase + min(base + NT, L.quadCount) - 1) >> 6) - gBase;
        const uint64_t passMask = (pg1 - pg0 >= 63) ? ~0ull : (((2ull << (pg1 - pg0)) - 1) << pg0);
        if (pg0 < 64 && (L.groups & passMask) == 0) continue;
        const uint64_t tScan = stamp();
        if (STATS && tid == 0) ctr[kStatPasses]++;
        const uint32_t waveFirst = base + wave * 64u;
        bool hit = false;
        QuadRec rec;
        if (waveFirst < L.quadCount) {
          const uint32_t wg0 = ((L.quadBase + waveFirst) >> 6) - gBase;
          const uint32_t wg1 = ((L.quadBase + min(waveFirst + 63u, L.quadCount - 1u)) >> 6) - gBase;
          const bool groupHit = wg1 >= 64 || ((L.groups >> wg0) & 1ull) || ((L.groups >> wg1) & 1ull);
          if (groupHit && base + tid < L.quadCount) {
            rec = J.recs[L.quadBase + base + tid];
            const uint32_t rx = rec.range & 0xFFFFu, ry = rec.range >> 16;
            hit = rx != 0u && box_touches(quad_box(rec.fb, rx, ry, uint32_t(bX), uint32_t(bY)), tbx0, tbx1, tby0, tby1) &&
                  !quad_misses_tile(rec, tbx0, tbx1, tby0, tby1, bX);
          }
        }
        const uint64_t hm = __ballot(hit);
        const uint32_t hpre =
            __builtin_amdgcn_mbcnt_hi(uint32_t(hm >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(hm), 0u));
        if (lane == 0) sm.wcnt[wave] = uint32_t(__popcll(hm));
        __syncthreads();
        uint32_t hoff = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < NW; ++w) {
          hoff += w < wave ? sm.wcnt[w] : 0u;
          total += sm.wcnt[w];
        }
        const uint32_t pos = hoff + hpre;
        if (STATS && tid == 0) ctr[kStatScanTicks] += stamp() - tScan;
        // Hits reload their record (just read, so from cache) rather than
        // hold it across the compaction, which made the compiler spill it.
        if (hit && pos < kQuadChunk) {
          sm.recs[pos] = J.recs[L.quadBase + base + tid];
          if (LABELS) sm.qid[pos] = L.quadBase + base + tid;
        }
        for (uint32_t k0 = 0; k0 < total; k0 += kQuadChunk) {
          const uint32_t n = min(kQuadChunk, total - k0);
          if (k0 > 0 && hit && pos >= k0 && pos < k0 + n) {
            sm.recs[pos - k0] = J.recs[L.quadBase + base + tid];
            if (LABELS) sm.qid[pos - k0] = L.quadBase + base + tid;
          }
          __syncthreads();
          const uint64_t tDraw = stamp();
          uint64_t tB = 0;
          if constexpr (TBX == 8 && TBY == 8) {
            tB = draw_chunk_block_major<TBX, TBY, NW, STATS, LABELS>(sm, n, tbx0, tbx1, tby0, tby1, bX, invBX, wave,
                                                                     lane, bc, J.winner);
          } else {
            for (uint32_t q = 0; q < n; ++q)
              raster_quad_wave<TBX, TBY, NW, STATS>(sm.recs[q], sm, tbx0, tbx1, tby0, tby1, bX, invBX, wave, lane, bc);
          }
          if (STATS && tid == 0) ctr[kStatQuads] += n;
          __syncthreads();
          if (STATS && tid == 0) {
            const uint64_t t = stamp();
            ctr[kStatDrawTicks] += t - tDraw;
            ctr[kStatChunks]++;
            if (TBX == 8 && TBY == 8) {
              ctr[kStatPhaseATicks] += tB - tDraw;
              ctr[kStatPhaseBTicks] += t - tB;
            }
          }
        }
      }
      if (STATS && tid == 0) ctr[kStatRasterTicks] += stamp() - tr;
      ++drawVersion;  // the tile's state changed: early answers that failed may now succeed
    }
    __syncthreads();
    wbase = wnext;
  }

  if (STATS) {
    if (tid == 0) ctr[kStatLoopTicks] += stamp() - tLoop;
    if (lane == 0) {
      atomicAdd(&sm.blockStats[0], bc.visited);
      atomicAdd(&sm.blockStats[1], bc.rejected);
      atomicAdd(&sm.blockStats[2], bc.written);
      atomicAdd(&sm.blockStats[3], bc.foldSteps);
    }
    __syncthreads();
    if (tid == 0) {
      ctr[kStatBlocksVisited] += sm.blockStats[0];
      ctr[kStatBlocksRejected] += sm.blockStats[1];
      ctr[kStatBlocksWritten] += sm.blockStats[2];
      ctr[kStatFoldSteps] += sm.blockStats[3];
    }
  }
  if constexpr (LABELS) {
    // Each block is counted by the wave that drew it, which wrote its
    // winners; blocks still cleared were not written this frame.
    for (int lb = 0; lb < NB; ++lb) {
      const int lbx = lb % TBX, lby = lb / TBX;
      if (block_owner<NW>(lbx, lby) != wave) continue;
      const int gx = tbx0 + lbx, gy = tby0 + lby;
      if (gx > tbx1 || gy > tby1 || sm.hiz[lb] == 1u) continue;
      const uint32_t w = J.winner[(size_t(gy) * size_t(bX) + size_t(gx)) * 64u + lane];
      bool active = w != kNoWinner;
      while (true) {
        const uint64_t m = __ballot(active);
        if (m == 0ull) break;
        const uint32_t v = uint32_t(__builtin_amdgcn_readlane(int(w), int(__builtin_ctzll(m))));
        if (active && w == v) active = false;
        if (lane == 0) atomicAdd(&J.labels[v], 1u);
      }
    }
  }
  // Write the tile back.
  for (int rrow = int(tid); rrow < NB * 8; rrow += int(NT)) {
    const int lb = rrow >> 3, y = rrow & 7;
    const int gx = tbx0 + lb % TBX, gy = tby0 + lb / TBX;
    if (gx <= tbx1 && gy <= tby1)
      *reinterpret_cast<uint4*>(J.depth + (size_t(gy) * bX + gx) * 64 + y * 8) =
          *reinterpret_cast<const uint4*>(&sm.depth[lb * 64 + y * 8]);
  }
  for (int lb = int(tid); lb < NB; lb += int(NT)) {
    const int gx = tbx0 + lb % TBX, gy = tby0 + lb / TBX;
    if (gx <= tbx1 && gy <= tby1) J.hiz[size_t(gy) * bX + gx] = sm.hiz[lb];
  }
  // Min hiZ of each 8x8-block sub-tile for the query kernel (blocks outside
  // the buffer hold 0xFFFF and do not lower it).
  constexpr int kSubX = TBX / 8, kSubY = TBY / 8;
  for (int sub = int(wave); sub < kSubX * kSubY; sub += NW) {
    const int sx = sub % kSubX, sy = sub / kSubX;
    const int lb = (sy * 8 + int(lane >> 3)) * TBX + sx * 8 + int(lane & 7u);
    const uint32_t mn = __ockl_wfred_min_u32(uint32_t(sm.hiz[lb]));
    const uint32_t px = tX * kSubX + uint32_t(sx), py = tY * kSubY + uint32_t(sy);
    if (lane == 0 && px < J.pyrX && py * 8 < uint32_t(bY)) J.tileMin[py * J.pyrX + px] = uint16_t(mn);
  }
}

// Two work-groups per CU, as the multi-view kernel (at most 128 VGPRs with
// eight waves per group).
template <int TBX, int TBY, int NW, bool STATS>
__global__ void __launch_bounds__(64 * NW, NW == 8 ? 4 : 2) rz_frame_kernel(FrameParams p) {
  __shared__ TileSmem<TBX, TBY> sm;
  // Observability counters, kept in LDS by thread 0 (unused unless STATS).
  __shared__ uint64_t ctr[kFrameStatWords];
  if (STATS && threadIdx.x < kFrameStatWords) ctr[threadIdx.x] = 0;
  const uint64_t tEntry = STATS ? __builtin_amdgcn_s_memrealtime() : 0;
  TileJob J;
  J.items = p.items;
  J.state = p.state;
  J.itemBegin = 0;
  J.itemEnd = p.nItems;
  J.bX = int(p.bX);
  J.bY = int(p.bY);
  J.tX = blockIdx.x % p.tilesX;
  J.tY = blockIdx.x / p.tilesX;
  J.recs = p.recs;
  J.groupBox = p.groupBox;
  J.depth = p.depth;
  J.hiz = p.hiz;
  J.tileMin = p.tileMin;
  J.pyrX = p.pyrX;
  J.done = p.done;
  J.err = p.err;
  J.timeoutTicks = p.timeoutTicks;
  J.winner = nullptr;
  J.labels = nullptr;
  J.list = nullptr;
  J.listLen = 0;
  bool covLoaded = false;
  process_tile<TBX, TBY, NW, STATS, false>(J, sm, p.cov, covLoaded, ctr);
  if (STATS && threadIdx.x == 0) {
    ctr[kStatTotalTicks] = __builtin_amdgcn_s_memrealtime() - tEntry;
    for (int k = 0; k < kFrameStatWords; ++k) p.stats[size_t(blockIdx.x) * kFrameStatWords + k] = ctr[k];
  }
}

// Multi-view persistent raster: each work-group takes (view, tile) units from
// p.queue until the queue is exhausted.  Units are handed out in view order.
// Classified items never wait.  Query-gated items (p.done) wait only on
// other tiles of the same view; when every view's tiles fit in the grid and
// the grid is co-resident (the host checks both), the earliest unfinished
// view always has all its tiles running, so every wait ends (and each is
// bounded by p.timeoutTicks in any case).
// Two work-groups per CU (the LDS allows two): with eight waves each, every
// SIMD must hold four waves, i.e. at most 128 VGPRs per lane.  NW is 2, 4 or
// 8 (the item window gathers one item per thread of its first 128).
template <int TBX, int TBY, int NW, bool STATS, bool LABELS>
__global__ void __launch_bounds__(64 * NW, NW == 8 ? 4 : 2) rz_views_kernel(ViewsParams p) {
  __shared__ TileSmem<TBX, TBY> sm;
  __shared__ uint32_t unitSh;
  __shared__ uint64_t ctr[kFrameStatWords];
  if (STATS && threadIdx.x < kFrameStatWords) ctr[threadIdx.x] = 0;
  const uint64_t tEntry = STATS ? __builtin_amdgcn_s_memrealtime() : 0;
  bool covLoaded = false;
  // With a compacted unit list, only units that have items are handed out,
  // each carrying its view; else every unit, its view found by search.
  const uint32_t nHanded = p.unitList ? *p.nActive : p.nUnits;
  while (true) {
    if (threadIdx.x == 0) {
      const uint32_t u = __hip_atomic_fetch_add(p.queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      unitSh = u >= nHanded ? 0xFFFFFFFFu : p.unitList ? p.unitList[u] : u;
    }
    __syncthreads();
    const uint32_t entry = unitSh;
    __syncthreads();
    if (entry == 0xFFFFFFFFu) break;
    uint32_t lo = 0, t = 0;
    if (p.unitList) {
      lo = entry >> 20;
      t = entry & 0xFFFFFu;
    } else {
      uint32_t hi = p.nViews - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (p.views[mid].unitBase <= entry) lo = mid;
        else hi = mid - 1;
      }
      t = entry - p.views[lo].unitBase;
    }
    const ViewDesc& V = p.views[lo];
    const uint32_t unit = V.unitBase + t;
    TileJob J;
    J.items = p.items;
    J.state = p.state;
    J.itemBegin = p.batchItems ? V.batchBase : V.itemBase;
    J.itemEnd = p.batchItems ? V.batchBase + V.nBatches : V.itemBase + V.nItems;
    J.bX = int(V.bX);
    J.bY = int(V.bY);
    J.tX = t % V.tilesX;
    J.tY = t / V.tilesX;
    J.recs = p.recs;
    J.groupBox = p.groupBox;
    J.depth = V.depth;
    J.hiz = V.hiz;
    J.tileMin = V.tileMin;
    J.pyrX = V.pyrX;
    J.done = p.done;
    J.err = p.err;
    J.timeoutTicks = p.timeoutTicks;
    J.winner = V.winner;
    J.labels = p.labels;
    J.list = p.lists ? p.lists + V.listBase + size_t(t) * (J.itemEnd - J.itemBegin) : nullptr;
    J.listLen = p.lists ? p.listLen[unit] : 0u;
    process_tile<TBX, TBY, NW, STATS, LABELS>(J, sm, p.cov, covLoaded, ctr);
    __syncthreads();
  }
  if (STATS) {
    // Words 0-2: block-loop counters; 3 onwards: the per-tile counters of
    // every unit summed, word 3 + kStatTotalTicks being the work-groups'
    // lifetimes.
    __syncthreads();
    if (threadIdx.x == 0) ctr[kStatTotalTicks] = __builtin_amdgcn_s_memrealtime() - tEntry;
    __syncthreads();
    if (threadIdx.x < kFrameStatWords)
      atomicAdd(&p.stats[3 + threadIdx.x], (unsigned long long)ctr[threadIdx.x]);
    if (threadIdx.x == 0) {
      atomicAdd(&p.stats[0], (unsigned long long)ctr[kStatBlocksVisited]);
      atomicAdd(&p.stats[1], (unsigned long long)ctr[kStatBlocksRejected]);
      atomicAdd(&p.stats[2], (unsigned long long)ctr[kStatBlocksWritten]);
    }
  }
}

// The units of one view (one work-group per view) that have items, in tile
// order, as view << 20 | tile, at unitTmp[view's unitBase ...], and their
// count in p.unitCounts[view].
__global__ void __launch_bounds__(256) rz_views_compact_kernel(ViewsParams p, uint32_t* __restrict__ unitTmp) {
  __shared__ uint32_t wcnt[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t v = blockIdx.x;
  const ViewDesc& V = p.views[v];
  // Units with more items first (buckets of floor(log2(items)), 8 and up
  // together at the top), tile order within a bucket: the tiles that the
  // others' query waits depend on most start first.  Any order within a view
  // keeps every wait finite (all of the earliest view's tiles run).
  uint32_t n = 0;
  for (int bucket = 7; bucket >= 0; --bucket) {
    for (uint32_t base = 0; base < V.nUnits; base += 256u) {
      const uint32_t t = base + tid;
      const uint32_t len = t < V.nUnits ? p.listLen[V.unitBase + t] : 0u;
      const int lg = len ? min(31 - __builtin_clz(len), 7) : -1;
      const bool has = lg == bucket;
      const uint64_t m = __ballot(has);
      const uint32_t pre = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
      if (lane == 0) wcnt[wave] = uint32_t(__popcll(m));
      __syncthreads();
      uint32_t off = n, total = 0;
#pragma unroll
      for (uint32_t w = 0; w < 4; ++w) {
        off += w < wave ? wcnt[w] : 0u;
        total += wcnt[w];
      }
      if (has) unitTmp[V.unitBase + off + pre] = (v << 20) | t;
      n += total;
      __syncthreads();
    }
  }
  if (tid == 0) p.unitCounts[v] = n;
}

// Concatenates the views' compacted units in view order into p.unitList
// (one work-group per view); the last view also writes the total.
__global__ void __launch_bounds__(256) rz_views_gather_kernel(ViewsParams p, const uint32_t* __restrict__ unitTmp) {
  const uint32_t v = blockIdx.x;
  uint32_t off = 0;
  for (uint32_t w = 0; w < v; ++w) off += p.unitCounts[w];
  const uint32_t n = p.unitCounts[v];
  const uint32_t src = p.views[v].unitBase;
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) p.unitList[off + k] = unitTmp[src + k];
  if (v == p.nViews - 1 && threadIdx.x == 0) *p.nActive = off + n;
}

// Per item of a multi-view launch, the box in 8x8-block tiles (u16 x0, x1,
// y0, y1) that bounds every tile it is relevant to: the union of its
// footprint's tiles and, for a query-gated item, its query rectangle's;
// empty (x0 > x1) when the item is not drawn.
__global__ void __launch_bounds__(256) rz_item_tiles_kernel(ViewsParams p, uint32_t n, uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ItemState& S = p.state[i];
  // Tiles of tileB blocks: block >> sh, pixel >> (sh + 3).
  const uint32_t sh = p.views[p.items[i].view].tileB == 16u ? 4u : 3u;
  uint32_t x0 = 0xFFFFu, x1 = 0u, y0 = 0xFFFFu, y1 = 0u;
  if (S.status != kStatusSkip) {
    if (S.fp[0] <= S.fp[1]) {
      x0 = S.fp[0] >> sh;
      x1 = S.fp[1] >> sh;
      y0 = S.fp[2] >> sh;
      y1 = S.fp[3] >> sh;
    }
    if (S.status == kStatusTest) {
      x0 = min(x0, S.qMinX >> (sh + 3u));
      x1 = max(x1, S.qMaxX >> (sh + 3u));
      y0 = min(y0, S.qMinY >> (sh + 3u));
      y1 = max(y1, S.qMaxY >> (sh + 3u));
    }
  }
  out[i] = uint64_t(min(x0, 0xFFFFu)) | uint64_t(min(x1, 0xFFFFu)) << 16 | uint64_t(min(y0, 0xFFFFu)) << 32 |
           uint64_t(min(y1, 0xFFFFu)) << 48;
}

// Per-tile item lists of a multi-view launch: for each (view, tile) unit, the
// indices of the view's items whose tile box (rz_item_tiles_kernel) holds the
// tile, in submission order, at lists[views[v].listBase + tile * (items in
// view)], and their count in listLen[unit].  A box can hold a tile the item is
// not relevant to; the raster kernel's exact test drops those.  One wave per
// unit tests the items 64 at a time and compacts them with ballots (no
// barriers: a view's items are about a hundred, so a work-group per unit
// spent its time being scheduled).
__global__ void __launch_bounds__(256) rz_views_bin_kernel(ViewsParams p, const uint64_t* __restrict__ tileBox) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = uint32_t(__builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)));
  for (uint32_t unit = blockIdx.x * 4u + wave; unit < p.nUnits; unit += gridDim.x * 4u) {
    uint32_t lo = 0, hi = p.nViews - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (p.views[mid].unitBase <= unit) lo = mid;
      else hi = mid - 1;
    }
    const ViewDesc& V = p.views[lo];
    const uint32_t t = unit - V.unitBase;
    const uint32_t tX = t % V.tilesX, tY = t / V.tilesX;
    const uint32_t begin = p.batchItems ? V.batchBase : V.itemBase;
    const uint32_t end = begin + (p.batchItems ? V.nBatches : V.nItems);
    uint32_t* out = p.lists + V.listBase + size_t(t) * (end - begin);
    uint32_t n = 0;
    for (uint32_t base = begin; base < end; base += 64u) {
      const uint32_t i = base + lane;
      bool rel = false;
      if (i < end) {
        const uint64_t b = tileBox[i];
        rel = uint32_t(b & 0xFFFF) <= tX && tX <= uint32_t((b >> 16) & 0xFFFF) && uint32_t((b >> 32) & 0xFFFF) <= tY &&
              tY <= uint32_t(b >> 48);
      }
      const uint64_t m = __ballot(rel);
      const uint32_t pre = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
      if (rel) out[n + pre] = i;
      n += uint32_t(__popcll(m));
    }
    if (lane == 0) p.listLen[unit] = n;
  }
}

hipError_t launch_views_bin(const ViewsParams& p, uint32_t nItems, uint64_t* tileBox, hipStream_t stream) {
  if (p.nUnits == 0 || !p.lists) return hipSuccess;
  if (nItems) hipLaunchKernelGGL(rz_item_tiles_kernel, dim3((nItems + 255) / 256), dim3(256), 0, stream, p, nItems, tileBox);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(rz_views_bin_kernel, dim3(std::min<uint32_t>((p.nUnits + 3u) / 4u, 65535u)), dim3(256), 0,
                     stream, p, tileBox);
  return hipGetLastError();
}

hipError_t launch_views_compact(const ViewsParams& p, uint32_t* unitTmp, hipStream_t stream) {
  if (p.nUnits == 0 || !p.unitList) return hipSuccess;
  hipLaunchKernelGGL(rz_views_compact_kernel, dim3(p.nViews), dim3(256), 0, stream, p, unitTmp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(rz_views_gather_kernel, dim3(p.nViews), dim3(256), 0, stream, p, static_cast<const uint32_t*>(unitTmp));
  return hipGetLastError();
}

// Per-item draw status of a frame: 0 not drawn, 1 rasterize<false>,
// 2 rasterize<true>.
__global__ void rz_status_kernel(const ItemState* __restrict__ state, const uint32_t* __restrict__ visible,
                                 uint32_t n, uint8_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t st = state[i].status;
  out[i] = uint8_t(st == kStatusTest ? ((visible[i] >> 16) != 0u ? 1u : 0u) : st);
}

// ---------------------------------------------------------------------------
// Queries, readback, clear.
// ---------------------------------------------------------------------------
// queryVisibility (Rasterizer.cpp:123-281), one wave per box.
// queryVisibility (Rasterizer.cpp:123-281) of a batch: lanes 0..BPW-1 of a
// wave each take one box for the depth-independent part (projection,
// frustum, near plane, rectangle), and the wave then runs query2D for its
// boxes that need it, one after another.  A frame's 65,536 boxes are too few
// to give each wave 64 of them (one wave per SIMD, long serial query2D
// chains), and too many for one wave per box (the wave launches dominate):
// four boxes per wave take 20 us per 65,536-box city batch, one 33 us, eight
// 28 us.
template <uint32_t BPW>
__global__ void __launch_bounds__(256) rz_query_kernel(QueryParams p) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t q = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * BPW + lane;
  const bool valid = lane < BPW && q < p.n;
  QueryPre Q = {0, 0, 0, 0, 0, 0};
  if (valid) {
    const float* b = p.bounds + 8 * size_t(q);
    const float bmin[4] = {b[0], b[1], b[2], b[3]};
    const float bmax[4] = {b[4], b[5], b[6], b[7]};
    Q = query_pre(p.mvp, bmin, bmax, p.width, p.height, p.rcp);
  }
  uint32_t vis = Q.kind == 1 ? 1u : 0u;
  const uint32_t clip = vis;
  uint64_t m = __ballot(valid && Q.kind == 3);
  while (m) {
    const int j = int(__builtin_ctzll(m));
    m &= m - 1ull;
    const uint32_t minX = uint32_t(__builtin_amdgcn_readlane(int(Q.minX), j));
    const uint32_t maxX = uint32_t(__builtin_amdgcn_readlane(int(Q.maxX), j));
    const uint32_t minY = uint32_t(__builtin_amdgcn_readlane(int(Q.minY), j));
    const uint32_t maxY = uint32_t(__builtin_amdgcn_readlane(int(Q.maxY), j));
    const uint32_t maxZ = uint32_t(__builtin_amdgcn_readlane(int(Q.maxZ), j));
    const bool v = query2d_wave(minX, maxX, minY, maxY, maxZ, p.depth, p.hiz, p.tileMin, p.bX, p.pyrX, lane);
    if (int(lane) == j) vis = v ? 1u : 0u;
  }
  if (valid) {
    p.visible[q] = uint8_t(vis);
    if (p.needsClipping) p.needsClipping[q] = uint8_t(clip);
    // bit 0 visible, bit 1 needsClipping, bit 2 needsClipping was written
    if (p.code) p.code[q] = uint8_t(vis | (clip << 1) | (Q.kind != 0 ? 4u : 0u));
  }
}

// queryVisibility over the queries of every view of a multi-view launch;
// blockIdx.y is the view.  A step holds many boxes, so one lane per box does
// the depth-independent part (projection, frustum, near plane, rectangle)
// and the wave then runs query2D for its boxes that need it.  (The
// single-view kernel gives each wave four boxes: its batches are too small
// to fill the GPU with 64 boxes per wave.)
__global__ void __launch_bounds__(256) rz_views_query_kernel(ViewsQueryParams p) {
  // The raster launch that sets the error word has completed.
  if (p.errOut && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *p.errOut = *p.err;
  const uint32_t lane = threadIdx.x & 63u;
  const ViewDesc& V = p.views[blockIdx.y];
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;  // 64 boxes per wave
  const bool valid = k < V.nQueries;
  QueryPre Q = {0, 0, 0, 0, 0, 0};
  if (valid) {
    const float* b = V.queryBounds + 8 * size_t(k);
    const float bmin[4] = {b[0], b[1], b[2], b[3]};
    const float bmax[4] = {b[4], b[5], b[6], b[7]};
    Q = query_pre(V.mvp, bmin, bmax, V.width, V.height, p.rcp);
  }
  uint32_t vis = Q.kind == 1 ? 1u : 0u;
  const uint32_t clip = vis;
  // The boxes that need query2D, one after another, each by the whole wave.
  uint64_t m = __ballot(valid && Q.kind == 3);
  while (m) {
    const int j = int(__builtin_ctzll(m));
    m &= m - 1ull;
    const uint32_t minX = uint32_t(__builtin_amdgcn_readlane(int(Q.minX), j));
    const uint32_t maxX = uint32_t(__builtin_amdgcn_readlane(int(Q.maxX), j));
    const uint32_t minY = uint32_t(__builtin_amdgcn_readlane(int(Q.minY), j));
    const uint32_t maxY = uint32_t(__builtin_amdgcn_readlane(int(Q.maxY), j));
    const uint32_t maxZ = uint32_t(__builtin_amdgcn_readlane(int(Q.maxZ), j));
    const bool v = query2d_wave(minX, maxX, minY, maxY, maxZ, V.depth, V.hiz, V.tileMin, V.bX, V.pyrX, lane);
    if (int(lane) == j) vis = v ? 1u : 0u;
  }
  if (valid) {
    V.queryVisible[k] = uint8_t(vis);
    V.queryClip[k] = uint8_t(clip);
  }
}

// clear (Rasterizer.cpp:107-121) of every view (blockIdx.y): hiZ := 1, the
// cleared mark, eight entries per 16-byte store (each view's hiZ buffer is its
// own allocation, holding whole groups of eight), and the per-tile min hiZ
// := 1.  The view table comes from `src`, the host's pinned copy with this
// step's matrices, read through its device mapping; the first work-group of
// each view copies the view's entry to `views`, which every later kernel of
// the step reads (no copy before the launch).
__global__ void rz_views_clear_kernel(const ViewDesc* __restrict__ src, ViewDesc* __restrict__ views) {
  const ViewDesc& V = src[blockIdx.y];
  if (blockIdx.x == 0 && threadIdx.x < sizeof(ViewDesc) / 16)
    reinterpret_cast<uint4*>(views + blockIdx.y)[threadIdx.x] =
        reinterpret_cast<const uint4*>(src + blockIdx.y)[threadIdx.x];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (V.bX * V.bY + 8u) / 8u) reinterpret_cast<uint4*>(V.hiz)[i] = make_uint4(0x10001u, 0x10001u, 0x10001u, 0x10001u);
  if (i < V.pyrX * V.pyrY) V.tileMin[i] = 1;
}

__global__ void rz_query2d_kernel(uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY, uint32_t maxZ,
                                  QueryParams p, uint8_t* out) {
  const bool v = query2d_wave(minX, maxX, minY, maxY, maxZ, p.depth, p.hiz, p.tileMin, p.bX, p.pyrX, threadIdx.x);
  if (threadIdx.x == 0) *out = v ? 1 : 0;
}

// Min of hiZ over each 8x8-block tile, one wave per tile.
__global__ void rz_tile_min_kernel(const uint16_t* __restrict__ hiz, uint32_t bX, uint32_t bY,
                                   uint16_t* __restrict__ tileMin, uint32_t pyrX, uint32_t pyrY) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (t >= pyrX * pyrY) return;
  const uint32_t bx = (t % pyrX) * 8 + (lane & 7u), by = (t / pyrX) * 8 + (lane >> 3);
  const uint32_t v = (bx < bX && by < bY) ? uint32_t(hiz[by * bX + bx]) : 0xFFFFu;
  const uint32_t mn = __ockl_wfred_min_u32(v);
  if (lane == 0) tileMin[t] = uint16_t(mn);
}

// readBackDepth (Rasterizer.cpp:351-399), one pixel per thread.
__global__ void __launch_bounds__(256) rz_readback_kernel(const uint16_t* __restrict__ depth,
                                                          const uint16_t* __restrict__ hiz, uint32_t width,
                                                          uint32_t height, uint32_t bX, uint32_t* __restrict__ rgba) {
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t py = blockIdx.y;
  if (px >= width || py >= height) return;
  const uint32_t b = (py >> 3) * bX + (px >> 3);
  uint32_t out = 0;
  if (hiz[b] != 1u) {
    const uint32_t v = depth[size_t(b) * 64 + (py & 7u) * 8 + (px & 7u)];
    const float f = bitsf(v << 12) * 3.9623753e+28f;
    const float l = 0.5f / (1000.25f - (1.0f - f) * 999.75f);
    const uint32_t d = uint32_t(25600.0f * l);
    out = (d / 100u & 0xFFu) | ((d % 256u) << 8) | 0xFF000000u;
  }
  rgba[size_t(py) * width + px] = out;
}

// clear (Rasterizer.cpp:107-121) of a single view: hiZ and the per-tile min
// hiZ := 1 (the cleared mark), both arrays in one launch; hiZ (whole groups
// of eight entries, in its own allocation) eight entries per 16-byte store.
__global__ void rz_clear_kernel(uint4* hiz, uint32_t nHiz8, uint16_t* tileMin, uint32_t nMin) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nHiz8) hiz[i] = make_uint4(0x10001u, 0x10001u, 0x10001u, 0x10001u);
  else if (i - nHiz8 < nMin) tileMin[i - nHiz8] = 1;
}

// ---------------------------------------------------------------------------
// Draw-order labels and batch features.
// ---------------------------------------------------------------------------
// Item whose quad range holds quad g (items are in quad order).
// 16 features per batch, from its quads' setup records and packets and the
// labels of the previous step:
//  0 drawn quads / 8          1 log2(1 + union block-box area)
//  2 min packed maxZ / 65535  3 max packed maxZ / 65535   4 mean packed maxZ / 65535
//  5 log2(1 + sum of block-box areas)  6 drawn with rasterize<true>
//  7, 8 union box centre / buffer size (x, y)   9 log2(1 + |W| of the batch centroid)
// 10 log2(1 + previous step's label)  11-13 mean, min, max of (unit quad normal . view direction)
// 14 position of the batch in the view's plain order   15 fraction of Convex quads
// Writes batch b's features; returns whether its occluder is drawn.
__device__ bool batch_features(const LearnParams& p, uint32_t b) {
  const uint32_t item = item_of_quad(p.items, p.nItems, b * 8u, p.groupItem);
  const ItemDesc& D = p.items[item];
  const ItemState& S = p.state[item];
  const ViewDesc& V = p.views[D.view];
  float f[kFeatures];
#pragma unroll
  for (int k = 0; k < kFeatures; ++k) f[k] = 0.0f;
  uint32_t n = 0, convex = 0;
  uint32_t c0 = 0xFFFFu, c1 = 0, r0 = 0xFFFFu, r1 = 0;
  float zmin = 1.0f, zmax = 0.0f, zsum = 0.0f, areaSum = 0.0f;
  const float vlen = sqrtf(V.mvp.raw[12] * V.mvp.raw[12] + V.mvp.raw[13] * V.mvp.raw[13] + V.mvp.raw[14] * V.mvp.raw[14]);
  const float vd[3] = {V.mvp.raw[12] / vlen, V.mvp.raw[13] / vlen, V.mvp.raw[14] / vlen};
  float ndSum = 0.0f, ndMin = 1.0f, ndMax = -1.0f, cen[3] = {0.0f, 0.0f, 0.0f};
  float ext[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) ext[j] = D.refMax[j] - D.refMin[j];
  // The batch's 4 packets of 8 quads: 32 contiguous words (word 8j + k is
  // vertex j of quad k), read as 16-byte loads.
  uint32_t W[32];
  {
    const uint4* src = reinterpret_cast<const uint4*>(D.packets + size_t(b * 8u - D.quadBase) * 4u);
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const uint4 x = src[v];
      W[4 * v] = x.x;
      W[4 * v + 1] = x.y;
      W[4 * v + 2] = x.z;
      W[4 * v + 3] = x.w;
    }
  }
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t g = b * 8u + k;
    float P[4][3];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t w = W[8 * j + k];
      const float X = float((int32_t(w) >> 21) + 1024), Y = float((w >> 10) & 2047u), Z = float(w & 1023u);
      P[j][0] = D.refMin[0] + X * (ext[0] / 2047.0f);
      P[j][1] = D.refMin[1] + Y * (ext[1] / 2047.0f);
      P[j][2] = D.refMin[2] + Z * (ext[2] / 1023.0f);
#pragma unroll
      for (int a = 0; a < 3; ++a) cen[a] += P[j][a] * (1.0f / 32.0f);
    }
    if (S.status == kStatusSkip) continue;  // not drawn: no valid quads (records not written)
    const uint32_t fz = p.featZ[g];
    if (!(fz & kFeatValid)) continue;  // empty rectangle
    ++n;
    const uint64_t bb = p.bbox[g];
    const uint32_t bc0 = uint32_t(bb & 0xFFFF), bc1 = uint32_t((bb >> 16) & 0xFFFF);
    const uint32_t br0 = uint32_t((bb >> 32) & 0xFFFF), br1 = uint32_t(bb >> 48);
    c0 = min(c0, bc0);
    c1 = max(c1, bc1);
    r0 = min(r0, br0);
    r1 = max(r1, br1);
    areaSum += float((bc1 - bc0 + 1) * (br1 - br0 + 1));
    const float z = float(fz >> 16) * (1.0f / 65535.0f);
    zmin = fminf(zmin, z);
    zmax = fmaxf(zmax, z);
    zsum += z;
    convex += (fz & 0xFFu) == kModeConvex ? 1u : 0u;
    float e1[3], e2[3], e3[3], nrm[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      e1[a] = P[1][a] - P[0][a];
      e2[a] = P[2][a] - P[0][a];
      e3[a] = P[3][a] - P[0][a];
    }
    nrm[0] = (e1[1] * e2[2] - e1[2] * e2[1]) + (e2[1] * e3[2] - e2[2] * e3[1]);
    nrm[1] = (e1[2] * e2[0] - e1[0] * e2[2]) + (e2[2] * e3[0] - e2[0] * e3[2]);
    nrm[2] = (e1[0] * e2[1] - e1[1] * e2[0]) + (e2[0] * e3[1] - e2[1] * e3[0]);
    const float nl = sqrtf(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    const float nd = nl > 0.0f ? (nrm[0] * vd[0] + nrm[1] * vd[1] + nrm[2] * vd[2]) / nl : 0.0f;
    ndSum += nd;
    ndMin = fminf(ndMin, nd);
    ndMax = fmaxf(ndMax, nd);
  }
  f[0] = float(n) * 0.125f;
  if (n > 0) {
    f[1] = log2f(1.0f + float((c1 - c0 + 1) * (r1 - r0 + 1)));
    f[2] = zmin;
    f[3] = zmax;
    f[4] = zsum / float(n);
    f[5] = log2f(1.0f + areaSum);
    f[7] = (0.5f * float(c0 + c1) + 0.5f) / float(V.bX);
    f[8] = (0.5f * float(r0 + r1) + 0.5f) / float(V.bY);
    f[11] = ndSum / float(n);
    f[12] = ndMin;
    f[13] = ndMax;
    f[15] = float(convex) / float(n);
  } else {
    f[7] = f[8] = 0.5f;
  }
  f[6] = S.status == kStatusClip ? 1.0f : 0.0f;
  const float w = cen[0] * V.mvp.raw[12] + cen[1] * V.mvp.raw[13] + cen[2] * V.mvp.raw[14] + V.mvp.raw[15];
  f[9] = log2f(1.0f + fabsf(w));
  f[10] = log2f(1.0f + float(p.labels[b]));
  f[14] = V.nBatches > 1 ? float(b - V.batchBase) / float(V.nBatches - 1) : 0.0f;
#pragma unroll
  for (int k = 0; k < kFeatures; ++k) p.features[size_t(b) * kFeatures + k] = f[k];
  return S.status != kStatusSkip;
}

__global__ void __launch_bounds__(256) rz_features_kernel(LearnParams p, uint32_t nBatches) {
  __shared__ uint32_t wgCount, wgBase;
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  const bool drawn = b < nBatches && batch_features(p, b);
  if (!p.drawn) return;
  // Append the work-group's drawn batches to the list with one atomic on the
  // list's count (the count is a single word: one atomic per wave serialised
  // there).
  if (threadIdx.x == 0) wgCount = 0;
  __syncthreads();
  const uint64_t m = __ballot(drawn);
  uint32_t waveOff = 0;
  if ((threadIdx.x & 63u) == 0 && m != 0ull) waveOff = atomicAdd(&wgCount, uint32_t(__popcll(m)));
  waveOff = uint32_t(__builtin_amdgcn_readlane(int(waveOff), 0));
  __syncthreads();
  if (threadIdx.x == 0) wgBase = wgCount != 0 ? atomicAdd(p.nDrawn, wgCount) : 0u;
  __syncthreads();
  const uint32_t pre = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
  if (drawn) p.drawn[wgBase + waveOff + pre] = b;
}

// Item k of a reordered draw: batch order[k] alone, with its occluder's
// classification and the footprint of the batch's drawn quads.
__global__ void __launch_bounds__(256) rz_batch_items_kernel(LearnParams p, const uint32_t* __restrict__ order,
                                                              uint32_t nBatches, ItemDesc* __restrict__ outItems,
                                                              ItemState* __restrict__ outState) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nBatches) return;
  const uint32_t b = order[k];
  const uint32_t item = item_of_quad(p.items, p.nItems, b * 8u, p.groupItem);
  ItemDesc D = p.items[item];
  ItemState S = p.state[item];
  D.quadBase = b * 8u;
  D.quadCount = 8u;
  uint32_t c0 = 0xFFFFFFFFu, c1 = 0, r0 = 0xFFFFFFFFu, r1 = 0;
  for (uint32_t j = 0; j < 8 && S.status != kStatusSkip; ++j) {  // no boxes are written for skipped occluders
    const uint64_t bb = p.bbox[b * 8u + j];
    const uint32_t bc0 = uint32_t(bb & 0xFFFF), bc1 = uint32_t((bb >> 16) & 0xFFFF);
    if (bc0 > bc1) continue;  // empty box: quad not drawn
    c0 = min(c0, bc0);
    c1 = max(c1, bc1);
    r0 = min(r0, uint32_t((bb >> 32) & 0xFFFF));
    r1 = max(r1, uint32_t(bb >> 48));
  }
  S.fp[0] = c0;
  S.fp[1] = c1;
  S.fp[2] = r0;
  S.fp[3] = r1;
  outItems[k] = D;
  outState[k] = S;
}

__global__ void __launch_bounds__(256) rz_item_scores_kernel(LearnParams p, const float* __restrict__ batchScores,
                                                              float* __restrict__ itemScores) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.nItems) return;
  const ItemDesc& D = p.items[i];
  float best = -3.402823466e+38f;
  for (uint32_t b = D.quadBase / 8u; b < (D.quadBase + D.quadCount) / 8u; ++b) best = fmaxf(best, batchScores[b]);
  itemScores[i] = best;
}

__global__ void __launch_bounds__(256) rz_sort_keys_kernel(const ViewDesc* __restrict__ views, uint32_t nViews,
                                                           const float* __restrict__ scores, uint32_t n, bool items,
                                                           uint64_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t lo = 0, hi = nViews - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if ((items ? views[mid].itemBase : views[mid].batchBase) <= i) lo = mid;
    else hi = mid - 1;
  }
  const uint32_t u = __float_as_uint(scores[i]);
  const uint32_t ordered = (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // ascending with the float order
  keys[i] = (uint64_t(lo) << 32) | uint64_t(~ordered);
}

// Squared distance of Main.cpp:184-190: _mm_dp_ps(d, d, 0x7f) sums the x
// and y products, then adds z's (each rounded; compiled without
// contraction, so no fused multiply-add).
__global__ void __launch_bounds__(256) rz_f2b_keys_kernel(const ItemDesc* __restrict__ items, uint32_t n,
                                                          const float* __restrict__ eyes, uint64_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ItemDesc& D = items[i];
  const float* e = eyes + 4 * D.view;
  uint32_t bits = 0;
  if (e[3] != 0.0f) {
    float p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float c = (D.boundsMax[a] + D.boundsMin[a]) * 0.5f;
      const float d = c - e[a];
      p[a] = d * d;
    }
    bits = __float_as_uint((p[0] + p[1]) + p[2]);  // >= +0 (or NaN, last): its bits order as the value
  }
  keys[i] = (uint64_t(D.view) << 32) | uint64_t(bits);
}

__global__ void __launch_bounds__(256) rz_item_reorder_kernel(LearnParams p, const uint32_t* __restrict__ order,
                                                               ItemDesc* __restrict__ outItems,
                                                               ItemState* __restrict__ outState) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.nItems) return;
  const uint32_t i = order[k];
  outItems[k] = p.items[i];
  outState[k] = p.state[i];
}

__global__ void rz_fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
static inline uint32_t groups(uint32_t n, uint32_t per) { return (n + per - 1) / per; }

hipError_t launch_prep(const PrepParams& p, hipStream_t stream) {
  if (p.nItems == 0) return hipSuccess;
  // One wave per work-group: a frame's few hundred items then spread over as
  // many CUs, each with its own reads in flight (from host memory, for
  // single-view draws).
  hipLaunchKernelGGL(rz_prep_kernel, dim3(groups(p.nItems, 64)), dim3(64), 0, stream, p);
  return hipGetLastError();
}

hipError_t launch_setup(const SetupParams& p, hipStream_t stream) {
  if (p.nQuads == 0 || p.nItems == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_setup_kernel, dim3(groups(p.nQuads, 256)), dim3(256), 0, stream, p);
  return hipGetLastError();
}

template <int T, int NW>
static const void* frame_fn(bool stats) {
  return stats ? reinterpret_cast<const void*>(&rz_frame_kernel<T, T, NW, true>)
               : reinterpret_cast<const void*>(&rz_frame_kernel<T, T, NW, false>);
}

static const void* frame_kernel_fn(int tileW, int waves, bool stats) {
  if (tileW == 8)
    return waves == 8 ? frame_fn<8, 8>(stats)
           : waves == 4 ? frame_fn<8, 4>(stats)
           : waves == 2 ? frame_fn<8, 2>(stats)
                        : nullptr;
  if (tileW == 16)
    return waves == 8 ? frame_fn<16, 8>(stats)
           : waves == 4 ? frame_fn<16, 4>(stats)
           : waves == 2 ? frame_fn<16, 2>(stats)
                        : nullptr;
  return nullptr;
}

// Launches a grid that the host has sized to be co-resident, cooperatively
// when asked.  The runtime can refuse a cooperative launch while another
// process on the device holds its cooperative (global wave sync) resources;
// the grid is then dispatched plainly, which the kernels allow: they never
// synchronise the whole grid, and each of their cross-tile waits is bounded
// in time (a grid that another process kept from being co-resident reports
// RZ_ERR_TIMEOUT instead of hanging).
static hipError_t launch_coresident(const void* fn, dim3 grid, dim3 block, void** args, hipStream_t stream,
                                    bool cooperative) {
  if (cooperative) {
    const hipError_t e = hipLaunchCooperativeKernel(fn, grid, block, args, 0, stream);
    if (e != hipErrorIllegalState && e != hipErrorCooperativeLaunchTooLarge) return e;
    (void)hipGetLastError();  // the refusal is handled: clear it
  }
  return hipLaunchKernel(fn, grid, block, args, 0, stream);
}

hipError_t launch_frame(const FrameParams& p, int tileW, hipStream_t stream, bool cooperative, int waves) {
  if (p.nItems == 0) return hipSuccess;
  FrameParams local = p;
  void* args[] = {&local};
  const dim3 grid(p.tilesX * p.tilesY), block(64 * waves);
  const void* fn = frame_kernel_fn(tileW, waves, p.stats != nullptr);
  if (!fn) return hipErrorInvalidValue;
  return launch_coresident(fn, grid, block, args, stream, cooperative);
}

hipError_t frame_max_groups(int tileW, int waves, int* maxGroups) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  int perCU = 1 << 30;
  for (bool stats : {false, true}) {
    const void* fn = frame_kernel_fn(tileW, waves, stats);
    if (!fn) return hipErrorInvalidValue;
    int n = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64 * waves, 0);
    if (e != hipSuccess) return e;
    perCU = n < perCU ? n : perCU;
  }
  *maxGroups = perCU * cus;
  return hipSuccess;
}

hipError_t launch_status(const ItemState* state, const uint32_t* visible, uint32_t n, uint8_t* out,
                         hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_status_kernel, dim3(groups(n, 256)), dim3(256), 0, stream, state, visible, n, out);
  return hipGetLastError();
}

hipError_t launch_query(const QueryParams& p, hipStream_t stream) {
  if (p.n == 0) return hipSuccess;
  constexpr uint32_t kBoxesPerWave = 4;  // measured over 1, 2, 4, 8, 16 and 32 on the city frames' batches
  hipLaunchKernelGGL(rz_query_kernel<kBoxesPerWave>, dim3(groups(groups(p.n, kBoxesPerWave), 4)), dim3(256), 0, stream,
                     p);
  return hipGetLastError();
}

hipError_t launch_query2d(uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY, uint32_t maxZ,
                          const QueryParams& p, uint8_t* out, hipStream_t stream) {
  hipLaunchKernelGGL(rz_query2d_kernel, dim3(1), dim3(64), 0, stream, minX, maxX, minY, maxY, maxZ, p, out);
  return hipGetLastError();
}

hipError_t launch_tile_min(const uint16_t* hiz, uint32_t bX, uint32_t bY, uint16_t* tileMin, uint32_t pyrX,
                           uint32_t pyrY, hipStream_t stream) {
  hipLaunchKernelGGL(rz_tile_min_kernel, dim3(groups(pyrX * pyrY, 4)), dim3(256), 0, stream, hiz, bX, bY, tileMin,
                     pyrX, pyrY);
  return hipGetLastError();
}

hipError_t launch_readback(const uint16_t* depth, const uint16_t* hiz, uint32_t width, uint32_t height, uint32_t bX,
                           uint32_t* rgba, hipStream_t stream) {
  hipLaunchKernelGGL(rz_readback_kernel, dim3(groups(width, 256), height), dim3(256), 0, stream, depth, hiz, width,
                     height, bX, rgba);
  return hipGetLastError();
}

hipError_t launch_clear(uint16_t* hiz, uint32_t nHiz, uint16_t* tileMin, uint32_t nMin, hipStream_t stream) {
  if (nHiz % 8 != 0 || reinterpret_cast<uintptr_t>(hiz) % 16 != 0) return hipErrorInvalidValue;
  const uint32_t n = nHiz / 8 + nMin;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_clear_kernel, dim3(groups(n, 256)), dim3(256), 0, stream, reinterpret_cast<uint4*>(hiz),
                     nHiz / 8, tileMin, nMin);
  return hipGetLastError();
}

// Labelled draws need 8x8-block tiles (the block-major draw); 16x16-block
// tiles have no labelled variant.
template <int NW>
static const void* views_fn(int tileB, bool stats, bool labels) {
  if (tileB == 16)
    return stats ? reinterpret_cast<const void*>(&rz_views_kernel<16, 16, NW, true, false>)
                 : reinterpret_cast<const void*>(&rz_views_kernel<16, 16, NW, false, false>);
  if (labels)
    return stats ? reinterpret_cast<const void*>(&rz_views_kernel<8, 8, NW, true, true>)
                 : reinterpret_cast<const void*>(&rz_views_kernel<8, 8, NW, false, true>);
  return stats ? reinterpret_cast<const void*>(&rz_views_kernel<8, 8, NW, true, false>)
               : reinterpret_cast<const void*>(&rz_views_kernel<8, 8, NW, false, false>);
}

hipError_t launch_views(const ViewsParams& p, uint32_t nGroups, hipStream_t stream, bool cooperative) {
  if (p.nUnits == 0 || nGroups == 0) return hipSuccess;
  const bool st = p.stats != nullptr, lab = p.labels != nullptr;
  if (p.tileB != 8 && (p.tileB != 16 || lab)) return hipErrorInvalidValue;
  const int tb = int(p.tileB);
  if (p.waves != 2 && p.waves != 4 && p.waves != 8) return hipErrorInvalidValue;
  const void* fn = p.waves == 8 ? views_fn<8>(tb, st, lab) : p.waves == 4 ? views_fn<4>(tb, st, lab)
                                                                          : views_fn<2>(tb, st, lab);
  ViewsParams local = p;
  void* args[] = {&local};
  const dim3 grid(nGroups), block(64 * p.waves);
  return launch_coresident(fn, grid, block, args, stream, cooperative);
}

hipError_t views_max_groups(int tileB, int waves, int* maxGroups) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  int perCU = 1 << 30;
  for (int variant = 0; variant < (tileB == 16 ? 2 : 4); ++variant) {
    int n = 0;
    const bool stats = variant & 1, labels = variant & 2;
    if (waves != 2 && waves != 4 && waves != 8) return hipErrorInvalidValue;
    const void* fn = waves == 8 ? views_fn<8>(tileB, stats, labels)
                     : waves == 4 ? views_fn<4>(tileB, stats, labels)
                                  : views_fn<2>(tileB, stats, labels);
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64 * waves, 0);
    if (e != hipSuccess) return e;
    perCU = n < perCU ? n : perCU;
  }
  *maxGroups = perCU * cus;
  return hipSuccess;
}

hipError_t launch_views_query(const ViewsQueryParams& p, hipStream_t stream) {
  if (p.nViews == 0 || (p.nQueries == 0 && !p.errOut)) return hipSuccess;
  hipLaunchKernelGGL(rz_views_query_kernel, dim3(std::max(groups(p.maxPerView, 256), 1u), p.nViews), dim3(256), 0,
                     stream, p);
  return hipGetLastError();
}

hipError_t launch_features(const LearnParams& p, uint32_t nBatches, hipStream_t stream) {
  if (nBatches == 0 || p.nItems == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_features_kernel, dim3(groups(nBatches, 256)), dim3(256), 0, stream, p, nBatches);
  return hipGetLastError();
}

hipError_t launch_batch_items(const LearnParams& p, const uint32_t* order, uint32_t nBatches, ItemDesc* outItems,
                              ItemState* outState, hipStream_t stream) {
  if (nBatches == 0 || p.nItems == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_batch_items_kernel, dim3(groups(nBatches, 256)), dim3(256), 0, stream, p, order, nBatches,
                     outItems, outState);
  return hipGetLastError();
}

hipError_t launch_item_scores(const LearnParams& p, const float* batchScores, float* itemScores,
                              hipStream_t stream) {
  if (p.nItems == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_item_scores_kernel, dim3(groups(p.nItems, 256)), dim3(256), 0, stream, p, batchScores,
                     itemScores);
  return hipGetLastError();
}

hipError_t launch_sort_keys(const ViewDesc* views, uint32_t nViews, const float* scores, uint32_t n, bool items,
                            uint64_t* keys, hipStream_t stream) {
  if (n == 0 || nViews == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_sort_keys_kernel, dim3(groups(n, 256)), dim3(256), 0, stream, views, nViews, scores, n, items,
                     keys);
  return hipGetLastError();
}

hipError_t launch_f2b_keys(const ItemDesc* items, uint32_t n, const float* eyes, uint64_t* keys, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_f2b_keys_kernel, dim3(groups(n, 256)), dim3(256), 0, stream, items, n, eyes, keys);
  return hipGetLastError();
}

hipError_t launch_item_reorder(const LearnParams& p, const uint32_t* order, ItemDesc* outItems, ItemState* outState,
                               hipStream_t stream) {
  if (p.nItems == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_item_reorder_kernel, dim3(groups(p.nItems, 256)), dim3(256), 0, stream, p, order, outItems,
                     outState);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_fill_u32_kernel, dim3(groups(n, 256)), dim3(256), 0, stream, p, n, v);
  return hipGetLastError();
}

hipError_t launch_views_clear(const ViewDesc* src, ViewDesc* views, uint32_t nViews, uint32_t maxEntries,
                              hipStream_t stream) {
  if (nViews == 0) return hipSuccess;
  hipLaunchKernelGGL(rz_views_clear_kernel, dim3(groups(std::max(maxEntries, 1u), 256), nViews), dim3(256), 0, stream,
                     src, views);
  return hipGetLastError();
}

}  // namespace rzd
