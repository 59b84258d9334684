/*
 * oracle.c -- scalar CPU restatement of rawrunprotected/rasterizer's
 * occlusion-culling hot path.  TEST INFRASTRUCTURE ONLY: see oracle.h.
 *
 * Every function cites the reference lines it restates (paths relative to
 * the reference's SoftwareRasterizer/ directory).  The reference works on 8
 * quads per AVX2 register; all of its per-quad arithmetic is lane-local, so
 * processing one quad at a time in submission order (packet, then lane) gives
 * the same buffer.  Built with -ffp-contract=off: every product and sum is
 * rounded on its own, and fmaf() appears exactly at the reference's fused
 * multiply-add sites.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------------------------------------------------------------------
 * x86 primitives, executed by the host CPU (single lane of the packed forms).
 * ------------------------------------------------------------------------- */

/* RCPPS (Rasterizer.cpp:717-727 use the 256-bit form, :218/:223 the 128-bit
 * form; both are the same approximation on the hosts probed). */
static inline float x_rcp256(float x) {
  __m256 r = _mm256_rcp_ps(_mm256_set1_ps(x));
  return _mm256_cvtss_f32(r);
}
static inline float x_rcp128(float x) { return _mm_cvtss_f32(_mm_rcp_ps(_mm_set1_ps(x))); }
/* MINPS/MAXPS: dst = a < b ? a : b (second operand on NaN or equal). */
static inline float x_min(float a, float b) { return _mm_cvtss_f32(_mm_min_ss(_mm_set_ss(a), _mm_set_ss(b))); }
static inline float x_max(float a, float b) { return _mm_cvtss_f32(_mm_max_ss(_mm_set_ss(a), _mm_set_ss(b))); }
/* CVTTPS2DQ: truncation, 0x80000000 on NaN / overflow. */
static inline int32_t x_cvtt(float x) { return _mm_cvttss_si32(_mm_set_ss(x)); }
static inline float x_round_ne(float x) {
  return _mm_cvtss_f32(_mm_round_ss(_mm_setzero_ps(), _mm_set_ss(x), _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC));
}
static inline float x_floor(float x) {
  return _mm_cvtss_f32(_mm_round_ss(_mm_setzero_ps(), _mm_set_ss(x), _MM_FROUND_TO_NEG_INF | _MM_FROUND_NO_EXC));
}
/* DPPS with immediate 0xFF / 0x7F, lane 0 of the broadcast result. */
static inline float x_dp4(const float a[4], const float b[4]) {
  return _mm_cvtss_f32(_mm_dp_ps(_mm_loadu_ps(a), _mm_loadu_ps(b), 0xFF));
}
static inline float x_dp3(const float a[4], const float b[4]) {
  return _mm_cvtss_f32(_mm_dp_ps(_mm_loadu_ps(a), _mm_loadu_ps(b), 0x7F));
}
static inline float x_rsqrt(float x) { return _mm_cvtss_f32(_mm_rsqrt_ps(_mm_set1_ps(x))); }

float or_x86_rcp(float x) { return x_rcp256(x); }
int32_t or_x86_cvtt(float x) { return x_cvtt(x); }

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline float xor_sign(float f, uint32_t signbits) { return bitsf(fbits(f) ^ signbits); }

/* packDepthPremultiplied (Rasterizer.cpp:508-525): PACKUSDW of (bits >> 12). */
static inline uint16_t pack_depth(float d) {
  int32_t v = ((int32_t)fbits(d)) >> 12;
  if (v < 0) return 0;
  if (v > 65535) return 65535;
  return (uint16_t)v;
}
static inline uint16_t avg_u16(uint16_t a, uint16_t b) { return (uint16_t)(((uint32_t)a + b + 1) >> 1); }

/* ---------------------------------------------------------------------------
 * Constants (Rasterizer.cpp:9-17), evaluated in float as the reference's
 * constant expressions are.
 * ------------------------------------------------------------------------- */
static const float kFloatCompressionBias = 2.5237386e-29f;
static const float kMinEdgeOffset = -0.45f;

enum { Culled = 0, Triangle0, Triangle1, ConcaveRight, ConcaveLeft, ConcaveCenter, Convex };

/* modeTable (Rasterizer.cpp:30-64): primitive mode per 8-bit configuration
 * (bits 0-3 = area sign of triangles 0..3, bits 4-7 = sign of w0..w3).  The
 * table is data the algorithm is defined by and is reproduced verbatim. */
static const uint8_t kModeTable[256] = {
  6,2,4,2,1,0,1,0, 3,2,0,2,1,0,1,0, 6,2,4,2,1,0,1,0, 0,2,5,2,1,0,1,0,
  6,0,4,2,1,0,1,0, 3,2,5,2,1,0,0,0, 6,2,4,0,1,0,1,0, 3,2,5,2,0,0,1,0,
  6,2,0,2,1,0,1,0, 3,2,5,2,1,0,1,0, 0,2,5,2,1,0,1,0, 5,2,5,2,1,0,1,0,
  6,2,4,2,1,0,0,0, 3,0,5,2,1,0,1,0, 2,2,2,2,0,0,0,0, 2,2,2,0,0,0,0,0,
  6,2,4,2,0,0,1,0, 3,2,5,0,1,0,1,0, 6,2,4,2,1,0,0,0, 3,0,5,2,1,0,1,0,
  0,2,4,2,1,0,1,0, 3,2,5,2,1,0,1,0, 3,2,0,2,1,0,1,0, 3,2,5,2,1,0,1,0,
  6,2,4,0,1,0,1,0, 3,2,5,2,0,0,1,0, 1,0,1,0,1,0,1,0, 1,0,1,0,1,0,0,0,
  4,2,4,2,1,0,1,0, 0,2,5,2,1,0,1,0, 0,0,0,0,0,0,0,0, 0,0,0,0,0,0,0,0,
};

struct OrRaster {
  uint32_t width, height, blocksX, blocksY;
  float mvp[16];      /* m_modelViewProjection (prebaked columns) */
  float mvpRaw[16];   /* m_modelViewProjectionRaw (rows) */
  uint64_t table[4096];
  uint16_t* depth;    /* blocksX*blocksY*64 */
  uint16_t* hiz;      /* blocksX*blocksY + 8 */
  uint64_t stats[7];
  uint32_t* blockVisits; /* optional per-block counters: [2*b] visited, [2*b+1] written */
  /* Write tracking (or_frame_labels): when set, winner[pixel] is the batch
   * whose write produced the pixel's current depth (first such write). */
  uint32_t* winner;
  uint32_t curBatch;
};

/* transposeMask (Rasterizer.cpp:527-545): pixel (x, y), sampled as bit
 * 8x+y, moves to bit 8x + 4 + y/2 for even y and 8x + (y-1)/2 for odd y. */
static uint64_t transpose_mask(uint64_t m) {
  uint64_t out = 0;
  for (int x = 0; x < 8; ++x)
    for (int y = 0; y < 8; ++y)
      if ((m >> (8 * x + y)) & 1) {
        int bit = 8 * x + ((y & 1) ? (y - 1) / 2 : 4 + y / 2);
        out |= (uint64_t)1 << bit;
      }
  return out;
}

/* precomputeRasterizationTable (Rasterizer.cpp:547-604) with the 128-bit
 * quantizeSlopeLookup (:470-480) and quantizeOffsetLookup (:496-506). */
static int build_table(uint64_t* table) {
  const uint32_t angularResolution = 2000, offsetResolution = 2000;
  memset(table, 0, 4096 * sizeof(uint64_t));
  const float maxOffset = -kMinEdgeOffset;
  const float offMul = 63.0f / (maxOffset - kMinEdgeOffset);
  const float offAdd = 0.5f - kMinEdgeOffset * offMul;
  const float slopeMul = (32 - 1) * 0.5f;
  const float slopeAdd = slopeMul + 0.5f;
  int ok = 1;
  for (uint32_t i = 0; i < angularResolution; ++i) {
    float angle = -0.1f + 6.4f * (float)i / (float)(angularResolution - 1);
    float nx = cosf(angle);
    float ny = sinf(angle);
    float l = 1.0f / (fabsf(nx) + fabsf(ny));
    nx *= l;
    ny *= l;
    int32_t yNeg = (ny < 0.0f) ? -1 : 0;
    int32_t q = x_cvtt(fmaf(nx, slopeMul, slopeAdd));
    uint32_t slopeLookup = (uint32_t)(((q << 1) - yNeg) << 6);
    for (uint32_t j = 0; j < offsetResolution; ++j) {
      float offset = -0.6f + 1.2f * (float)j / (float)(angularResolution - 1);
      float lookup = offset * offMul + offAdd;
      int32_t o = x_cvtt(lookup);
      if (o < 0) o = 0;
      if (o > 63) o = 63;
      uint64_t block = 0;
      for (int x = 0; x < 8; ++x)
        for (int y = 0; y < 8; ++y) {
          float edgeDistance = offset + ((float)x - 3.5f) / 8.0f * nx + ((float)y - 3.5f) / 8.0f * ny;
          if (edgeDistance <= 0.0f) block |= (uint64_t)1 << (8 * x + y);
        }
      table[slopeLookup | (uint32_t)o] |= transpose_mask(block);
    }
    /* invariant checked by the reference with __debugbreak (:592-602) */
    if (table[slopeLookup] != ~(uint64_t)0) ok = 0;
    if (table[slopeLookup + 63] != 0) ok = 0;
  }
  return ok;
}

OrRaster* or_create(uint32_t width, uint32_t height) {
  if (width == 0 || height == 0 || (width % 8) || (height % 8)) return NULL;
  OrRaster* r = (OrRaster*)calloc(1, sizeof(OrRaster));
  if (!r) return NULL;
  r->width = width;
  r->height = height;
  r->blocksX = width / 8;
  r->blocksY = height / 8;
  size_t nb = (size_t)r->blocksX * r->blocksY;
  r->depth = (uint16_t*)calloc(nb * 64, sizeof(uint16_t));
  r->hiz = (uint16_t*)calloc(nb + 8, sizeof(uint16_t));
  if (!r->depth || !r->hiz || !build_table(r->table)) {
    or_destroy(r);
    return NULL;
  }
  return r;
}

void or_destroy(OrRaster* r) {
  if (!r) return;
  free(r->blockVisits);
  free(r->depth);
  free(r->hiz);
  free(r);
}

/* setModelViewProjection (Rasterizer.cpp:76-105). */
void or_set_mvp(OrRaster* r, const float m[16]) {
  float rows[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) rows[i][j] = m[4 * j + i]; /* transpose */
  memcpy(r->mvpRaw, rows, sizeof rows);
  const float sx = (float)r->width * 0.5f - 4.0f;
  const float sy = (float)r->height * 0.5f - 4.0f;
  const float sz = 0.5f * kFloatCompressionBias;
  float b[4][4];
  for (int j = 0; j < 4; ++j) {
    b[0][j] = (rows[0][j] + rows[3][j]) * sx;
    b[1][j] = (rows[1][j] + rows[3][j]) * sy;
    b[2][j] = (rows[3][j] - rows[2][j]) * sz;
    b[3][j] = rows[3][j];
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r->mvp[4 * i + j] = b[j][i]; /* transpose back */
}

/* clear (Rasterizer.cpp:107-121). */
void or_clear(OrRaster* r) {
  /* Whole 8-entry stores over size / 8 groups (Rasterizer.cpp:115-120): the
   * last (blocks + 8) % 8 padding entries keep their value. */
  size_t n = (((size_t)r->blocksX * r->blocksY + 8) / 8) * 8;
  for (size_t i = 0; i < n; ++i) r->hiz[i] = 1;
}

/* ---------------------------------------------------------------------------
 * rasterize<possiblyNearClipped> (Rasterizer.cpp:606-1295)
 * ------------------------------------------------------------------------- */
typedef struct {
  float T[4][4];  /* prologue matrix: T[out][in], out = X,Y,Z,W; in = Xf,Yf,Zf,1 */
  float c0, c1;
} Prologue;

/* Per-occluder prologue (Rasterizer.cpp:616-655). */
static void prologue(const OrRaster* r, const float refMin[4], const float refMax[4], Prologue* P) {
  float mat[4][4];
  for (int c = 0; c < 4; ++c)
    for (int j = 0; j < 4; ++j) mat[c][j] = r->mvp[4 * c + j];
  float ext[4];
  for (int j = 0; j < 4; ++j) ext[j] = refMax[j] - refMin[j];
  const float k0 = 1.0f / (float)(2047ull << 21);
  const float k1 = 1.0f / (float)(2047 << 10);
  const float k2 = 1.0f / (float)1023;
  float m0[4], m1[4], m2[4], m3[4];
  for (int j = 0; j < 4; ++j) {
    m3[j] = fmaf(mat[0][j], refMin[0], fmaf(mat[1][j], refMin[1], fmaf(mat[2][j], refMin[2], mat[3][j])));
    m0[j] = mat[0][j] * (ext[0] * k0);
    m1[j] = mat[1][j] * (ext[1] * k1);
    m2[j] = mat[2][j] * (ext[2] * k2);
  }
  for (int j = 0; j < 4; ++j) m3[j] = fmaf(m0[j], (float)(1024ull << 21), m3[j]);
  for (int j = 0; j < 4; ++j) {
    m1[j] = m1[j] - m0[j];
    m2[j] = m2[j] - m0[j];
  }
  for (int o = 0; o < 4; ++o) {
    P->T[o][0] = m0[o];
    P->T[o][1] = m1[o];
    P->T[o][2] = m2[o];
    P->T[o][3] = m3[o];
  }
  const float s[4] = {(float)(1 << 21), (float)(1 << 10), 1.0f, 1.0f};
  float Za = P->T[2][3], Zb = x_dp4(P->T[2], s);
  float Wa = P->T[3][3], Wb = x_dp4(P->T[3], s);
  float q = (Za - Zb) / (Wa - Wb);
  P->c0 = q;
  /* _mm_fnmadd_ps(q, Wa, Za) as the reference calls it (Rasterizer.cpp:654):
   * a NaN q (0/0 from a degenerate reference box) propagates with its sign,
   * where fmaf(-q, ...) would flip it. */
  P->c1 = _mm_cvtss_f32(_mm_fnmadd_ss(_mm_set_ss(q), _mm_set_ss(Wa), _mm_set_ss(Za)));
}

typedef struct {
  int valid;
  int mode;
  int32_t minX, minY, rangeX, rangeY;
  uint32_t firstBlock;
  uint16_t maxZ;
  float dp0, dpDx, dpDy;
  float nx[4], ny[4], off[4];
  uint32_t slope[4];
} QuadSetup;

/* normalizeEdge (Rasterizer.cpp:452-468). */
static void normalize_edge(int clip, float* nx, float* ny, uint32_t flip) {
  float invLen = x_rcp256(fabsf(*nx) + fabsf(*ny));
  const float maxOffset = -kMinEdgeOffset;
  float mul = (float)(64 - 1) / (maxOffset - kMinEdgeOffset);
  if (clip) mul = xor_sign(mul, flip);
  invLen = mul * invLen;
  *nx = *nx * invLen;
  *ny = *ny * invLen;
}

/* quantizeSlopeLookup, 256-bit form (Rasterizer.cpp:482-493): note the
 * ny <= 0 test (the table build uses ny < 0). */
static uint32_t quantize_slope(float nx, float ny) {
  const float maxOffset = -kMinEdgeOffset;
  const float mul = (32 - 1) * 0.5f / ((64 - 1) / (maxOffset - kMinEdgeOffset));
  const float add = (32 - 1) * 0.5f + 0.5f;
  int32_t yNeg = (ny <= 0.0f) ? -1 : 0;
  int32_t q = x_cvtt(fmaf(nx, mul, add));
  return (uint32_t)((int32_t)((uint32_t)q << 1) - yNeg) << 6;
}

/* Per-quad setup of one lane (Rasterizer.cpp:660-1063). */
static void setup_quad(const OrRaster* r, const Prologue* P, const uint32_t I[4], int clip, QuadSetup* Q) {
  const float maxInvW = sqrtf(3.40282347e+38f);
  float x[4], y[4], invW[4];
  uint32_t wSign[4];
  for (int k = 0; k < 4; ++k) {
    float Xf = (float)(int32_t)I[k];
    float Yf = (float)(int32_t)(I[k] & (2047u << 10));
    float Zf = (float)(int32_t)(I[k] & 1023u);
    float X = fmaf(Xf, P->T[0][0], fmaf(Yf, P->T[0][1], fmaf(Zf, P->T[0][2], P->T[0][3])));
    float Y = fmaf(Xf, P->T[1][0], fmaf(Yf, P->T[1][1], fmaf(Zf, P->T[1][2], P->T[1][3])));
    float W = fmaf(Xf, P->T[3][0], fmaf(Yf, P->T[3][1], fmaf(Zf, P->T[3][2], P->T[3][3])));
    if (clip)
      invW[k] = x_min(maxInvW, x_max(-maxInvW, x_rcp256(W)));
    else
      invW[k] = x_rcp256(W);
    x[k] = x_round_ne(X * invW[k]) * 0.125f;
    y[k] = x_round_ne(Y * invW[k]) * 0.125f;
    wSign[k] = clip ? (fbits(invW[k]) & 0x80000000u) : 0u;
  }
  float eX[5], eY[5];
  eX[0] = y[1] - y[0]; eX[1] = y[2] - y[1]; eX[2] = y[3] - y[2]; eX[3] = y[0] - y[3];
  eY[0] = x[0] - x[1]; eY[1] = x[1] - x[2]; eY[2] = x[2] - x[3]; eY[3] = x[3] - x[0];
  float area0 = fmaf(eX[0], eY[1], -(eX[1] * eY[0]));
  float area1 = fmaf(eX[1], eY[2], -(eX[2] * eY[1]));
  float area2 = fmaf(eX[2], eY[3], -(eX[3] * eY[2]));
  float area3 = (area0 + area2) - area1;

  /* Area signs, 0 counts as negative (:775-791). */
  int s0, s1, s2, s3;
  if (clip) {
    s0 = xor_sign(xor_sign(area0, wSign[0]), wSign[1] ^ wSign[2]) <= 0.0f;
    s1 = xor_sign(xor_sign(area1, wSign[1]), wSign[2] ^ wSign[3]) <= 0.0f;
    s2 = xor_sign(xor_sign(area2, wSign[0]), wSign[2] ^ wSign[3]) <= 0.0f;
    s3 = xor_sign(xor_sign(area3, wSign[1]), wSign[0] ^ wSign[3]) <= 0.0f;
  } else {
    s0 = area0 <= 0.0f;
    s1 = area1 <= 0.0f;
    s2 = area2 <= 0.0f;
    s3 = area3 <= 0.0f;
  }
  uint32_t config = (uint32_t)s0 | ((uint32_t)s1 << 1) | ((uint32_t)s2 << 2) | ((uint32_t)s3 << 3);
  if (clip)
    config |= ((wSign[0] >> 27) | (wSign[1] >> 26) | (wSign[2] >> 25) | (wSign[3] >> 24));
  int mode = kModeTable[config];
  Q->mode = mode;
  Q->valid = 0;
  if (mode == Culled) return;

  /* Bounding box (:816-919). */
  float minFx, minFy, maxFx, maxFy;
  if (clip) {
    const float infP = 10000.0f, infN = -10000.0f;
    float minPx[4], minPy[4], maxPx[4], maxPy[4], minNx[4], minNy[4], maxNx[4], maxNy[4];
    for (int k = 0; k < 4; ++k) {
      int neg = wSign[k] != 0;
      minPx[k] = neg ? infP : x[k];
      minPy[k] = neg ? infP : y[k];
      maxPx[k] = xor_sign(minPx[k], wSign[k]);
      maxPy[k] = xor_sign(minPy[k], wSign[k]);
      minNx[k] = neg ? x[k] : infP;
      minNy[k] = neg ? y[k] : infP;
      maxNx[k] = neg ? x[k] : infN;
      maxNy[k] = neg ? y[k] : infN;
    }
    float mPx = x_min(x_min(minPx[0], minPx[1]), x_min(minPx[2], minPx[3]));
    float mPy = x_min(x_min(minPy[0], minPy[1]), x_min(minPy[2], minPy[3]));
    float MPx = x_max(x_max(maxPx[0], maxPx[1]), x_max(maxPx[2], maxPx[3]));
    float MPy = x_max(x_max(maxPy[0], maxPy[1]), x_max(maxPy[2], maxPy[3]));
    float mNx = x_min(x_min(minNx[0], minNx[1]), x_min(minNx[2], minNx[3]));
    float mNy = x_min(x_min(minNy[0], minNy[1]), x_min(minNy[2], minNy[3]));
    float MNx = x_max(x_max(maxNx[0], maxNx[1]), x_max(maxNx[2], maxNx[3]));
    float MNy = x_max(x_max(maxNy[0], maxNy[1]), x_max(maxNy[2], maxNy[3]));
    float incAx = (MNx > mPx) ? infN : mPx;
    float incAy = (MNy > mPy) ? infN : mPy;
    float incBx = (MPx > mNx) ? infP : MPx;
    float incBy = (MPy > mNy) ? infP : MPy;
    minFx = x_min(incAx, incBx);
    minFy = x_min(incAy, incBy);
    maxFx = x_max(incAx, incBx);
    maxFy = x_max(incAy, incBy);
  } else {
    minFx = x_min(x_min(x[0], x[1]), x_min(x[2], x[3]));
    minFy = x_min(x_min(y[0], y[1]), x_min(y[2], y[3]));
    maxFx = x_max(x_max(x[0], x[1]), x_max(x[2], x[3]));
    maxFy = x_max(x_max(y[0], y[1]), x_max(y[2], y[3]));
  }
  /* Clamp and round (:921-935). */
  int32_t minX = x_cvtt(minFx + 4.9999f / 8.0f);
  int32_t minY = x_cvtt(minFy + 4.9999f / 8.0f);
  int32_t maxX = x_cvtt(maxFx + 11.0f / 8.0f);
  int32_t maxY = x_cvtt(maxFy + 11.0f / 8.0f);
  if (minX < 0) minX = 0;
  if (minY < 0) minY = 0;
  if (maxX > (int32_t)r->blocksX) maxX = (int32_t)r->blocksX;
  if (maxY > (int32_t)r->blocksY) maxY = (int32_t)r->blocksY;
  if (!(maxX > minX && maxY > minY)) return;
  Q->valid = 1;
  Q->minX = minX;
  Q->minY = minY;
  Q->rangeX = maxX - minX;
  Q->rangeY = maxY - minY;

  /* Depth from 1/w (:941-961). */
  float z[4];
  for (int k = 0; k < 4; ++k) z[k] = fmaf(invW[k], P->c1, P->c0);
  float maxZ = x_max(x_max(z[0], z[1]), x_max(z[2], z[3]));
  if (clip && ((wSign[0] | wSign[1] | wSign[2] | wSign[3]) != 0)) maxZ = 1.0f;
  Q->maxZ = pack_depth(maxZ);

  /* Depth plane (:963-999). */
  int greaterArea = fabsf(area0) < fabsf(area2);
  int tri0 = mode == Triangle0, tri1 = mode == Triangle1;
  greaterArea = !tri0 && (tri1 || greaterArea);
  float selArea = greaterArea ? area2 : area0;
  float invArea = clip ? 1.0f / selArea : x_rcp256(selArea);
  float z12 = z[1] - z[2], z20 = z[2] - z[0], z30 = z[3] - z[0];
  eX[4] = y[0] - y[2];
  eY[4] = x[2] - x[0];
  float dpx = greaterArea ? fmaf(-z20, eX[3], z30 * eX[4]) : fmaf(z20, eX[1], -(z12 * eX[4]));
  float dpy = greaterArea ? fmaf(-z20, eY[3], z30 * eY[4]) : fmaf(z20, eY[1], -(z12 * eY[4]));
  float dp1 = invArea * dpx;
  float dp2 = invArea * dpy;
  float x0 = x[0] - (float)minX;
  float y0 = y[0] - (float)minY;
  float dp0 = fmaf(-x0, dp1, fmaf(-y0, dp2, z[0]));
  Q->dp0 = dp0;
  Q->dpDx = dp1;
  Q->dpDy = dp2;

  /* Edge replacement for triangle modes (:1001-1005). */
  if (tri0) { eX[2] = eX[4]; eY[2] = eY[4]; }
  if (tri1) { eX[0] = -eX[4]; eY[0] = -eY[4]; }
  /* Flip masks (:1007-1022). */
  uint32_t f0 = 0, f1 = 0, f2 = 0, f3 = 0;
  if (clip) {
    f0 = wSign[0] ^ (tri1 ? wSign[2] : wSign[1]);
    f1 = wSign[1] ^ wSign[2];
    f2 = wSign[2] ^ (tri0 ? wSign[0] : wSign[3]);
    f3 = wSign[0] ^ wSign[3];
  }
  normalize_edge(clip, &eX[0], &eY[0], f0);
  normalize_edge(clip, &eX[1], &eY[1], f1);
  normalize_edge(clip, &eX[2], &eY[2], f2);
  normalize_edge(clip, &eX[3], &eY[3], f3);

  /* Edge offsets at the rectangle origin (:1030-1045). */
  const float maxOffset = -kMinEdgeOffset;
  const float add = 0.5f - kMinEdgeOffset * (64 - 1) / (maxOffset - kMinEdgeOffset);
  float o0 = fmaf(-x0, eX[0], fmaf(-y0, eY[0], add));
  float o1 = fmaf(-x[1], eX[1], fmaf(-y[1], eY[1], add));
  float o2 = fmaf(-x[2], eX[2], fmaf(-y[2], eY[2], add));
  float o3 = fmaf(-x[3], eX[3], fmaf(-y[3], eY[3], add));
  o1 = fmaf((float)minX, eX[1], o1);
  o2 = fmaf((float)minX, eX[2], o2);
  o3 = fmaf((float)minX, eX[3], o3);
  o1 = fmaf((float)minY, eY[1], o1);
  o2 = fmaf((float)minY, eY[2], o2);
  o3 = fmaf((float)minY, eY[3], o3);
  Q->off[0] = o0; Q->off[1] = o1; Q->off[2] = o2; Q->off[3] = o3;
  for (int e = 0; e < 4; ++e) {
    Q->nx[e] = eX[e];
    Q->ny[e] = eY[e];
    Q->slope[e] = quantize_slope(eX[e], eY[e]);
  }
  /* 16-bit multiply of the block index (:1054). */
  Q->firstBlock = (((uint32_t)minY * r->blocksX) & 0xFFFFu) + (uint32_t)minX;
}

static inline int mask_bit(int x, int y) { return 8 * x + ((y & 1) ? (y - 1) / 2 : 4 + y / 2); }

/* Block traversal for one quad (Rasterizer.cpp:1095-1292). */
static void traverse_quad(OrRaster* r, const QuadSetup* Q) {
  const float dDx = Q->dpDx, dDy = Q->dpDy;
  const float sp = -0.5f + 1.0f / 16.0f;
  float lineDepth[8];
  for (int k = 0; k < 8; ++k) {
    float sx = sp + 0.125f * (float)(k & 3);
    float sy = sp + 0.125f * (float)(k >> 2);
    lineDepth[k] = fmaf(dDx, sx, fmaf(dDy, sy, Q->dp0));
  }
  float lineOffset[4] = {Q->off[0], Q->off[1], Q->off[2], Q->off[3]};
  const uint32_t bX = r->blocksX;
  for (int32_t by = 0; by < Q->rangeY; ++by) {
    float offset[4], depth[8];
    memcpy(offset, lineOffset, sizeof offset);
    memcpy(depth, lineDepth, sizeof depth);
    int anyBlockHit = 0;
    for (int32_t bx = 0; bx < Q->rangeX; ++bx) {
      uint32_t blk = Q->firstBlock + (uint32_t)by * bX + (uint32_t)bx;
      uint16_t* hizp = &r->hiz[blk];
      uint16_t* out = &r->depth[(size_t)blk * 64];
      uint16_t hz = *hizp;
      r->stats[0]++;
      if (r->blockVisits) r->blockVisits[2 * blk]++;
      uint64_t mask = 0;
      int skip = 0;
      if (hz >= Q->maxZ) {
        r->stats[1]++;
        skip = 1;
      } else if (Q->mode == Convex) {
        if (offset[0] >= 63.0f || offset[1] >= 63.0f || offset[2] >= 63.0f || offset[3] >= 63.0f) {
          r->stats[2]++;
          if (anyBlockHit) break;
          skip = 1;
        } else {
          anyBlockHit = 1;
          uint64_t A[4];
          for (int e = 0; e < 4; ++e) {
            int32_t o = x_cvtt(offset[e]);
            if (o < 0) o = 0;
            A[e] = r->table[Q->slope[e] | (uint32_t)o];
          }
          mask = (A[0] & A[1]) & (A[2] & A[3]);
        }
      } else {
        uint64_t A[4];
        for (int e = 0; e < 4; ++e) {
          int32_t o = x_cvtt(offset[e]);
          if (o < 0) o = 0;
          if (o > 63) o = 63;
          A[e] = r->table[Q->slope[e] | (uint32_t)o];
        }
        switch (Q->mode) {
          case Triangle0: mask = A[0] & A[1] & A[2]; break;
          case Triangle1: mask = A[0] & A[2] & A[3]; break;
          case ConcaveRight: mask = (A[0] | A[3]) & (A[1] & A[2]); break;
          case ConcaveCenter: mask = (A[0] & A[1]) | (A[2] & A[3]); break;
          default: mask = (A[0] & A[3]) & (A[1] | A[2]); break; /* ConcaveLeft */
        }
        if (!mask) {
          r->stats[3]++;
          skip = 1;
        }
      }
      if (!skip) {
        uint16_t lo[2][8], hi[2][8];
        for (int p = 0; p < 2; ++p)
          for (int xx = 0; xx < 8; ++xx) {
            float v = depth[(xx & 3) + 4 * p];
            if (xx >= 4) v = fmaf(dDx, 0.5f, v);
            lo[p][xx] = pack_depth(v);
            hi[p][xx] = pack_depth(dDy + v);
          }
        uint16_t rows[8][8];
        for (int p = 0; p < 2; ++p)
          for (int xx = 0; xx < 8; ++xx) {
            uint16_t d0 = lo[p][xx], d4 = hi[p][xx];
            uint16_t d2 = avg_u16(d0, d4);
            uint16_t d1 = avg_u16(d0, d2);
            uint16_t d3 = avg_u16(d2, d4);
            rows[0 + p][xx] = d0;
            rows[2 + p][xx] = d1;
            rows[4 + p][xx] = d2;
            rows[6 + p][xx] = d3;
          }
        uint16_t mn = 0xFFFF;
        for (int yy = 0; yy < 8; ++yy)
          for (int xx = 0; xx < 8; ++xx) {
            uint16_t d = rows[yy][xx];
            if (!((mask >> mask_bit(xx, yy)) & 1)) d = 0;
            uint32_t* w = r->winner ? &r->winner[(size_t)blk * 64 + 8 * yy + xx] : NULL;
            if (hz != 1) {
              uint16_t old = out[8 * yy + xx];
              if (w && d > old) *w = r->curBatch;
              if (old > d) d = old;
            } else if (w) {
              /* First write of a cleared block: every pixel is replaced. */
              *w = d ? r->curBatch : 0xFFFFFFFFu;
            }
            out[8 * yy + xx] = d;
            if (d < mn) mn = d;
          }
        *hizp = mn;
        r->stats[4]++;
        if (r->blockVisits) r->blockVisits[2 * blk + 1]++;
      }
      for (int k = 0; k < 8; ++k) depth[k] = dDx + depth[k];
      for (int e = 0; e < 4; ++e) offset[e] = Q->nx[e] + offset[e];
    }
    for (int k = 0; k < 8; ++k) lineDepth[k] = lineDepth[k] + dDy;
    for (int e = 0; e < 4; ++e) lineOffset[e] = lineOffset[e] + Q->ny[e];
  }
}

void or_prologue(const OrRaster* r, const float refMin[4], const float refMax[4], float out[18]) {
  Prologue P;
  prologue(r, refMin, refMax, &P);
  for (int o = 0; o < 4; ++o)
    for (int j = 0; j < 4; ++j) out[4 * o + j] = P.T[o][j];
  out[16] = P.c0;
  out[17] = P.c1;
}

int or_setup_dump(const OrRaster* r, const uint32_t* packets, uint32_t packetCount, const float refMin[4],
                  const float refMax[4], int clip, float* out) {
  Prologue P;
  prologue(r, refMin, refMax, &P);
  int n = 0;
  for (uint32_t p = 0; p + 4 <= packetCount; p += 4)
    for (int lane = 0; lane < 8; ++lane) {
      uint32_t I[4];
      for (int k = 0; k < 4; ++k) I[k] = packets[(size_t)(p + k) * 8 + lane];
      QuadSetup Q;
      setup_quad(r, &P, I, clip ? 1 : 0, &Q);
      float* o = out + 24 * n++;
      o[0] = (float)Q.valid; o[1] = (float)Q.mode; o[2] = (float)Q.minX; o[3] = (float)Q.minY;
      o[4] = (float)Q.rangeX; o[5] = (float)Q.rangeY; o[6] = (float)Q.maxZ; o[7] = Q.dp0;
      for (int e = 0; e < 4; ++e) { o[8 + e] = Q.off[e]; o[12 + e] = Q.nx[e]; o[16 + e] = Q.ny[e]; }
      o[20] = Q.dpDx; o[21] = Q.dpDy; o[22] = (float)Q.firstBlock; o[23] = 0;
    }
  return n;
}

void or_rasterize(OrRaster* r, const uint32_t* packets, uint32_t packetCount, const float refMin[4],
                  const float refMax[4], int clip) {
  Prologue P;
  prologue(r, refMin, refMax, &P);
  for (uint32_t p = 0; p + 4 <= packetCount; p += 4) {
    for (int lane = 0; lane < 8; ++lane) {
      uint32_t I[4];
      for (int k = 0; k < 4; ++k) I[k] = packets[(size_t)(p + k) * 8 + lane];
      QuadSetup Q;
      setup_quad(r, &P, I, clip ? 1 : 0, &Q);
      r->stats[5]++;
      if (!Q.valid) continue;
      r->stats[6]++;
      traverse_quad(r, &Q);
    }
  }
}

/* ---------------------------------------------------------------------------
 * queryVisibility / query2D (Rasterizer.cpp:123-349)
 * ------------------------------------------------------------------------- */
int or_query2d(const OrRaster* r, uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY, uint32_t maxZ) {
  uint32_t blockMinX = minX / 8, blockMaxX = maxX / 8;
  uint32_t blockMinY = minY / 8, blockMaxY = maxY / 8;
  for (uint32_t blockY = blockMinY; blockY <= blockMaxY; ++blockY) {
    int32_t sy = (int32_t)(minY - 8 * blockY);
    int32_t ey = (int32_t)(maxY - 8 * blockY);
    uint32_t startY = (uint32_t)(sy > 0 ? sy : 0);
    uint32_t endY = (uint32_t)(ey < 7 ? ey : 7);
    int interiorLine = (startY == 0) && (endY == 7);
    for (uint32_t blockX = blockMinX; blockX <= blockMaxX; ++blockX) {
      uint32_t b = blockY * r->blocksX + blockX;
      if (maxZ <= r->hiz[b]) continue;
      int32_t sx = (int32_t)(minX - blockX * 8);
      int32_t ex = (int32_t)(maxX - blockX * 8);
      uint32_t startX = (uint32_t)(sx > 0 ? sx : 0);
      uint32_t endX = (uint32_t)(ex < 7 ? ex : 7);
      if (interiorLine && startX == 0 && endX == 7) return 1;
      uint16_t rowSelector = (uint16_t)((0xFFFFu << (2 * startX)) & (0xFFFFu >> (2 * (7 - endX))));
      const uint16_t* d = &r->depth[(size_t)b * 64];
      for (uint32_t y = startY; y <= endY; ++y) {
        uint32_t vis = 0;
        for (int x = 0; x < 8; ++x) {
          uint16_t v = d[8 * y + x];
          uint16_t mn = v < maxZ ? v : (uint16_t)maxZ;
          if (mn != (uint16_t)maxZ) vis |= 3u << (2 * x);
        }
        if (rowSelector & vis) return 1;
      }
    }
  }
  return 0;
}

/* queryVisibility; with depthTest == 0 it stops before query2D and returns 1
 * for every box that reaches it (the depth-independent classification). */
static int query_impl_ex(const OrRaster* r, const float bmin[4], const float bmax[4], int* needsClipping,
                         int depthTest) {
  float ext[4], cen[4];
  for (int j = 0; j < 4; ++j) {
    ext[j] = bmax[j] - bmin[j];
    cen[j] = bmax[j] + bmin[j];
  }
  const float* row0 = r->mvpRaw + 0;
  const float* row1 = r->mvpRaw + 4;
  const float* row2 = r->mvpRaw + 8;
  const float* row3 = r->mvpRaw + 12;
  /* Frustum planes (:135-167). */
  uint32_t signs = 0;
  for (int pl = 0; pl < 6; ++pl) {
    const float* rr = pl < 2 ? row0 : (pl < 4 ? row1 : row2);
    float plane[4], off[4];
    for (int j = 0; j < 4; ++j) {
      plane[j] = (pl & 1) ? (row3[j] - rr[j]) : (row3[j] + rr[j]);
      off[j] = cen[j] + xor_sign(ext[j], fbits(plane[j]) & 0x80000000u);
    }
    signs |= fbits(x_dp4(plane, off));
  }
  if (signs & 0x80000000u) return 0;

  /* Corners (:169-198). */
  const float* col0 = r->mvp + 0;
  const float* col1 = r->mvp + 4;
  const float* col2 = r->mvp + 8;
  const float* col3 = r->mvp + 12;
  float e0[4], e1[4], e2[4], c[8][4];
  for (int j = 0; j < 4; ++j) {
    e0[j] = col0[j] * ext[0];
    e1[j] = col1[j] * ext[1];
    e2[j] = col2[j] * ext[2];
    c[0][j] = fmaf(col0[j], bmin[0], fmaf(col1[j], bmin[1], fmaf(col2[j], bmin[2], col3[j])));
  }
  for (int j = 0; j < 4; ++j) {
    c[1][j] = c[0][j] + e0[j];
    c[2][j] = c[0][j] + e1[j];
    c[4][j] = c[0][j] + e2[j];
  }
  for (int j = 0; j < 4; ++j) {
    c[3][j] = c[1][j] + e1[j];
    c[5][j] = c[4][j] + e0[j];
    c[6][j] = c[2][j] + e2[j];
  }
  for (int j = 0; j < 4; ++j) c[7][j] = c[6][j] + e0[j];

  /* Near-plane epsilon per lane (:204-213). */
  float me1[4], me2[4];
  for (int j = 0; j < 4; ++j) me1[j] = x_max(ext[j], ext[j ^ 2]);
  for (int j = 0; j < 4; ++j) me2[j] = x_max(me1[j], me1[j ^ 1]);
  int close = 0;
  for (int j = 0; j < 4; ++j) {
    float eps = me2[j] * 0.001f;
    if (c[j][3] < eps || c[4 + j][3] < eps) close = 1;
  }
  if (close) {
    *needsClipping = 1;
    return 1;
  }
  *needsClipping = 0;

  /* Projection and bounds (:217-264). */
  float px[8], py[8], pz[8];
  for (int k = 0; k < 8; ++k) {
    float rw = x_rcp128(c[k][3]);
    px[k] = c[k][0] * rw;
    py[k] = c[k][1] * rw;
    pz[k] = c[k][2] * rw;
  }
  float minsX[4], maxsX[4], minsY[4], maxsY[4];
  for (int j = 0; j < 4; ++j) {
    minsX[j] = x_min(px[j], px[4 + j]);
    maxsX[j] = x_max(px[j], px[4 + j]);
    minsY[j] = x_min(py[j], py[4 + j]);
    maxsY[j] = x_max(py[j], py[4 + j]);
  }
  float mXY[4] = {x_min(minsX[0], minsX[2]), x_min(minsY[0], minsY[2]), x_min(minsX[1], minsX[3]),
                  x_min(minsY[1], minsY[3])};
  float MXY[4] = {x_max(maxsX[0], maxsX[2]), x_max(maxsY[0], maxsY[2]), x_max(maxsX[1], maxsX[3]),
                  x_max(maxsY[1], maxsY[3])};
  const float lim[4] = {(float)(r->width - 1), (float)(r->height - 1), (float)(r->width - 1), (float)(r->height - 1)};
  for (int j = 0; j < 4; ++j) {
    mXY[j] = x_max(mXY[j], 0.0f);
    MXY[j] = x_min(MXY[j], lim[j]);
    MXY[j] = -MXY[j];
  }
  float bf[4] = {x_min(mXY[0], mXY[2]), x_min(MXY[0], MXY[2]), x_min(mXY[1], mXY[3]), x_min(MXY[1], MXY[3])};
  int32_t bi[4];
  for (int j = 0; j < 4; ++j) bi[j] = x_cvtt(x_floor(bf[j]));
  bi[1] = -bi[1];
  bi[3] = -bi[3];
  if (bi[0] >= bi[1] || bi[2] >= bi[3]) return 0;
  if (!depthTest) return 1;

  /* Max depth of the corners (:271-273). */
  uint16_t maxZ = 0;
  for (int k = 0; k < 8; ++k) {
    uint16_t d = pack_depth(pz[k]);
    if (d > maxZ) maxZ = d;
  }
  return or_query2d(r, (uint32_t)bi[0], (uint32_t)bi[1], (uint32_t)bi[2], (uint32_t)bi[3], maxZ);
}

static int query_impl(const OrRaster* r, const float bmin[4], const float bmax[4], int* needsClipping) {
  return query_impl_ex(r, bmin, bmax, needsClipping, 1);
}

int or_query(OrRaster* r, const float bmin[4], const float bmax[4], int* needsClipping) {
  return query_impl(r, bmin, bmax, needsClipping);
}

void or_query_batch(const OrRaster* r, uint32_t n, const float* bounds, uint8_t* visible, uint8_t* needsClipping) {
  for (uint32_t i = 0; i < n; ++i) {
    int clip = 0;
    visible[i] = (uint8_t)query_impl(r, bounds + 8 * (size_t)i, bounds + 8 * (size_t)i + 4, &clip);
    if (needsClipping) needsClipping[i] = (uint8_t)clip;
  }
}

/* readBackDepth (Rasterizer.cpp:351-399). */
void or_readback(const OrRaster* r, uint8_t* rgba) {
  const float bias = 3.9623753e+28f;
  for (uint32_t blockY = 0; blockY < r->blocksY; ++blockY)
    for (uint32_t blockX = 0; blockX < r->blocksX; ++blockX) {
      uint32_t b = blockY * r->blocksX + blockX;
      for (uint32_t y = 0; y < 8; ++y) {
        uint8_t* dest = rgba + 4 * (8 * blockX + r->width * (8 * blockY + y));
        if (r->hiz[b] == 1) {
          memset(dest, 0, 32);
          continue;
        }
        for (uint32_t x = 0; x < 8; ++x) {
          uint16_t v = r->depth[(size_t)b * 64 + 8 * y + x];
          float f = bitsf((uint32_t)v << 12) * bias;
          float l = (2 * 0.25f) / ((0.25f + 1000.0f) - (1.0f - f) * (1000.0f - 0.25f));
          uint32_t d = (uint32_t)(int64_t)(100 * 256 * l);
          dest[4 * x + 0] = (uint8_t)(d / 100);
          dest[4 * x + 1] = (uint8_t)(d % 256);
          dest[4 * x + 2] = 0;
          dest[4 * x + 3] = 255;
        }
      }
    }
}

void or_get_raw(const OrRaster* r, uint16_t* depth, uint16_t* hiz) {
  size_t nb = (size_t)r->blocksX * r->blocksY;
  if (depth) memcpy(depth, r->depth, nb * 64 * sizeof(uint16_t));
  if (hiz) memcpy(hiz, r->hiz, (nb + 8) * sizeof(uint16_t));
}
void or_set_raw(OrRaster* r, const uint16_t* depth, const uint16_t* hiz) {
  size_t nb = (size_t)r->blocksX * r->blocksY;
  if (depth) memcpy(r->depth, depth, nb * 64 * sizeof(uint16_t));
  if (hiz) memcpy(r->hiz, hiz, (nb + 8) * sizeof(uint16_t));
}
void or_get_table(const OrRaster* r, uint64_t* t) { memcpy(t, r->table, sizeof r->table); }
void or_get_mvp(const OrRaster* r, float prebaked[16], float raw[16]) {
  if (prebaked) memcpy(prebaked, r->mvp, 64);
  if (raw) memcpy(raw, r->mvpRaw, 64);
}
void or_get_stats(const OrRaster* r, uint64_t s[7]) { memcpy(s, r->stats, sizeof r->stats); }
void or_reset_stats(OrRaster* r) {
  memset(r->stats, 0, sizeof r->stats);
  if (r->blockVisits) memset(r->blockVisits, 0, sizeof(uint32_t) * 2 * r->blocksX * r->blocksY);
}

void or_block_stats(OrRaster* r, uint32_t* out) {
  size_t nb = (size_t)r->blocksX * r->blocksY;
  if (!r->blockVisits) r->blockVisits = (uint32_t*)calloc(2 * nb, sizeof(uint32_t));
  if (out && r->blockVisits) memcpy(out, r->blockVisits, sizeof(uint32_t) * 2 * nb);
}

/* Main.cpp:192-206. */
void or_frame(OrRaster* r, uint32_t n, const OrOccluder* occ, uint8_t* status) {
  for (uint32_t i = 0; i < n; ++i) {
    int clip = 0;
    uint8_t st = 0;
    if (query_impl(r, occ[i].boundsMin, occ[i].boundsMax, &clip)) {
      or_rasterize(r, occ[i].packets, occ[i].packetCount, occ[i].refMin, occ[i].refMax, clip);
      st = clip ? 2 : 1;
    }
    if (status) status[i] = st;
  }
}

void or_frame_labels(OrRaster* r, uint32_t n, const OrOccluder* occ, const uint32_t* order, uint8_t* status,
                     uint32_t* labels) {
  uint32_t nb = 0;
  for (uint32_t i = 0; i < n; ++i) {
    int clip = 0;
    status[i] = query_impl_ex(r, occ[i].boundsMin, occ[i].boundsMax, &clip, 0) ? (clip ? 2 : 1) : 0;
    nb += occ[i].packetCount / 4;
  }
  /* batch k -> (occluder, first packet) */
  uint32_t* owner = (uint32_t*)malloc((nb ? nb : 1) * sizeof(uint32_t));
  uint32_t* first = (uint32_t*)malloc((nb ? nb : 1) * sizeof(uint32_t));
  for (uint32_t i = 0, k = 0; i < n; ++i)
    for (uint32_t b = 0; b < occ[i].packetCount / 4; ++b, ++k) {
      owner[k] = i;
      first[k] = 4 * b;
    }
  const size_t blocks = (size_t)r->blocksX * r->blocksY;
  uint32_t* winner = (uint32_t*)malloc(blocks * 64 * sizeof(uint32_t));
  for (size_t i = 0; i < blocks * 64; ++i) winner[i] = 0xFFFFFFFFu;
  r->winner = winner;
  for (uint32_t t = 0; t < nb; ++t) {
    const uint32_t k = order ? order[t] : t;
    const OrOccluder* o = &occ[owner[k]];
    if (!status[owner[k]]) continue;
    r->curBatch = k;
    or_rasterize(r, o->packets + (size_t)first[k] * 8, 4, o->refMin, o->refMax, status[owner[k]] == 2);
  }
  r->winner = NULL;
  for (uint32_t k = 0; k < nb; ++k) labels[k] = 0;
  for (size_t b = 0; b < blocks; ++b) {
    if (r->hiz[b] == 1) continue; /* not written this frame: winners are stale */
    /* Distinct winners of the block, each counted once. */
    for (int p = 0; p < 64; ++p) {
      uint32_t w = winner[b * 64 + p];
      if (w == 0xFFFFFFFFu) continue;
      int seen = 0;
      for (int q = 0; q < p && !seen; ++q) seen = winner[b * 64 + q] == w;
      if (!seen) labels[w]++;
    }
  }
  free(winner);
  free(owner);
  free(first);
}

void or_frame_classified(OrRaster* r, uint32_t n, const OrOccluder* occ, uint8_t* status) {
  for (uint32_t i = 0; i < n; ++i) {
    int clip = 0;
    uint8_t st = 0;
    if (query_impl_ex(r, occ[i].boundsMin, occ[i].boundsMax, &clip, 0)) {
      or_rasterize(r, occ[i].packets, occ[i].packetCount, occ[i].refMin, occ[i].refMax, clip);
      st = clip ? 2 : 1;
    }
    if (status) status[i] = st;
  }
}

/* ---------------------------------------------------------------------------
 * Occluder::bake (Occluder.cpp:7-181) with VectorMath.h cross/normal/normalize.
 * ------------------------------------------------------------------------- */
static void v_cross(const float a[4], const float b[4], float out[4]) {
  out[0] = a[1] * b[2] - a[2] * b[1];
  out[1] = a[2] * b[0] - a[0] * b[2];
  out[2] = a[0] * b[1] - a[1] * b[0];
  out[3] = a[3] * b[3] - a[3] * b[3];
}
static void v_normal(const float* v0, const float* v1, const float* v2, float out[4]) {
  float a[4], b[4];
  for (int j = 0; j < 4; ++j) {
    a[j] = v1[j] - v0[j];
    b[j] = v2[j] - v0[j];
  }
  v_cross(a, b, out);
}
/* normalize: v * rsqrt(dot3(v, v)) (VectorMath.h:20-23). */
static void v_normalize(float v[4]) {
  float rs = x_rsqrt(x_dp3(v, v));
  for (int j = 0; j < 4; ++j) v[j] = v[j] * rs;
}

/* Occluder.cpp:11-93: per-quad normals, the 6-centroid k-means by normal
 * direction and the stable regroup; order[i] = the quad placed i-th. */
static int bake_order(const float* vertices, uint32_t nq, uint32_t* order) {
  float* normals = (float*)malloc(sizeof(float) * 4 * nq);
  uint32_t* assign = (uint32_t*)calloc(nq, sizeof(uint32_t));
  if (!normals || !assign) {
    free(normals); free(assign);
    return -2;
  }
  for (uint32_t q = 0; q < nq; ++q) {
    const float* v0 = vertices + 16 * q;
    float a[4], b[4];
    v_normal(v0, v0 + 4, v0 + 8, a);
    v_normal(v0, v0 + 8, v0 + 12, b);
    for (int j = 0; j < 4; ++j) normals[4 * q + j] = a[j] + b[j];
    v_normalize(normals + 4 * q);
  }
  float cent[6][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, -1, 0, 0}, {0, 0, -1, 0}, {-1, 0, 0, 0}};
  int anyChanged = 1;
  for (int iter = 0; iter < 10 && anyChanged; ++iter) {
    anyChanged = 0;
    for (uint32_t q = 0; q < nq; ++q) {
      float best = -INFINITY;
      uint32_t bestK = 0;
      for (uint32_t k = 0; k < 6; ++k) {
        float d = x_dp3(cent[k], normals + 4 * q);
        if (_mm_comige_ss(_mm_set_ss(d), _mm_set_ss(best))) {
          best = d;
          bestK = k;
        }
      }
      if (assign[q] != bestK) {
        assign[q] = bestK;
        anyChanged = 1;
      }
    }
    memset(cent, 0, sizeof cent);
    for (uint32_t q = 0; q < nq; ++q)
      for (int j = 0; j < 4; ++j) cent[assign[q]][j] = cent[assign[q]][j] + normals[4 * q + j];
    for (int k = 0; k < 6; ++k) v_normalize(cent[k]);
  }
  uint32_t cnt = 0;
  for (uint32_t k = 0; k < 6; ++k)
    for (uint32_t q = 0; q < nq; ++q)
      if (assign[q] == k) order[cnt++] = q;
  free(normals);
  free(assign);
  return 0;
}

int or_bake_order(const float* vertices, uint32_t nQuads, uint32_t* orderOut) {
  if (nQuads == 0) return -1;
  return bake_order(vertices, nQuads, orderOut);
}

int or_bake(const float* vertices, uint32_t n, const float refMin[4], const float refMax[4], uint32_t* packetsOut,
            float boundsMin[4], float boundsMax[4], float center[4]) {
  if (n == 0 || n % 32) return -1;
  uint32_t nq = n / 4;
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * nq);
  if (!order) return -2;
  if (bake_order(vertices, nq, order) != 0) {
    free(order);
    return -2;
  }

  float invExt[4];
  for (int j = 0; j < 4; ++j) invExt[j] = 1.0f / (refMax[j] - refMin[j]);
  const float scale[3] = {2047.0f, 2047.0f, 1023.0f};
  for (uint32_t batch = 0; batch < nq / 8; ++batch) {
    for (int k = 0; k < 4; ++k)
      for (int lane = 0; lane < 8; ++lane) {
        const float* v = vertices + 16 * order[8 * batch + lane] + 4 * k;
        float t[3];
        for (int j = 0; j < 3; ++j) t[j] = fmaf((v[j] - refMin[j]) * invExt[j], scale[j], 0.5f);
        int32_t X = x_cvtt(t[0]) - 1024;
        int32_t Y = x_cvtt(t[1]);
        int32_t Z = x_cvtt(t[2]);
        packetsOut[(size_t)(4 * batch + k) * 8 + lane] = ((uint32_t)X << 21) | ((uint32_t)Y << 10) | (uint32_t)Z;
      }
  }
  float mn[4] = {INFINITY, INFINITY, INFINITY, INFINITY}, mx[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < n; ++i)
    for (int j = 0; j < 4; ++j) {
      mn[j] = x_min(vertices[4 * i + j], mn[j]);
      mx[j] = x_max(vertices[4 * i + j], mx[j]);
    }
  mn[3] = 1.0f;
  mx[3] = 1.0f;
  for (int j = 0; j < 4; ++j) {
    boundsMin[j] = mn[j];
    boundsMax[j] = mx[j];
    center[j] = (mx[j] + mn[j]) * 0.5f;
  }
  free(order);
  return 0;
}

/* ---------------------------------------------------------------------------
 * CPU baselines (one oracle instance per thread).
 * ------------------------------------------------------------------------- */
typedef struct {
  uint32_t w, h, n, frames, threads, tid;
  const OrOccluder* occ;
  const float* mvps;
  const uint32_t* order;
  OrOccluder* scratch;
} FrameJob;

static void* frame_worker(void* arg) {
  FrameJob* j = (FrameJob*)arg;
  OrRaster* r = or_create(j->w, j->h);
  if (!r) return NULL;
  for (uint32_t f = j->tid; f < j->frames; f += j->threads) {
    for (uint32_t i = 0; i < j->n; ++i) j->scratch[i] = j->occ[j->order[(size_t)f * j->n + i]];
    or_clear(r);
    or_set_mvp(r, j->mvps + 16 * (size_t)f);
    or_frame(r, j->n, j->scratch, NULL);
  }
  or_destroy(r);
  return NULL;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

double or_bench_frames(uint32_t w, uint32_t h, uint32_t n, const OrOccluder* occ, uint32_t frames, const float* mvps,
                       const uint32_t* order, uint32_t threads) {
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
  FrameJob* jobs = (FrameJob*)calloc(threads, sizeof(FrameJob));
  for (uint32_t t = 0; t < threads; ++t) {
    jobs[t] = (FrameJob){w, h, n, frames, threads, t, occ, mvps, order, (OrOccluder*)malloc(sizeof(OrOccluder) * n)};
  }
  double t0 = now_s();
  for (uint32_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, frame_worker, &jobs[t]);
  for (uint32_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  double t1 = now_s();
  for (uint32_t t = 0; t < threads; ++t) free(jobs[t].scratch);
  free(jobs);
  free(th);
  return t1 - t0;
}

typedef struct {
  const OrRaster* r;
  uint32_t nq, reps, threads, tid;
  const float* bounds;
} QueryJob;

static void* query_worker(void* arg) {
  QueryJob* j = (QueryJob*)arg;
  volatile int sink = 0;
  for (uint32_t rep = 0; rep < j->reps; ++rep)
    for (uint32_t q = j->tid; q < j->nq; q += j->threads) {
      int clip = 0;
      sink += query_impl(j->r, j->bounds + 8 * (size_t)q, j->bounds + 8 * (size_t)q + 4, &clip);
    }
  (void)sink;
  return NULL;
}

double or_bench_queries(const OrRaster* r, uint32_t nq, const float* bounds, uint32_t reps, uint32_t threads) {
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
  QueryJob* jobs = (QueryJob*)calloc(threads, sizeof(QueryJob));
  for (uint32_t t = 0; t < threads; ++t) jobs[t] = (QueryJob){r, nq, reps, threads, t, bounds};
  double t0 = now_s();
  for (uint32_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, query_worker, &jobs[t]);
  for (uint32_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  double t1 = now_s();
  free(jobs);
  free(th);
  return t1 - t0;
}
