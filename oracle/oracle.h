/*
 * oracle.h -- CPU oracle for the occlusion-culling hot path of
 * rawrunprotected/rasterizer (SoftwareRasterizer/Rasterizer.cpp, Occluder.cpp).
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (rasterizer_amd) never links or calls it.
 *
 * It is a scalar restatement (one quad / one AABB at a time) of the
 * reference's 8-wide AVX2 code.  The x86 primitives whose results are
 * implementation-defined or approximate (RCPPS, RSQRTPS, CVTTPS2DQ, MINPS /
 * MAXPS operand order, DPPS summation order, ROUNDPS) are executed with the
 * host's own SSE/AVX instructions, so on a given host the oracle reproduces
 * the reference's arithmetic lane for lane.  Everything else is plain IEEE
 * float32 with contraction off and fmaf() exactly where the reference calls
 * _mm*_fmadd/_fmsub/_fnmadd.
 *
 * The reference itself cannot be compiled in this image: it includes
 * MSVC's <intrin.h> and calls _aligned_malloc / _BitScanForward /
 * __debugbreak, which would need stand-in headers.
 */
#ifndef RZ_ORACLE_H
#define RZ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrRaster OrRaster;

/* Rasterizer::Rasterizer (Rasterizer.cpp:66-74): allocates depth + hiZ and
 * builds the 4096-entry coverage table (Rasterizer.cpp:547-604). */
OrRaster* or_create(uint32_t width, uint32_t height);
void or_destroy(OrRaster* r);

/* setModelViewProjection (Rasterizer.cpp:76-105). */
void or_set_mvp(OrRaster* r, const float matrix[16]);
/* clear (Rasterizer.cpp:107-121): hiZ := 1, depth untouched. */
void or_clear(OrRaster* r);

/* rasterize<possiblyNearClipped> (Rasterizer.cpp:606-1295).  packets are the
 * Occluder::m_vertexData words: packet k of batch b holds vertex k of quads
 * 8b..8b+7, i.e. word[(4b+k)*8 + lane]. */
void or_rasterize(OrRaster* r, const uint32_t* packets, uint32_t packetCount,
                  const float refMin[4], const float refMax[4], int possiblyNearClipped);

/* Debug: per-quad setup of rasterize<clip> (24 floats per quad: valid, mode,
 * minX, minY, rangeX, rangeY, maxZ, dp0, off[4], nx[4], ny[4], dDx, dDy,
 * firstBlock, 0).  Returns the number of quads. */
/* Rasterizer.cpp:616-655 for one occluder: the folded matrix T[4][4]
 * (row-major) and c0, c1. */
void or_prologue(const OrRaster* r, const float refMin[4], const float refMax[4], float out[18]);
int or_setup_dump(const OrRaster* r, const uint32_t* packets, uint32_t packetCount, const float refMin[4],
                  const float refMax[4], int clip, float* out);

/* queryVisibility (Rasterizer.cpp:123-281).  Returns visible; *needsClipping
 * is written only when the box survives the frustum test (as the reference). */
int or_query(OrRaster* r, const float boundsMin[4], const float boundsMax[4], int* needsClipping);
/* queryVisibility over n boxes (8 floats each: min xyzw, max xyzw);
 * needsClipping is 0 for frustum-culled boxes. */
void or_query_batch(const OrRaster* r, uint32_t n, const float* bounds, uint8_t* visible, uint8_t* needsClipping);
/* query2D (Rasterizer.cpp:283-349). */
int or_query2d(const OrRaster* r, uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY, uint32_t maxZ);

/* readBackDepth (Rasterizer.cpp:351-399): width*height*4 bytes. */
void or_readback(const OrRaster* r, uint8_t* rgba);

/* Raw state access (tests only). depth: blocksX*blocksY*64 u16 (block-linear,
 * 8 rows of 8 pixels per block); hiz: blocksX*blocksY + 8 u16. */
void or_get_raw(const OrRaster* r, uint16_t* depth, uint16_t* hiz);
void or_set_raw(OrRaster* r, const uint16_t* depth, const uint16_t* hiz);
void or_get_table(const OrRaster* r, uint64_t* table4096);
void or_get_mvp(const OrRaster* r, float prebaked[16], float raw[16]);

/* Counters of the block loop (Rasterizer.cpp:1140-1290) since the last reset:
 * [0] (quad, block) pairs visited, [1] rejected by hiZ, [2] rejected by the
 * Convex offset test, [3] empty non-convex masks, [4] blocks written,
 * [5] quads set up, [6] quads valid after cull. */
void or_get_stats(const OrRaster* r, uint64_t stats[7]);
void or_reset_stats(OrRaster* r);
/* Per-block counters (2 u32 per block: visited, written); the first call
 * enables counting and may pass out = NULL. */
void or_block_stats(OrRaster* r, uint32_t* out);

/* One occluder of a frame: its baked packets, ref box and bounds. */
typedef struct OrOccluder {
  const uint32_t* packets;
  uint32_t packetCount;
  float refMin[4], refMax[4];
  float boundsMin[4], boundsMax[4];
} OrOccluder;

/* The query-gated raster loop of Main.cpp:192-206 over occluders already in
 * draw order.  status[i] = 0 not drawn, 1 rasterize<false>, 2 rasterize<true>. */
void or_frame(OrRaster* r, uint32_t n, const OrOccluder* occ, uint8_t* status);
/* The batched path's sequence for one view: each occluder is classified by
 * the depth-independent part of queryVisibility (Rasterizer.cpp:123-264:
 * frustum cull, near-plane test, empty screen rectangle) and rasterized with
 * the resulting needsClipping, with no occlusion test between draws. */
void or_frame_classified(OrRaster* r, uint32_t n, const OrOccluder* occ, uint8_t* status);
/* Labelled classify-then-draw of one frame (caller has cleared and set the
 * matrix): status[i] as or_frame_classified (0 not drawn, 1 rasterize<false>,
 * 2 rasterize<true>); then the 8-quad batches of the drawn occluders in
 * `order` (batch numbers counted in occluder order; NULL: occluder order).
 * Each pixel tracks the batch whose write produced its depth: the first write
 * of a cleared block (hiZ == 1) sets every pixel's winner (none where the
 * written depth is 0), a later write only where it raises the depth.  So the
 * winner is the first batch in submission order that wrote the final depth.
 * labels[k] = number of written blocks in which batch k wins some pixel. */
void or_frame_labels(OrRaster* r, uint32_t n, const OrOccluder* occ, const uint32_t* order, uint8_t* status,
                     uint32_t* labels);

/* Occluder::bake (Occluder.cpp:7-181).  vertices: n*4 floats, n a multiple
 * of 32 (8 quads).  packetsOut: n words.  Returns 0, or -1 on bad n. */
int or_bake(const float* vertices, uint32_t n, const float refMin[4], const float refMax[4],
            uint32_t* packetsOut, float boundsMin[4], float boundsMax[4], float center[4]);

/* Occluder.cpp:11-93 alone: the order in which bake places the quads (the
 * k-means groups by normal, each in quad order).  orderOut: nQuads words. */
int or_bake_order(const float* vertices, uint32_t nQuads, uint32_t* orderOut);

/* Host x86 primitives as used above (exposed for tests). */
float or_x86_rcp(float x);
int32_t or_x86_cvtt(float x);

/* CPU baseline: run `frames` frames (each: clear, set MVP, query-gated
 * raster of occluders in the given per-frame order) on `threads` threads,
 * each thread owning its own OrRaster (one instance per core).  Thread t
 * renders frames t, t+threads, ...  mvps: frames*16 floats; order:
 * frames*n indices.  Returns wall seconds. */
double or_bench_frames(uint32_t width, uint32_t height, uint32_t n, const OrOccluder* occ,
                       uint32_t frames, const float* mvps, const uint32_t* order, uint32_t threads);
/* CPU baseline for queries: `nq` AABBs (8 floats each: min xyzw, max xyzw)
 * against the current buffer of r, repeated `reps` times on `threads`
 * threads sharing r read-only.  Returns wall seconds. */
double or_bench_queries(const OrRaster* r, uint32_t nq, const float* bounds, uint32_t reps, uint32_t threads);

#ifdef __cplusplus
}
#endif
#endif
