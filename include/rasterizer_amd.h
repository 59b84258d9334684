/*
 * rasterizer_amd.h -- C ABI of the MI355X occlusion-culling rasterizer.
 *
 * Drop-in GPU replacement for the hot path of rawrunprotected/rasterizer
 * (SoftwareRasterizer/Rasterizer.h, Rasterizer.cpp, Occluder.h, Occluder.cpp).
 * Every entry point below cites the reference interface it replaces.  The C++
 * class API with the reference's exact signatures (Rasterizer, Occluder) sits
 * on top of this ABI in include/rasterizer_amd/Rasterizer.h and Occluder.h.
 *
 * Conventions
 *  - Every function returns an int status: RZ_OK (0) or a negative RZ_ERR_*.
 *  - Handles are opaque and destroyed by the caller.
 *  - A context is one Rasterizer: one depth buffer + hiZ on one GPU, bound to
 *    one HIP stream it creates.  A context is not thread-safe; independent
 *    contexts are.
 *  - rz_rasterize / rz_draw_frame / rz_clear / rz_set_mvp are stream-ordered
 *    and do not block the host.  rz_query*, rz_readback_* and rz_sync block
 *    until the stream has drained, and report any device-side error raised by
 *    earlier asynchronous calls.
 *  - Vectors are 4 floats (x, y, z, w), as the reference's __m128.
 *  - Nothing here falls back to the CPU: without a usable GPU, rz_create fails.
 */
#ifndef RASTERIZER_AMD_H
#define RASTERIZER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RZ_OK 0
#define RZ_ERR_INVALID_ARGUMENT (-1)
#define RZ_ERR_OUT_OF_MEMORY (-2)
#define RZ_ERR_DEVICE (-3)
#define RZ_ERR_TIMEOUT (-4)
#define RZ_ERR_UNSUPPORTED_HOST (-5)
#define RZ_ERR_TOO_LARGE (-6)
#define RZ_ERR_NOT_IDENTICAL (-7)
#define RZ_ERR_IO (-8)

typedef struct rz_context rz_context;
typedef struct rz_occluder rz_occluder;

/* Human-readable text for a status code (static storage). */
const char* rz_status_string(int status);

/* Rasterizer::Rasterizer(uint32_t width, uint32_t height)
 * (Rasterizer.h:13, Rasterizer.cpp:66-74, table build :547-604).
 * width and height must be non-zero multiples of 8.  Builds the 4096-entry
 * coverage table and the host CPU's RCPPS table on the host, uploads both, and
 * allocates depth (2 B/pixel, 8x8-block-linear) and hiZ (1 u16 per block + 8).
 * The depth buffer starts at 0 and hiZ at 0, as the reference's vectors do. */
int rz_create(uint32_t width, uint32_t height, rz_context** out);
int rz_destroy(rz_context* ctx);

/* Rasterizer::setModelViewProjection(const float* matrix)
 * (Rasterizer.h:15, Rasterizer.cpp:76-105).  16 floats, row-major for row
 * vectors (a DirectXMath XMMATRIX memcpy'd to float[16]). */
int rz_set_mvp(rz_context* ctx, const float matrix[16]);

/* Rasterizer::clear() (Rasterizer.h:17, Rasterizer.cpp:107-121).  Sets every
 * hiZ entry to 1 (the "cleared" flag); depth is left as it is. */
int rz_clear(rz_context* ctx);

/* Device copy of a baked Occluder (Occluder.h:7-21): its packets
 * (m_vertexData, 4 x 8 words per 8-quad batch), m_refMin/m_refMax and
 * m_boundsMin/m_boundsMax.  packetCount must be a multiple of 4.  The packets
 * are copied; the host array may be freed afterwards. */
int rz_occluder_create(rz_context* ctx, const uint32_t* packets, uint32_t packetCount, const float refMin[4],
                       const float refMax[4], const float boundsMin[4], const float boundsMax[4],
                       rz_occluder** out);
/* The same from packets already in device memory (d_packets, e.g. written by
 * rz_bake_device or rz_prepare_scene_device): copied device to device; the
 * boxes are host arrays. */
int rz_occluder_create_device(rz_context* ctx, const uint32_t* d_packets, uint32_t packetCount, const float refMin[4],
                              const float refMax[4], const float boundsMin[4], const float boundsMax[4],
                              rz_occluder** out);
int rz_occluder_destroy(rz_occluder* occluder);

/* template<bool possiblyNearClipped> Rasterizer::rasterize(const Occluder&)
 * (Rasterizer.h:19-20, Rasterizer.cpp:606-1295; instantiations :1297-1299).
 * Enqueued on the context stream; the occluder must outlive the call's
 * completion (rz_sync or any blocking call). */
int rz_rasterize(rz_context* ctx, const rz_occluder* occluder, int possiblyNearClipped);

/* The query-gated draw loop of Main.cpp:192-206, on the device in one pass:
 * for i in order: if queryVisibility(bounds_i, needsClipping) then
 * rasterize<needsClipping>(occluder_i).  Results are identical to issuing
 * those calls one by one.  statusOut (optional, n bytes, written when the
 * stream reaches it; valid after rz_sync): 0 not drawn, 1 drawn with
 * rasterize<false>, 2 drawn with rasterize<true>. */
int rz_draw_frame(rz_context* ctx, uint32_t n, const rz_occluder* const* occluders, uint8_t* statusOut);

/* Bound of each cross-tile wait of the query-gated draw, in ticks of the
 * device's 100 MHz s_memrealtime clock (0 restores the default, 2 s).  No
 * counterpart in the reference, whose loop is sequential.  A wait that runs
 * out ends every other wait of the launch, every kernel exits normally, the
 * frame's contents are unspecified and the next blocking call (rz_sync)
 * returns RZ_ERR_TIMEOUT. */
int rz_set_wait_timeout(rz_context* ctx, uint64_t ticks);

/* rz_draw_frame without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device or pinned memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * draw's working buffers (item descriptors and states, query done words,
 * quad setup records and boxes) live in one static device array of the
 * library, about 47 MB, that each GPU the library runs on holds until the
 * process ends.  Limits: n <= RZ_NOALLOC_MAX_ITEMS occluders and at most
 * RZ_NOALLOC_MAX_QUADS quads (8 per batch) in total, else RZ_ERR_TOO_LARGE
 * and nothing is enqueued.  Every call on one GPU shares that array: calls
 * whose draws overlap on one GPU (on different contexts or host threads)
 * race, so order them on the device (one stream, or events).  d_statusOut
 * (optional) is DEVICE memory of n bytes, written with rz_draw_frame's
 * statuses when the stream reaches the draw.  Results are identical to
 * rz_draw_frame's. */
#define RZ_NOALLOC_MAX_ITEMS 4096u
#define RZ_NOALLOC_MAX_QUADS 524288u
int rz_draw_frame_noalloc(rz_context* ctx, uint32_t n, const rz_occluder* const* occluders, uint8_t* d_statusOut);

/* Rasterizer::queryVisibility(__m128 boundsMin, __m128 boundsMax, bool& needsClipping)
 * (Rasterizer.h:22, Rasterizer.cpp:123-281).  Blocks.  *needsClipping is
 * written only when the box passes the frustum test, as in the reference. */
int rz_query(rz_context* ctx, const float boundsMin[4], const float boundsMax[4], int* visible,
             int* needsClipping);

/* Rasterizer::query2D(minX, maxX, minY, maxY, maxZ) (Rasterizer.h:24,
 * Rasterizer.cpp:283-349).  Blocks.  Pixel bounds are inclusive, maxX <
 * width, maxY < height. */
int rz_query2d(rz_context* ctx, uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY, uint32_t maxZ,
               int* visible);

/* Batched queryVisibility over n boxes, host memory: bounds is n x 8 floats
 * (min xyzw, max xyzw); visible and needsClipping are n bytes each
 * (needsClipping is 0 when the box was frustum-culled).  Blocks. */
int rz_query_batch(rz_context* ctx, uint32_t n, const float* bounds, uint8_t* visible, uint8_t* needsClipping);

/* Same on device-resident arrays (device pointers), enqueued without
 * blocking. */
int rz_query_batch_device(rz_context* ctx, uint32_t n, const float* d_bounds, uint8_t* d_visible,
                          uint8_t* d_needsClipping);

/* Rasterizer::readBackDepth(void* target) (Rasterizer.h:26,
 * Rasterizer.cpp:351-399).  width*height*4 bytes, row-major.  Blocks. */
int rz_readback_depth(rz_context* ctx, void* rgba);

/* Raw buffer access (the reference's private m_depthBuffer / m_hiZ):
 * depth blocksX*blocksY*64 u16, hiz blocksX*blocksY + 8 u16; either may be
 * NULL.  rz_upload_raw replaces the buffers.  Blocking. */
int rz_readback_raw(rz_context* ctx, uint16_t* depth, uint16_t* hiz);
int rz_upload_raw(rz_context* ctx, const uint16_t* depth, const uint16_t* hiz);

/* Waits for the context stream and returns the first device-side error
 * raised since the last blocking call (RZ_OK if none). */
int rz_sync(rz_context* ctx);

/* The HIP stream (hipStream_t) the context enqueues on. */
int rz_get_stream(rz_context* ctx, void** stream);

/* Context facts: width, height, blocksX, blocksY, tile width and height in
 * blocks used by the frame kernel. */
int rz_get_info(rz_context* ctx, uint32_t info[6]);

/* The two host-built tables as uploaded to the device: the 4096-entry u64
 * coverage table (Rasterizer.cpp:547-604) and the 4096-entry RCPPS table of
 * the host CPU (mantissa index = top 12 bits, value = RCPPS(1.m)). */
int rz_get_tables(rz_context* ctx, uint64_t* coverage4096, uint32_t* rcp4096);

/* Prebaked MVP state as setModelViewProjection stores it
 * (m_modelViewProjection, m_modelViewProjectionRaw). */
int rz_get_mvp(rz_context* ctx, float prebaked[16], float raw[16]);

/* Occluder::bake(const std::vector<__m128>& vertices, __m128 refMin, __m128 refMax)
 * (Occluder.h:9, Occluder.cpp:7-181), host side.  vertices: n x 4 floats, 4
 * per quad, n a multiple of 32 (8 quads).  packets: n words out.  bounds and
 * center as m_boundsMin, m_boundsMax, m_center. */
int rz_bake(const float* vertices, uint32_t n, const float refMin[4], const float refMax[4], uint32_t* packets,
            float boundsMin[4], float boundsMax[4], float center[4]);
/* Occluder::bake on the device, for many occluders at once (occluders built
 * at run time): occluder i's vertices are d_vertices[4 * vertexOffsets[i] ..
 * 4 * vertexOffsets[i + 1]) (device, xyzw floats; a positive multiple of 32
 * vertices, i.e. whole 8-quad batches; offsets multiples of 32), its
 * reference box refMinMax[8i .. 8i + 8) (host: min xyzw, max xyzw).  Writes
 * the packets, one word per vertex at the same offsets, to d_packets, and
 * boundsMin, boundsMax and centre (12 floats per occluder) to d_bounds.
 * Bit-identical to rz_bake; RZ_ERR_UNSUPPORTED_HOST if this CPU's RSQRTPS
 * cannot be tabulated.  vertexOffsets: n + 1 host words.  Blocks. */
int rz_bake_device(rz_context* ctx, uint32_t n, const uint32_t* vertexOffsets, const float* d_vertices,
                   const float* refMinMax, uint32_t* d_packets, float* d_bounds);
/* SurfaceAreaHeuristic::generateBatches(aabbs, targetSize, splitGranularity)
 * (SurfaceAreaHeuristic.h, SurfaceAreaHeuristic.cpp:96-104) on the device.
 * d_boxes: n quad boxes as std::vector<Aabb> lays them out (min xyzw, max
 * xyzw; w ignored; 16-byte aligned), finite, min <= max and
 * |coordinate| <= 2^40.  Writes the
 * batches' quad indices, batch after batch, to d_indices (n words, device)
 * and the batch boundaries to batchOffsets (host, *nBatches + 1 words; room
 * for n + 1 suffices), in the reference's depth-first order, bit-identical
 * to its splits: stable sorts by box centre (-0 equal to +0), DPPS surface
 * areas, the first strict minimum of the cost over (axis, split index).
 * Needs splitGranularity >= 1 and targetSize > 2 * splitGranularity (else a
 * child could have no split position); a set of at most 2 * splitGranularity
 * quads, where the reference's split is undefined, is one batch in input
 * order.  Boxes outside the stated range: RZ_ERR_INVALID_ARGUMENT.  Blocks. */
int rz_sah_batches(rz_context* ctx, uint32_t n, const float* d_boxes, uint32_t targetSize, uint32_t splitGranularity,
                   uint32_t* d_indices, uint32_t* batchOffsets, uint32_t* nBatches);
/* Main.cpp:86-128 on the device, for occluder sets that live in HBM: pad to
 * whole 8-quad batches with copies of vertex 0 (:90-94), one Aabb per quad
 * (:96-105), SurfaceAreaHeuristic::generateBatches (:107, as
 * rz_sah_batches), the scene box (:109-113) and Occluder::bake of each batch
 * against it (:115-128, as rz_bake_device).  The input is the quads' own
 * vertex list, so the scene box is taken over the quads' vertices; Main.cpp
 * takes it over the whole indexed vertex buffer, which can hold vertices no
 * quad references (a caller with such a mesh gets a different box there).
 * d_vertices: nQuads * 4 vertices, xyzw floats, 16-byte aligned, device;
 * x, y and z finite with |coordinate| <= 2^40 (else RZ_ERR_INVALID_ARGUMENT,
 * as for rz_sah_batches' boxes).
 * Writes occluder i's packets, one word per vertex, to
 * d_packets[4 * quadOffsets[i] .. 4 * quadOffsets[i + 1]) (room for
 * 4 * nQuads rounded up to 32 words), its boundsMin, boundsMax and centre to
 * d_bounds[12 i ..] (room for 12 * (nQuads + 7) / 8 floats), the batch
 * boundaries in padded quads to quadOffsets (host; room for
 * (nQuads + 7) / 8 + 1 words), the count to *nOccluders and the scene box
 * (min xyzw, max xyzw) to refMinMax.  Bit-identical to the host preparation
 * (rz_bake per batch of the reference's SAH batches); where a coordinate's
 * extreme is a zero of both signs, the scene box takes -0 as its minimum and
 * +0 as its maximum.  splitGranularity must be a multiple of 8 and
 * targetSize > 2 * splitGranularity.  Blocks. */
int rz_prepare_scene_device(rz_context* ctx, uint32_t nQuads, const float* d_vertices, uint32_t targetSize,
                            uint32_t splitGranularity, uint32_t* d_packets, float* d_bounds, uint32_t* quadOffsets,
                            uint32_t* nOccluders, float refMinMax[8]);
/* Main.cpp:86-128 on the device from the indexed triangle mesh, as Main.cpp
 * loads it: QuadDecomposition::decompose (as rz_quad_decompose_device), then
 * the padding, quad boxes, SAH batches and bake of rz_prepare_scene_device,
 * with the scene box taken over the whole vertex buffer (:109-113), which
 * can hold vertices no quad references.  d_indices: 3 * nTriangles words;
 * d_vertices: nVertices x 4 floats (xyzw, 16-byte aligned), every vertex's
 * x, y and z finite with |coordinate| <= 2^40 (else
 * RZ_ERR_INVALID_ARGUMENT); both device.  Outputs as rz_prepare_scene_device
 * for at most nTriangles quads: room for 4 * nTriangles rounded up to 32
 * words of packets, 12 * (nTriangles + 7) / 8 floats of bounds and
 * (nTriangles + 7) / 8 + 1 host words of quadOffsets.  Bit-identical to
 * rz_quad_decompose followed by the host preparation with the vertex
 * buffer's box.  Blocks. */
int rz_prepare_mesh_device(rz_context* ctx, uint32_t nTriangles, const uint32_t* d_indices, const float* d_vertices,
                           uint32_t nVertices, uint32_t targetSize, uint32_t splitGranularity, uint32_t* d_packets,
                           float* d_bounds, uint32_t* quadOffsets, uint32_t* nOccluders, float refMinMax[8]);

/* QuadDecomposition::decompose(indices, vertices) (QuadDecomposition.h:8,
 * QuadDecomposition.cpp:311-445), host side, no device needed: pairs the
 * triangles (indices: nIndices words, three per triangle; a trailing partial
 * triangle is ignored) into quads along shared edges whose two triangles keep
 * the far vertex within 0.5 units of each other's plane (RSQRTPS-normalised
 * normals, as the reference), as a maximum matching grown greedily and then
 * by augmenting paths; a triangle left single becomes the degenerate quad
 * (i0, i2, i1, i0).  Writes 4 indices per quad to out (room for
 * 4 * (nIndices / 3) words) and the number of words to *nOut; identical to
 * the reference's output, quad for quad.  vertices: nVertices x 4 floats
 * (xyzw); an index >= nVertices is RZ_ERR_INVALID_ARGUMENT. */
int rz_quad_decompose(const uint32_t* indices, size_t nIndices, const float* vertices, size_t nVertices, uint32_t* out,
                      size_t* nOut);
/* QuadDecomposition::decompose (QuadDecomposition.cpp:311-445) for a mesh in
 * HBM: the same quads as rz_quad_decompose, quad for quad.  The candidate
 * graph (half-edges grouped by a device radix sort, the 0.5-unit plane tests
 * with the host CPU's RSQRTPS, the adjacency lists in the reference's link
 * order) and the quads are built on the device; the maximum matching, a
 * sequential augmenting-path search, runs on the host over the adjacency
 * lists copied back.  d_indices: 3 * nTriangles words (device); d_vertices:
 * nVertices x 4 floats (xyzw, device, 16-byte aligned); an index >=
 * nVertices is RZ_ERR_INVALID_ARGUMENT.  Writes 4 indices per quad to d_out
 * (device, room for 4 * nTriangles words) and the number of words to *nOut.
 * Blocks. */
int rz_quad_decompose_device(rz_context* ctx, uint32_t nTriangles, const uint32_t* d_indices, const float* d_vertices,
                             uint32_t nVertices, uint32_t* d_out, uint32_t* nOut);

/* Timing helpers on the context stream (HIP events): rz_timer_start records
 * a start event, rz_timer_stop a stop event and returns (after waiting for
 * it) the elapsed milliseconds between them. */
int rz_timer_start(rz_context* ctx);
int rz_timer_stop(rz_context* ctx, float* ms);

/* Per-kernel timing with HIP events recorded around each launch on the
 * context stream.  rz_profile_enable(ctx, 1) starts (and clears) the
 * accumulation, 0 stops it.  rz_profile_read blocks and returns the summed
 * milliseconds of the prep, setup and frame kernels of all draws and of the
 * query-batch kernel, and the numbers of draws and query batches timed. */
int rz_profile_enable(rz_context* ctx, int enable);
int rz_profile_read(rz_context* ctx, double ms[4], uint32_t counts[2]);

/* Per-tile counters of the frame kernel for the most recent draw
 * (observability).  rz_frame_stats_enable(ctx, 1) switches draws to the
 * counting variant of the kernel.  rz_frame_stats_read blocks and copies
 * tilesX*tilesY records of 22 u64 into out (when out is non-NULL and *count
 * is large enough) and sets *count to the number of records: total ticks,
 * ticks waiting on other tiles, waits, ticks in local query2D work, query2D
 * contributions, ticks drawing (scan + draw), quads drawn, items touching the
 * tile, scan ticks, scan passes, tile load ticks, item-loop ticks, draw
 * ticks, (quad, block) pairs visited, of which rejected, of which written,
 * draw ticks staging per-row folds, draw ticks in the per-block loops, chunks
 * drawn, fold additions staged, query partials evaluated ahead of their turn
 * and of those contributed early (ticks of the 100 MHz s_memrealtime clock). */
int rz_frame_stats_enable(rz_context* ctx, int enable);
int rz_frame_stats_read(rz_context* ctx, uint64_t* out, uint32_t* count);

/* ---------------------------------------------------------------------------
 * Launch configuration and autotuning (no counterpart in the reference, whose
 * raster is one CPU thread).  Every configuration computes identical bits;
 * only time differs.
 * ------------------------------------------------------------------------- */
typedef struct rz_frame_config {
  uint32_t tileW; /* tile edge in 8x8 blocks: 8 (64-pixel bins) or 16 (128-pixel bins) */
  uint32_t waves; /* wave64s per tile work-group: 2, 4 or 8 */
  float ms;       /* rz_tune_frame: median time of one rz_draw_frame; -1 if the grid cannot be co-resident */
} rz_frame_config;

int rz_get_config(rz_context* ctx, rz_frame_config* config);
/* RZ_ERR_TOO_LARGE (configuration unchanged) when the tile grid cannot be
 * co-resident on the device. */
int rz_set_config(rz_context* ctx, uint32_t tileW, uint32_t waves);
/* Autotuner: for every configuration (tileW {8, 16} x waves {2, 4, 8}) that
 * fits, clears and
 * draws the frame of occ (rz_draw_frame with the context's matrix), requires
 * identical depth, hiZ and statuses from all of them (else
 * RZ_ERR_NOT_IDENTICAL, configuration unchanged), and times `reps` draws
 * (HIP events, median).  Applies the fastest and, with a non-NULL cachePath,
 * records it there for this device and buffer size.  candidates (optional, 6
 * entries) receives every configuration tried.  Leaves the drawn frame in
 * the buffer.  Blocks. */
int rz_tune_frame(rz_context* ctx, uint32_t n, const rz_occluder* const* occ, uint32_t reps, const char* cachePath,
                  rz_frame_config candidates[6], uint32_t* nCandidates);
/* Applies the configuration cachePath records for this device and buffer
 * size; *found = 0 when it has none.  rz_create does this itself when the
 * environment variable RZ_TUNE_CACHE names a cache file. */
int rz_load_config(rz_context* ctx, const char* cachePath, int* found);

/* Device memory for callers without an allocator of their own (e.g. the
 * inputs of rz_query_batch_device): rz_device_alloc allocates `bytes` on the
 * context's device and copies `init` into it when non-NULL; rz_device_read
 * copies back after the context stream has drained; rz_device_free releases
 * (after a device-wide synchronization). */
int rz_device_alloc(rz_context* ctx, size_t bytes, const void* init, void** devptr);
int rz_device_read(rz_context* ctx, const void* devptr, size_t bytes, void* dst);
int rz_device_free(void* devptr);

/* ---------------------------------------------------------------------------
 * Batched multi-view launches.  No counterpart in the reference class: they
 * run the reference's per-view sequence for many views in one step.  For each
 * view: clear (Rasterizer.cpp:107-121); then, per occluder in the view's
 * order, the depth-independent part of queryVisibility (Rasterizer.cpp:123-
 * 264: frustum cull, near-plane test, empty screen rectangle) and
 * rasterize<needsClipping> (:606-1295) of the survivors, with no occlusion
 * test between draws; then queryVisibility (:123-281) of the view's boxes
 * against the finished buffer.  Each view's buffer, statuses and query bits
 * equal that sequence on one Rasterizer.
 * ------------------------------------------------------------------------- */
typedef struct rz_views rz_views;

/* nViews buffers of widths[i] x heights[i] (non-zero multiples of 8) on the
 * current device, each zero-initialised with an identity matrix. */
int rz_views_create(uint32_t nViews, const uint32_t* widths, const uint32_t* heights, rz_views** out);
int rz_views_destroy(rz_views* views);
/* setModelViewProjection of one view (takes effect at the next run). */
int rz_views_set_mvp(rz_views* views, uint32_t view, const float matrix[16]);
/* The view's occluders in draw order (handles from rz_occluder_create on a
 * context of the same device; they must outlive their use here). */
int rz_views_set_occluders(rz_views* views, uint32_t view, uint32_t n, const rz_occluder* const* occluders);
/* The view's query boxes (n x 8 floats: min xyzw, max xyzw), copied. */
int rz_views_set_queries(rz_views* views, uint32_t view, uint32_t n, const float* bounds);
/* Enqueues one step of every view.  useGraph != 0 replays a hipGraph of the
 * step, captured at the first run after an occluder list or query count
 * changed; 0 launches the same kernels directly. */
int rz_views_run(rz_views* views, int useGraph);
/* Waits for the last step; *ms (optional) is its GPU time between events. */
int rz_views_sync(rz_views* views, float* ms);
/* Results of the last step for one view (block until it has finished):
 * raw depth + hiZ as rz_readback_raw; query bits as rz_query_batch; per
 * occluder status 0 skipped, 1 rasterize<false>, 2 rasterize<true>. */
int rz_views_read_raw(rz_views* views, uint32_t view, uint16_t* depth, uint16_t* hiz);
int rz_views_read_queries(rz_views* views, uint32_t view, uint8_t* visible, uint8_t* needsClipping);
int rz_views_read_status(rz_views* views, uint32_t view, uint8_t* status);
/* Switches every view to the reference's query-gated draw loop (Main.cpp:
 * 192-206): per occluder in the view's order, queryVisibility (Rasterizer.cpp:
 * 123-281) against the view's buffer as drawn so far, then
 * rasterize<needsClipping> (:606-1295) if visible; statuses as rz_draw_frame.
 * Each view equals rz_draw_frame on a Rasterizer whose buffer holds what the
 * view's buffer held (clear leaves depth stale, and query2D reads stale depth
 * in partly covered cleared blocks, Rasterizer.cpp:283-349).  Tiles of a view wait for
 * each other's query answers, so a step needs every view's tiles to fit
 * co-resident (a view too large for 8x8-block tiles moves the step to
 * 16x16-block tiles; beyond those run returns RZ_ERR_TOO_LARGE), is launched
 * cooperatively and is not captured into a graph.  It combines with a
 * front-to-back order (rz_views_set_eye) and with a model that orders whole
 * occluders (rz_views_set_model with byOccluder != 0): the query-gated loop
 * then runs over the occluders in that order, one query per occluder as in
 * the reference; batch orders, explicit or learned, would query per batch and
 * are refused (RZ_ERR_INVALID_ARGUMENT at the next run).  0 restores
 * classify-then-draw. */
int rz_views_set_query_gated(rz_views* views, int gated);
/* Bound of each cross-tile wait of query-gated steps, in 100 MHz
 * s_memrealtime ticks (0 restores the default, 2 s); as rz_set_wait_timeout:
 * a wait that runs out ends the launch's other waits, every kernel exits
 * normally and rz_views_sync returns RZ_ERR_TIMEOUT. */
int rz_views_set_wait_timeout(rz_views* views, uint64_t ticks);

/* Observability of the last step (no counterpart in the reference): the
 * view's per-occluder prologue state (128 B records: the folded matrix T[4][4]
 * and c0, c1 of Rasterizer.cpp:616-655, then the draw status) and its quads'
 * setup records (80 B each, in the view's occluder order: edge normals and
 * offsets, depth plane, first block, block ranges, mode | packed maxZ << 16,
 * slope lookups; Rasterizer.cpp:657-1081).  Either pointer may be NULL. */
int rz_views_read_items(rz_views* views, uint32_t view, void* state, void* setup);

/* Phase timing with HIP events on the views stream (steps then run without
 * graph capture while it is on): rz_views_profile_read blocks and returns the
 * milliseconds of the last step's clear + prep + setup (+ learned-order work),
 * of the raster launch alone, and of the query launch. */
int rz_views_profile_enable(rz_views* views, int enable);
int rz_views_profile_read(rz_views* views, float ms[3]);

/* info[0..7]: views, items, quads, (view, tile) units, queries, raster
 * work-groups, graph captured, plan out of date. */
int rz_views_get_info(rz_views* views, uint32_t info[8]);

/* Launch configuration of the multi-view raster kernel (no counterpart in the
 * reference): the tile (bin) edge in 8x8 blocks, 8 (64-pixel bins) or 16
 * (128-pixel bins), wave64s per (view, tile) work-group, 2, 4 or 8, and
 * persistent work-groups per CU (0: as many as fit).  Labelled steps (labels
 * on, or a model that reads them) draw with 8x8-block tiles whatever the
 * setting.  A query-gated step whose views' 8x8-block tiles cannot all be
 * co-resident (2048^2, 3840x2160) draws with 16x16-block tiles by itself. */
typedef struct rz_views_config {
  uint32_t tileW;
  uint32_t waves;
  uint32_t groupsPerCU;
  float ms; /* rz_views_tune: median step time; -1 if it does not fit */
} rz_views_config;
int rz_views_get_config(rz_views* views, rz_views_config* config);
int rz_views_set_config(rz_views* views, uint32_t tileW, uint32_t waves, uint32_t groupsPerCU);
/* Autotuner over tileW {8, 16} x waves {2, 4, 8} x groupsPerCU {1, all that
 * fit}: runs the current step (matrices, occluders, queries, order) under
 * each, requires identical buffers, query bits and labels from all (else
 * RZ_ERR_NOT_IDENTICAL, configuration unchanged; configurations whose grid
 * cannot hold a query-gated view's tiles, and 16x16 tiles in labelled steps,
 * are skipped), times `reps` steps each, applies the fastest and records it
 * in cachePath (if not NULL) for this device and set of view sizes.
 * candidates (optional) receives up to 12 entries.  Blocks. */
int rz_views_tune(rz_views* views, uint32_t reps, const char* cachePath, rz_views_config candidates[12],
                  uint32_t* nCandidates);
/* As rz_load_config; rz_views_create does it itself with RZ_TUNE_CACHE. */
int rz_views_load_config(rz_views* views, const char* cachePath, int* found);

/* Draw-order learning (no counterpart in the reference).  With labels on,
 * every step also computes, per 8-quad batch (the k-th batch of a view is the
 * k-th in its occluder order): 16 features before the draw and, after it,
 * the label = number of written blocks in which the batch wins some pixel.
 * The draw tracks each pixel's winner: the first write of a cleared block
 * sets it for every pixel (none where the written depth is 0), a later write
 * only where it raises the depth; so a pixel's winner is the first batch in
 * submission order whose write produced its final depth.  Counted by the
 * draw itself, at no extra raster pass. */
int rz_views_enable_labels(rz_views* views, int enable);
/* Draws the view's batches in the given order (a permutation of its batch
 * numbers 0..n-1), each with its occluder's classification; n = 0 restores
 * occluder order. */
int rz_views_set_batch_order(rz_views* views, uint32_t view, uint32_t n, const uint32_t* order);
/* Labels (u32 per batch) and features (16 floats per batch) of the last step. */
int rz_views_read_labels(rz_views* views, uint32_t view, uint32_t* labels, float* features);

/* The batch-scoring MLP (no counterpart in the reference): 16 features ->
 * 64 -> 64 -> 1 with GELU, BF16 operands on MFMA with FP32 accumulation.
 * params: 5313 floats (W1[16][64], b1[64], W2[64][64], b2[64], W3[64], b3);
 * norm: 16 feature means then 16 inverse deviations. */
typedef struct rz_mlp rz_mlp;
int rz_mlp_create(uint32_t seed, rz_mlp** out);
int rz_mlp_destroy(rz_mlp* mlp);
int rz_mlp_set_params(rz_mlp* mlp, const float* params, const float* norm);
int rz_mlp_get_params(rz_mlp* mlp, float* params, float* norm);
/* Whether the model reads feature 10, the batch's label from the previous
 * step (default 1).  A model trained with that feature held at 0 is set to 0:
 * multi-view steps ordered by it then feed 0 there and draw without tracking
 * labels, which saves the labelled raster variant's cost. */
int rz_mlp_set_uses_labels(rz_mlp* mlp, int usesLabels);
/* Scores of n samples (16 floats each); host arrays. */
int rz_mlp_forward(rz_mlp* mlp, uint32_t n, const float* x, float* scores);
/* Mean pairwise ranking loss softplus(s_j - s_i) over pairs (i, j) (sample i
 * should score higher) and its gradient in the parameters, without update. */
int rz_mlp_gradients(rz_mlp* mlp, uint32_t n, const float* x, uint32_t nPairs, const uint32_t* pairs, float* grad,
                     float* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss. */
int rz_mlp_train_step(rz_mlp* mlp, uint32_t n, const float* x, uint32_t nPairs, const uint32_t* pairs, float lr,
                      float* loss);
/* Learned draw order for a multi-view batch: every step scores each view's
 * batches with the model (from this step's features and the previous step's
 * labels).  byOccluder == 0 draws the batches by descending score; != 0 draws
 * whole occluders by their best batch score (no batch items, so it scales to
 * large occluder sets).  Only batches of drawn (not frustum-culled)
 * occluders are scored; the others score -infinity, so they come last (their
 * occluders score -FLT_MAX, as occluders without batches).  Ties keep
 * occluder order.  NULL restores occluder order.  The model must outlive its
 * use. */
int rz_views_set_model(rz_views* views, rz_mlp* model, int byOccluder);
/* The order of the last learned or front-to-back step: view-local batch
 * numbers, or occluder numbers when ordering by occluder or front to back. */
int rz_views_read_order(rz_views* views, uint32_t view, uint32_t* order);
/* Front-to-back draw order computed on the device every step, replacing
 * Main.cpp:184-190's std::sort: the view's occluders by ascending squared
 * distance from eye to their centre ((boundsMin + boundsMax) * 0.5,
 * Occluder.cpp:178), float32 summed as DPPS sums it ((dx^2 + dy^2) + dz^2);
 * ties keep submission order (std::sort leaves them unspecified).  Works in
 * both launch modes; statuses and query results stay in submission order.
 * NULL restores submission order.  Cannot be combined with a learned model or
 * an explicit batch order (RZ_ERR_INVALID_ARGUMENT at the next run). */
int rz_views_set_eye(rz_views* views, uint32_t view, const float eye[3]);
/* Block-loop counters of each step (observability): rz_views_read_stats
 * returns, for the last step over all views, the (quad, block) pairs
 * visited, those rejected by the hiZ or edge tests, and blocks written. */
int rz_views_enable_stats(rz_views* views, int enable);
int rz_views_read_stats(rz_views* views, uint64_t out[3]);
/* The per-tile counters of the last step's raster launch (stats enabled),
 * summed over all (view, tile) units, in the 22-word layout of
 * rz_frame_stats_read; word 0 is the sum of the persistent work-groups'
 * lifetimes, in 100 MHz ticks like the other tick words. */
int rz_views_read_tile_stats(rz_views* views, uint64_t out[22]);

/* On-disk formats (no counterpart in the reference, which re-bakes its
 * scene at every start): versioned little-endian files with a magic, the
 * payload size and an FNV-1a 64 checksum of the payload.  Host only; no
 * device needed.  RZ_ERR_IO when a file cannot be written or read, or is not
 * this format, version or checksum. */

/* A baked occluder as the reference holds it (Occluder.h:7-21,
 * Occluder.cpp:153-176): packetCount packets of 8 words (m_vertexData) and
 * its reference box and bounds. */
typedef struct rz_occluder_data {
  const uint32_t* packets;
  uint32_t packetCount; /* multiple of 4 (whole 8-quad batches) */
  float refMin[4], refMax[4], boundsMin[4], boundsMax[4];
} rz_occluder_data;
int rz_write_occluders(const char* path, uint32_t n, const rz_occluder_data* occluders);
/* With occluders == NULL: sets *n and *totalPackets from the file.  Else *n
 * and *totalPackets give the capacity of occluders and packetWords
 * (totalPackets * 8 words); the descriptors are filled with their packets
 * pointing into packetWords, and *n, *totalPackets are set to the file's. */
int rz_read_occluders(const char* path, uint32_t* n, uint32_t* totalPackets, rz_occluder_data* occluders,
                      uint32_t* packetWords);
/* A frame dump: raw depth and hiZ as rz_readback_raw returns them
 * ((width/8)*(height/8)*64 and (width/8)*(height/8) + 8 u16) and the query
 * bits of nQueries queries (NULL when nQueries is 0).  rz_read_frame sets
 * the sizes and fills the arrays that are not NULL (call it with NULL arrays
 * first to learn the sizes). */
int rz_write_frame(const char* path, uint32_t width, uint32_t height, const uint16_t* depth, const uint16_t* hiz,
                   uint32_t nQueries, const uint8_t* visible, const uint8_t* needsClipping);
int rz_read_frame(const char* path, uint32_t* width, uint32_t* height, uint32_t* nQueries, uint16_t* depth,
                  uint16_t* hiz, uint8_t* visible, uint8_t* needsClipping);

#ifdef __cplusplus
}
#endif
#endif
