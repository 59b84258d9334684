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
  // Multi-view launches: per-item view table (width, height and matrix come
  // from views[item.view]) and the work-queue counter that prep resets.
  const ViewDesc* views;
  uint32_t* queue;
  // Single-view draws: the descriptors as the host wrote them, in pinned
  // host memory mapped for the device.  Prep reads each there and copies it
  // to itemsOut, the device table the later kernels read (one transfer
  // fewer than a copy before the launch).  Null: `items` holds them.
  const ItemDesc* hostItems;
  ItemDesc* itemsOut;
  // Optional: for every group of 64 quads, the item that holds its first
  // quad (the setup kernel's item search starts there).
  uint32_t* groupItem;
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
  const ViewDesc* views;  // multi-view launches: bX, bY from views[item.view]
  const uint32_t* groupItem;  // optional: per group of 64 quads, the item of its first quad
  // Optional (learned-order steps): per quad of a drawn item, its record's
  // modeZ | kFeatValid when the quad is drawn (non-empty rectangle), else 0:
  // the features' per-quad input in 4 bytes instead of the 80-byte record.
  uint32_t* featZ;
};
constexpr uint32_t kFeatValid = 0x100u;  // bits 8-15 of modeZ are unused

// Persistent multi-view raster launch: work units are (view, tile) pairs taken
// from an atomic queue in view order; units wait only on tiles of their own
// view, and only for query-gated items.
struct ViewsParams {
  const ItemDesc* items;
  const ItemState* state;
  const ViewDesc* views;
  uint32_t nViews, nUnits;
  const QuadRec* recs;
  const uint64_t* groupBox;
  const uint64_t* cov;
  uint32_t* queue;
  // Draw batch items instead of occluder items: view v's items are
  // [views[v].batchBase, + nBatches) of items/state.
  uint32_t batchItems;
  uint32_t waves;  // wave64s per work-group: 4 or 8
  uint32_t tileB;  // tile edge in 8x8 blocks of every view of the launch: 8, or 16 (no labels)
  // Optional block-loop counters of the launch (visited, rejected by hiZ or
  // the edge tests, written), added to stats[0..2].
  unsigned long long* stats;
  // Optional labelled draw: per-batch counts of the written blocks in which
  // each batch wins a pixel (zeroed by the caller), with each view's
  // per-pixel winners in views[v].winner.
  uint32_t* labels;
  // Query-gated items (kItemQuery): per-item done words (query partials and
  // visible tiles), the error word and the bounded wait, as the frame
  // kernel's.  Null done: no gated items.
  uint32_t* done;
  uint32_t* err;
  uint64_t timeoutTicks;
  // Optional per-tile item lists (launch_views_bin): tile t of view v walks
  // lists[views[v].listBase + t * (items in view) ...], listLen[unit] long.
  uint32_t* lists;
  uint32_t* listLen;
  // Optional (with lists): the units that have items, as view << 20 | tile,
  // in view order and within a view by descending item count (log2
  // buckets, tile order within a bucket), and their count; the queue hands
  // out only these.
  uint32_t* unitList;
  uint32_t* nActive;
  uint32_t* unitCounts;  // scratch: per view
};

// queryVisibility over the queries of every view of a multi-view launch.
struct ViewsQueryParams {
  const ViewDesc* views;
  uint32_t nViews, nQueries;
  uint32_t maxPerView;  // most queries of any view (grid x extent, 256 per work-group)
  const uint32_t* rcp;
  // Optional: the step's error word, copied by the first thread to errOut
  // (host memory mapped for the device), so that the host reads it after
  // the step without a copy of its own.  The kernel is then launched even
  // without queries.
  const uint32_t* err;
  uint32_t* errOut;
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
  kStatBlocksVisited = 13,   // (quad, block) pairs visited by the block loop
  kStatBlocksRejected = 14,  // of which rejected by hiZ or the edge tests
  kStatBlocksWritten = 15,   // of which written
  kStatPhaseATicks = 16,     // block-major draw: staging the per-row folds (with its barrier)
  kStatPhaseBTicks = 17,     // block-major draw: the per-block candidate loops (with its barrier)
  kStatChunks = 18,          // staged chunks drawn
  kStatFoldSteps = 19,       // sequential fold additions made by phase A (all waves)
  kStatEarlyTried = 20,      // query partials evaluated ahead of their item's turn
  kStatEarlyAnswered = 21,   // of which contributed early (provably not visible)
  kFrameStatWords = 22,
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
// cooperative == false launches the same grid as a plain dispatch (for tools
// that cannot trace cooperative dispatches); the grid is sized to be
// co-resident either way and every cross-tile wait is bounded in time.
// waves: 4 or 8 wave64s per tile (block ownership interleaves over them).
hipError_t launch_frame(const FrameParams& p, int tileW, hipStream_t stream, bool cooperative = true, int waves = 8);
hipError_t frame_max_groups(int tileW, int waves, int* maxGroups);
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
hipError_t launch_clear(uint16_t* hiz, uint32_t nHiz, uint16_t* tileMin, uint32_t nMin, hipStream_t stream);
// Multi-view: grid of `groups` persistent work-groups over p.nUnits units of
// 8x8-block tiles.
hipError_t launch_views(const ViewsParams& p, uint32_t groups, hipStream_t stream, bool cooperative = false);
hipError_t views_max_groups(int tileB, int waves, int* maxGroups);
// Builds p.lists / p.listLen from p.state (after prep and setup); tileBox:
// nItems u64 of scratch (the items of p.items / p.state, batch items
// included).
hipError_t launch_views_bin(const ViewsParams& p, uint32_t nItems, uint64_t* tileBox, hipStream_t stream);
// Builds p.unitList / p.nActive from p.listLen (after launch_views_bin);
// unitTmp: nUnits words of scratch.
hipError_t launch_views_compact(const ViewsParams& p, uint32_t* unitTmp, hipStream_t stream);
hipError_t launch_views_query(const ViewsQueryParams& p, hipStream_t stream);
// hiZ := 1 (whole 8-entry groups) and tile min := 1 for every view.
hipError_t launch_views_clear(const ViewDesc* src, ViewDesc* views, uint32_t nViews, uint32_t maxEntries,
                              hipStream_t stream);

// ---------------------------------------------------------------------------
// Draw-order learning (rz_learn.hip).  Batches are numbered over the launch:
// batch b holds quads 8b .. 8b+7 of the setup records.
// ---------------------------------------------------------------------------
struct LearnParams {
  const ItemDesc* items;
  const ItemState* state;
  uint32_t nItems;
  const ViewDesc* views;
  uint32_t nViews;
  const uint32_t* featZ;  // per quad, as SetupParams::featZ
  const uint64_t* bbox;
  uint32_t nQuads;
  const uint64_t* cov;
  uint32_t* labels;    // per batch
  float* features;     // per batch, kFeatures floats
  const uint32_t* groupItem;  // optional: per group of 64 quads, the item of its first quad
  uint32_t* drawn;     // optional: the batches whose occluder is drawn, in no particular order
  uint32_t* nDrawn;    // their count (zero it before the features launch)
};
constexpr int kFeatures = 16;

hipError_t launch_features(const LearnParams& p, uint32_t nBatches, hipStream_t stream);
// Batch items in the order `order` (per view segment, batch numbers): item k
// draws batch order[k] with its occluder's status and the batch's footprint.
hipError_t launch_batch_items(const LearnParams& p, const uint32_t* order, uint32_t nBatches, ItemDesc* outItems,
                              ItemState* outState, hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t stream);
// Occluder-granular learned order: each item's score is the largest score of
// its batches; then items are copied in the order `order` (per view segment).
hipError_t launch_item_scores(const LearnParams& p, const float* batchScores, float* itemScores,
                              hipStream_t stream);
// Radix keys that order n scored elements (batches, or items when items is
// true) by view, then by descending score: (view << 32) | ~(the score's
// order-preserving bits).  A stable ascending sort over the low
// 32 + (bits of nViews - 1) bits then gives each view's elements by
// descending score, ties in element order.
hipError_t launch_sort_keys(const ViewDesc* views, uint32_t nViews, const float* scores, uint32_t n, bool items,
                            uint64_t* keys, hipStream_t stream);
// Front-to-back keys (Main.cpp:184-190) of n items: (view << 32) | the bits
// of the squared distance from eyes[view] (x, y, z; w != 0 when the view is
// sorted) to the item's centre (boundsMin + boundsMax) * 0.5
// (Occluder.cpp:178), summed as DPPS sums it; views without an eye get key
// (view << 32), so a stable sort keeps their submission order.
hipError_t launch_f2b_keys(const ItemDesc* items, uint32_t n, const float* eyes, uint64_t* keys, hipStream_t stream);
hipError_t launch_item_reorder(const LearnParams& p, const uint32_t* order, ItemDesc* outItems, ItemState* outState,
                               hipStream_t stream);

// ---------------------------------------------------------------------------
// Batch-scoring MLP 16 -> 64 -> 64 -> 1, GELU, BF16 operands on MFMA with FP32
// accumulation (rz_mlp.hip).  Parameters, fp32, in one array:
//   W1 [16][64] at 0, b1 [64] at 1024, W2 [64][64] at 1088, b2 [64] at 5184,
//   W3 [64] at 5248, b3 at 5312.  Inputs are normalised per feature by
//   (x - mean) * invStd (norm = 16 means then 16 inverse deviations).
// ---------------------------------------------------------------------------
constexpr int kMlpIn = 16, kMlpHidden = 64, kMlpParams = 5313;
constexpr int kMlpW1 = 0, kMlpB1 = 1024, kMlpW2 = 1088, kMlpB2 = 5184, kMlpW3 = 5248, kMlpB3 = 5312;

struct MlpParams {
  const float* w;
  const float* norm;
  const float* x;       // n x 16
  uint32_t n;
  const uint32_t* rows;   // optional (forward): score only samples rows[0 .. *nRows), nRows[0] <= n
  const uint32_t* nRows;
  float* scores;        // forward output
  const float* dscore;  // backward input: dLoss/dscore per sample
  float* grad;          // backward output, accumulated (zero it first)
};
hipError_t launch_mlp_forward(const MlpParams& p, hipStream_t stream);
// Work-groups of the scoring kernel resident on the current device at once
// (queried once per process; call outside stream capture before the first
// captured launch).
uint32_t mlp_forward_max_groups();

// Occluder::bake on the device (rz_bake.hip): occluder o's vertices are
// vertices[offsets[o] .. offsets[o + 1]) (4 floats each, a multiple of 32),
// its reference box ref[8o .. 8o + 8) (min xyzw, max xyzw); packets get one
// word per vertex at the same offsets, bounds 12 floats per occluder
// (boundsMin, boundsMax, centre).  Scratch per quad: normals, cluster, order.
struct BakeParams {
  const float* vertices;
  const uint32_t* offsets;
  const float* ref;
  uint32_t* packets;
  float* bounds;
  float4* normals;
  uint32_t* cluster;
  uint32_t* order;
  const uint32_t* rsqrt;  // 8192-entry host RSQRTPS table
};
hipError_t launch_bake(const BakeParams& p, uint32_t nOccluders, hipStream_t stream);
// SAH batching of n quad boxes on the stream (rz_sah.hip); an RZ_* status.
int sah_batches_run(hipStream_t stream, uint32_t n, const float* d_boxes, uint32_t target, uint32_t gran,
                    uint32_t* d_indices, uint32_t* batchOffsets, uint32_t* nBatches);
// QuadDecomposition of nTri device triangles on the stream (rz_quad.hip): the
// candidate graph and the quads on the device, the matching on the host;
// writes the quads' indices to d_out and their word count to *nOut.  An
// RZ_* status.
int quad_decompose_run(hipStream_t stream, uint32_t nTri, const uint32_t* d_idx, const float* d_vertices,
                       uint32_t nVertices, const uint32_t* hostRsqrt, uint32_t* d_out, uint32_t* nOut);
// Scene preparation around it (rz_sah.hip): boxes of nPad padded quads and
// the scene box as ordered-float keys (min xyzw, max xyzw; the caller sets
// them to ~0 and 0 first) and sceneKeys[8] set non-zero when a vertex has a
// coordinate (x, y, z) that is not finite or exceeds 2^40 (zero it first);
// the padded quads' vertices gathered in `order`.
hipError_t launch_scene_boxes(const float* vertices, uint32_t nQuads, uint32_t nPad, float* boxes,
                              uint32_t* sceneKeys, hipStream_t stream);
hipError_t launch_scene_gather(const float* vertices, uint32_t nQuads, uint32_t nPad, const uint32_t* order,
                               float* out, hipStream_t stream);
// The scene box's ordered keys (and the bad-vertex word) over a whole vertex
// buffer of n xyzw vertices, folded into sceneKeys.
hipError_t launch_vertex_keys(const float* vertices, uint32_t n, uint32_t* sceneKeys, hipStream_t stream);
// Quad q's four vertices, vertices[quadIdx[4q + j]], for nWords / 4 quads.
hipError_t launch_quad_gather(const float* vertices, const uint32_t* quadIdx, uint32_t nWords, float* out,
                              hipStream_t stream);
hipError_t launch_mlp_backward(const MlpParams& p, hipStream_t stream);
// Pairwise ranking loss over pairs (i, j) whose sample i should rank above j:
// mean softplus(s_j - s_i); adds its gradient to dscore and its value to *loss.
hipError_t launch_mlp_pair_loss(const float* scores, const uint32_t* pairs, uint32_t nPairs, float* dscore,
                                float* loss, hipStream_t stream);
hipError_t launch_mlp_adam(float* w, const float* grad, float* m, float* v, uint32_t n, float lr, float beta1,
                           float beta2, float eps, float corr1, float corr2, hipStream_t stream);
hipError_t launch_fill_f32(float* p, uint32_t n, float v, hipStream_t stream);

}  // namespace rzd
