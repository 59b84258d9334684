// Batched multi-view launches (C ABI in rasterizer_amd.h, rz_views_*).
//
// A step renders every view of the batch: clear, the depth-independent part
// of queryVisibility per occluder (frustum cull, near-plane test, empty
// rectangle; Rasterizer.cpp:123-264), rasterize<needsClipping> of the
// survivors in each view's draw order, then the view's query boxes.  One
// prep launch and one setup launch cover every (view, occluder) item and
// every quad; one persistent raster launch takes (view, tile) units from an
// atomic queue; one query launch covers every view's boxes.  The sequence is
// captured once into a hipGraph and replayed per step; matrices live in a
// pinned host table that the graph copies to the device at the start of each
// replay, so changing them does not need a new capture.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <new>
#include <limits>
#include <vector>

#include "rasterizer_amd.h"
#include "rz_host.h"
#include "rz_internal.h"
#include "rz_kernels.h"

namespace {
constexpr uint64_t kGatedTimeoutTicks = 200000000ull;  // 2 s of the 100 MHz s_memrealtime clock
}

using namespace rzd;

namespace {

struct ViewHost {
  uint32_t width = 0, height = 0, bX = 0, bY = 0, pyrX = 0, pyrY = 0;  // pyr: the 8x8-block tile grid of tileMin
  uint16_t* d_depth = nullptr;
  uint16_t* d_hiz = nullptr;
  uint16_t* d_tileMin = nullptr;
  std::vector<const rz_occluder*> occ;
  uint32_t nQueries = 0;
  float* d_qbounds = nullptr;
  uint8_t* d_qvis = nullptr;
  uint8_t* d_qclip = nullptr;
  uint32_t* d_winner = nullptr;  // per pixel, when labels are on
  std::vector<uint32_t> order;   // explicit batch order (view-local batch numbers), empty = occluder order
  Mvp mvp{};
  float eye[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // front-to-back order from this eye when eye[3] != 0
};

template <typename T>
hipError_t dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

}  // namespace

struct rz_views {
  hipStream_t stream = nullptr;
  uint64_t timeoutTicks = kGatedTimeoutTicks;  // bound of each cross-tile wait (rz_views_set_wait_timeout)
  std::vector<ViewHost> views;
  uint64_t* d_cov = nullptr;
  uint32_t* d_rcp = nullptr;
  // plan, rebuilt when an occluder list or a query count changes
  bool dirty = true;
  uint32_t nItems = 0, nQuads = 0, nUnits = 0, nQueries = 0, nGroups = 0, maxClear = 0;
  ItemDesc* d_items = nullptr;
  ItemState* d_state = nullptr;
  uint32_t* d_done = nullptr;
  uint32_t* d_visible = nullptr;
  QuadRec* d_recs = nullptr;
  uint32_t* d_featZ = nullptr;  // per quad: modeZ | kFeatValid of drawn quads (learned-order steps)
  uint64_t* d_bbox = nullptr;
  uint64_t* d_groupBox = nullptr;
  uint32_t* d_groupItem = nullptr;  // per group of 64 quads: the item of its first quad
  ViewDesc* d_views = nullptr;
  ViewDesc* h_views = nullptr;  // pinned, coherent; the clear kernel copies it to d_views every step
  const ViewDesc* d_hViews = nullptr;  // h_views' device mapping
  float* h_eyes = nullptr;      // pinned, 4 floats per view; copied to d_eyes by front-to-back steps
  float* d_eyes = nullptr;
  bool f2b = false;             // some view is drawn front to back (order computed on the device)
  uint32_t* d_queue = nullptr;
  // draw-order learning
  bool labels = false;    // compute features before and labels after the draw
  bool batchOrder = false;  // draw batch items in d_order
  uint32_t nBatches = 0, maxBlocks = 0;
  uint32_t* d_labels = nullptr;
  float* d_features = nullptr;
  uint32_t* d_order = nullptr;
  ItemDesc* d_bitems = nullptr;
  ItemState* d_bstate = nullptr;
  // learned order: scores of the batches by the model, sorted per view
  rz_mlp* model = nullptr;
  bool byOccluder = false;  // learned order of whole occluders (by their best batch) instead of batches
  float* d_itemScores = nullptr;
  uint32_t* d_itemIota = nullptr;
  uint32_t* d_itemOrder = nullptr;
  ItemDesc* d_ritems = nullptr;
  ItemState* d_rstate = nullptr;
  float* d_scores = nullptr;
  uint32_t* d_iota = nullptr;
  uint64_t* d_sortKeys = nullptr;     // radix keys (view, descending score) of batches or items
  uint64_t* d_sortKeysOut = nullptr;
  int sortEndBit = 32;
  uint32_t* d_drawn = nullptr;     // batches of drawn occluders (scored by the model), count in d_drawn[nBatches]
  void* d_sortTemp = nullptr;
  size_t sortTempBytes = 0;
  int waves = 4;  // wave64s per raster work-group
  int tileB = 8;  // raster tile edge in 8x8 blocks (8 or 16; labelled steps use 8)
  int planTileB = 8;  // the tile edge of the current plan
  int groupsPerCU = 0;  // persistent raster work-groups per CU (0: as many as fit)
  uint32_t* d_lists = nullptr;    // per-tile item lists (null: tiles walk their view's item range)
  uint32_t* d_listLen = nullptr;  // per (view, tile) unit
  uint64_t* d_tileBox = nullptr;  // per drawn item (occluder or batch): tile box, for binning
  uint32_t* d_unitList = nullptr;  // units with items, as view << 20 | tile (with lists)
  uint32_t* d_unitTmp = nullptr;
  uint32_t* d_unitCounts = nullptr;  // per view
  uint32_t* d_nActive = nullptr;
  bool gated = false;   // query-gated draw loop per view (Main.cpp:192-206) instead of classify-then-draw
  bool cooperative = true;  // gated launches: cooperative dispatch (RZ_FRAME_LAUNCH=plain: plain)
  uint32_t* d_err = nullptr;  // device error word of gated waits
  uint32_t* h_err = nullptr;  // pinned, coherent: d_err as the last step's query kernel left it
  uint32_t* d_hErr = nullptr;  // h_err's device mapping
  unsigned long long* d_stats = nullptr;  // block-loop counters of the last step, when on
  bool stats = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Cooperative (gated) steps: the launches before the raster launch and the
  // queries after it, captured as two graphs around the directly enqueued
  // raster launch, whose parameters are kept in rp.
  hipGraph_t graphPre = nullptr, graphPost = nullptr;
  hipGraphExec_t execPre = nullptr, execPost = nullptr;
  ViewsParams rp{};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // Phase timing of direct (not captured) steps, when on: events before the
  // clear, before the raster launch, after it and after the queries.
  bool profile = false;
  hipEvent_t ph[4] = {nullptr, nullptr, nullptr, nullptr};
  bool phRecorded = false;
  bool ran = false;
};

static void free_plan(rz_views* v) {
  if (v->exec) (void)hipGraphExecDestroy(v->exec);
  if (v->graph) (void)hipGraphDestroy(v->graph);
  v->exec = nullptr;
  v->graph = nullptr;
  for (hipGraphExec_t* e : {&v->execPre, &v->execPost}) {
    if (*e) (void)hipGraphExecDestroy(*e);
    *e = nullptr;
  }
  for (hipGraph_t* g : {&v->graphPre, &v->graphPost}) {
    if (*g) (void)hipGraphDestroy(*g);
    *g = nullptr;
  }
  hipFree(v->d_items);
  hipFree(v->d_state);
  hipFree(v->d_done);
  hipFree(v->d_visible);
  hipFree(v->d_recs);
  hipFree(v->d_featZ);
  v->d_featZ = nullptr;
  hipFree(v->d_bbox);
  hipFree(v->d_groupBox);
  hipFree(v->d_groupItem);
  v->d_groupItem = nullptr;
  hipFree(v->d_views);
  hipFree(v->d_labels);
  hipFree(v->d_features);
  hipFree(v->d_order);
  hipFree(v->d_bitems);
  hipFree(v->d_bstate);
  hipFree(v->d_lists);
  hipFree(v->d_listLen);
  hipFree(v->d_tileBox);
  hipFree(v->d_unitList);
  hipFree(v->d_unitTmp);
  hipFree(v->d_unitCounts);
  hipFree(v->d_nActive);
  v->d_lists = v->d_listLen = nullptr;
  v->d_tileBox = nullptr;
  v->d_unitList = v->d_unitTmp = v->d_unitCounts = v->d_nActive = nullptr;
  hipFree(v->d_itemScores);
  hipFree(v->d_itemIota);
  hipFree(v->d_itemOrder);
  hipFree(v->d_ritems);
  hipFree(v->d_rstate);
  v->d_itemScores = nullptr;
  v->d_itemIota = v->d_itemOrder = nullptr;
  v->d_ritems = nullptr;
  v->d_rstate = nullptr;
  hipFree(v->d_scores);
  hipFree(v->d_sortKeys);
  hipFree(v->d_sortKeysOut);
  hipFree(v->d_iota);
  hipFree(v->d_sortTemp);
  hipFree(v->d_drawn);
  v->d_scores = nullptr;
  v->d_sortKeys = v->d_sortKeysOut = nullptr;
  v->d_iota = v->d_drawn = nullptr;
  v->d_sortTemp = nullptr;
  v->sortTempBytes = 0;
  v->d_labels = nullptr;
  v->d_features = nullptr;
  v->d_order = nullptr;
  v->d_bitems = nullptr;
  v->d_bstate = nullptr;
  if (v->h_views) (void)hipHostFree(v->h_views);
  if (v->h_eyes) (void)hipHostFree(v->h_eyes);
  hipFree(v->d_eyes);
  v->h_eyes = nullptr;
  v->d_eyes = nullptr;
  v->d_items = nullptr;
  v->d_state = nullptr;
  v->d_done = v->d_visible = nullptr;
  v->d_recs = nullptr;
  v->d_bbox = v->d_groupBox = nullptr;
  v->d_views = nullptr;
  v->h_views = nullptr;
  v->d_hViews = nullptr;
}

static void free_views(rz_views* v) {
  if (!v) return;
  if (v->stream) (void)hipStreamSynchronize(v->stream);
  free_plan(v);
  hipFree(v->d_err);
  v->d_err = nullptr;
  if (v->h_err) (void)hipHostFree(v->h_err);
  v->h_err = nullptr;
  v->d_hErr = nullptr;
  for (ViewHost& h : v->views) {
    hipFree(h.d_depth);
    hipFree(h.d_hiz);
    hipFree(h.d_tileMin);
    hipFree(h.d_qbounds);
    hipFree(h.d_qvis);
    hipFree(h.d_qclip);
    hipFree(h.d_winner);
  }
  hipFree(v->d_cov);
  hipFree(v->d_rcp);
  hipFree(v->d_queue);
  hipFree(v->d_stats);
  if (v->t0) (void)hipEventDestroy(v->t0);
  if (v->t1) (void)hipEventDestroy(v->t1);
  for (hipEvent_t e : v->ph)
    if (e) (void)hipEventDestroy(e);
  if (v->stream) (void)hipStreamDestroy(v->stream);
  delete v;
}

extern "C" int rz_views_create(uint32_t nViews, const uint32_t* widths, const uint32_t* heights, rz_views** out) {
  if (!out || !widths || !heights || nViews == 0) return RZ_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  for (uint32_t i = 0; i < nViews; ++i) {
    const uint32_t w = widths[i], h = heights[i];
    if (w == 0 || h == 0 || (w % 8) || (h % 8) || w > 65528 || h > 65528) return RZ_ERR_INVALID_ARGUMENT;
    if (uint64_t(w / 8) * (h / 8) >= (1u << 24)) return RZ_ERR_TOO_LARGE;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RZ_ERR_DEVICE;
  const rzi::HostTables& tabs = rzi::host_tables();
  if (tabs.status != RZ_OK) return tabs.status;
  rz_views* v = new (std::nothrow) rz_views();
  if (!v) return RZ_ERR_OUT_OF_MEMORY;
  auto fail = [&](int code) {
    free_views(v);
    return code;
  };
  if (hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking) != hipSuccess) return fail(RZ_ERR_DEVICE);
  if (hipEventCreate(&v->t0) != hipSuccess || hipEventCreate(&v->t1) != hipSuccess) return fail(RZ_ERR_DEVICE);
  for (hipEvent_t& e : v->ph)
    if (hipEventCreate(&e) != hipSuccess) return fail(RZ_ERR_DEVICE);
  if (dev_alloc(&v->d_cov, 4096) != hipSuccess || dev_alloc(&v->d_rcp, 4096) != hipSuccess ||
      dev_alloc(&v->d_queue, 1) != hipSuccess || dev_alloc(&v->d_err, 1) != hipSuccess)
    return fail(RZ_ERR_OUT_OF_MEMORY);
  if (hipMemset(v->d_err, 0, sizeof(uint32_t)) != hipSuccess) return fail(RZ_ERR_DEVICE);
  if (hipHostMalloc(reinterpret_cast<void**>(&v->h_err), sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent) !=
      hipSuccess)
    return fail(RZ_ERR_OUT_OF_MEMORY);
  *v->h_err = 0;
  if (hipHostGetDevicePointer(reinterpret_cast<void**>(&v->d_hErr), v->h_err, 0) != hipSuccess)
    return fail(RZ_ERR_DEVICE);
  if (const char* mode = std::getenv("RZ_FRAME_LAUNCH")) v->cooperative = std::strcmp(mode, "plain") != 0;
  if (hipMemcpy(v->d_cov, tabs.cov, sizeof tabs.cov, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(v->d_rcp, tabs.rcp, sizeof tabs.rcp, hipMemcpyHostToDevice) != hipSuccess)
    return fail(RZ_ERR_DEVICE);
  v->views.resize(nViews);
  const float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  for (uint32_t i = 0; i < nViews; ++i) {
    ViewHost& h = v->views[i];
    h.width = widths[i];
    h.height = heights[i];
    h.bX = h.width / 8;
    h.bY = h.height / 8;
    h.pyrX = (h.bX + 7) / 8;
    h.pyrY = (h.bY + 7) / 8;
    const size_t nb = size_t(h.bX) * h.bY;
    if (dev_alloc(&h.d_depth, nb * 64) != hipSuccess || dev_alloc(&h.d_hiz, nb + 8) != hipSuccess ||
        dev_alloc(&h.d_tileMin, size_t(h.pyrX) * h.pyrY) != hipSuccess)
      return fail(RZ_ERR_OUT_OF_MEMORY);
    // Zero-initialised like the reference's std::vector buffers.
    if (hipMemset(h.d_depth, 0, nb * 64 * 2) != hipSuccess || hipMemset(h.d_hiz, 0, (nb + 8) * 2) != hipSuccess ||
        hipMemset(h.d_tileMin, 0, size_t(h.pyrX) * h.pyrY * 2) != hipSuccess)
      return fail(RZ_ERR_DEVICE);
    rzh::prebake_mvp(h.width, h.height, ident, h.mvp.prebaked, h.mvp.raw);
  }
  if (const char* cache = rzi::tune_cache_env()) {
    int found = 0;
    (void)rz_views_load_config(v, cache, &found);  // a stale or unusable entry leaves the default
  }
  *out = v;
  return RZ_OK;
}

extern "C" int rz_views_destroy(rz_views* v) {
  if (!v) return RZ_ERR_INVALID_ARGUMENT;
  free_views(v);
  return RZ_OK;
}

// Waits for the last step so that the pinned view table can be rewritten.
static int wait_last_step(rz_views* v) {
  if (v->ran) RZ_HIP(hipEventSynchronize(v->t1));
  return RZ_OK;
}

extern "C" int rz_views_set_mvp(rz_views* v, uint32_t view, const float matrix[16]) {
  if (!v || !matrix || view >= v->views.size()) return RZ_ERR_INVALID_ARGUMENT;
  ViewHost& h = v->views[view];
  rzh::prebake_mvp(h.width, h.height, matrix, h.mvp.prebaked, h.mvp.raw);
  return RZ_OK;
}

extern "C" int rz_views_set_occluders(rz_views* v, uint32_t view, uint32_t n, const rz_occluder* const* occluders) {
  if (!v || view >= v->views.size() || (n && !occluders)) return RZ_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n; ++i)
    if (!occluders[i]) return RZ_ERR_INVALID_ARGUMENT;
  v->views[view].occ.assign(occluders, occluders + n);
  v->dirty = true;
  return RZ_OK;
}

extern "C" int rz_views_set_queries(rz_views* v, uint32_t view, uint32_t n, const float* bounds) {
  if (!v || view >= v->views.size() || (n && !bounds)) return RZ_ERR_INVALID_ARGUMENT;
  int s = wait_last_step(v);
  if (s != RZ_OK) return s;
  ViewHost& h = v->views[view];
  if (n != h.nQueries) {
    hipFree(h.d_qbounds);
    hipFree(h.d_qvis);
    hipFree(h.d_qclip);
    h.d_qbounds = nullptr;
    h.d_qvis = h.d_qclip = nullptr;
    h.nQueries = 0;
    if (dev_alloc(&h.d_qbounds, size_t(n) * 8) != hipSuccess || dev_alloc(&h.d_qvis, n) != hipSuccess ||
        dev_alloc(&h.d_qclip, n) != hipSuccess)
      return RZ_ERR_OUT_OF_MEMORY;
    h.nQueries = n;
    v->dirty = true;
  }
  if (n) RZ_HIP(hipMemcpy(h.d_qbounds, bounds, size_t(n) * 32, hipMemcpyHostToDevice));
  return RZ_OK;
}

// Builds item and view tables for the current occluder lists and captures the
// step into a graph.
static int build_plan(rz_views* v) {
  free_plan(v);
  const uint32_t nv = uint32_t(v->views.size());
  // Labelled draws need 8x8-block tiles; a query-gated view whose 8x8-block
  // tiles cannot all be co-resident moves the launch to 16x16-block tiles.
  const bool labelled = v->labels || (v->model && v->model->usesLabels);
  int tileB = labelled ? 8 : v->tileB;
  if (v->gated && tileB == 8 && !labelled) {
    int maxGroups = 0;
    RZ_HIP(views_max_groups(8, v->waves, &maxGroups));
    for (const ViewHost& h : v->views)
      if (((h.bX + 7) / 8) * ((h.bY + 7) / 8) > uint32_t(maxGroups)) tileB = 16;
  }
  v->planTileB = tileB;
  std::vector<ItemDesc> items;
  uint64_t quads = 0;
  uint32_t units = 0, queries = 0, maxClear = 0;
  v->maxBlocks = 0;
  if (hipHostMalloc(reinterpret_cast<void**>(&v->h_views), nv * sizeof(ViewDesc),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    return RZ_ERR_OUT_OF_MEMORY;
  {
    void* mapped = nullptr;
    RZ_HIP(hipHostGetDevicePointer(&mapped, v->h_views, 0));
    v->d_hViews = static_cast<const ViewDesc*>(mapped);
  }
  for (uint32_t i = 0; i < nv; ++i) {
    ViewHost& h = v->views[i];
    ViewDesc& d = v->h_views[i];
    std::memset(&d, 0, sizeof d);
    d.width = h.width;
    d.height = h.height;
    d.bX = h.bX;
    d.bY = h.bY;
    d.tilesX = (h.bX + uint32_t(tileB) - 1) / uint32_t(tileB);
    d.tilesY = (h.bY + uint32_t(tileB) - 1) / uint32_t(tileB);
    d.tileB = uint32_t(tileB);
    d.pyrX = h.pyrX;
    d.pyrY = h.pyrY;
    d.itemBase = uint32_t(items.size());
    d.nItems = uint32_t(h.occ.size());
    d.unitBase = units;
    d.nUnits = d.tilesX * d.tilesY;
    d.queryBase = queries;
    d.nQueries = h.nQueries;
    d.batchBase = uint32_t(quads / 8);
    d.depth = h.d_depth;
    d.hiz = h.d_hiz;
    d.tileMin = h.d_tileMin;
    d.queryBounds = h.d_qbounds;
    d.queryVisible = h.d_qvis;
    d.queryClip = h.d_qclip;
    d.mvp = h.mvp;
    units += d.nUnits;
    queries += d.nQueries;
    maxClear = std::max(maxClear, std::max((h.bX * h.bY + 8) / 8, h.pyrX * h.pyrY));  // 16-byte hiZ stores
    for (const rz_occluder* o : h.occ) {
      ItemDesc it{};
      it.packets = o->d_packets;
      it.quadBase = uint32_t(quads);
      it.quadCount = (o->packetCount / 4) * 8;
      it.kind = v->gated ? kItemQuery : kItemClassify;
      it.view = i;
      std::memcpy(it.refMin, o->refMin, 16);
      std::memcpy(it.refMax, o->refMax, 16);
      std::memcpy(it.boundsMin, o->boundsMin, 16);
      std::memcpy(it.boundsMax, o->boundsMax, 16);
      quads += it.quadCount;
      items.push_back(it);
    }
    d.nBatches = uint32_t(quads / 8) - d.batchBase;
    d.winner = h.d_winner;
    v->maxBlocks = std::max(v->maxBlocks, h.bX * h.bY);
    if (!h.order.empty() && h.order.size() != d.nBatches) return RZ_ERR_INVALID_ARGUMENT;
  }
  if (quads >= (1ull << 31)) return RZ_ERR_TOO_LARGE;
  v->nBatches = uint32_t(quads / 8);
  v->nItems = uint32_t(items.size());
  v->nQuads = uint32_t(quads);
  v->nUnits = units;
  v->nQueries = queries;
  v->maxClear = maxClear;
  int maxGroups = 0;
  RZ_HIP(views_max_groups(tileB, v->waves, &maxGroups));
  if (v->groupsPerCU > 0) {
    int dev = 0, cus = 0;
    RZ_HIP(hipGetDevice(&dev));
    RZ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    maxGroups = std::min(maxGroups, v->groupsPerCU * cus);
  }
  v->nGroups = std::min<uint32_t>(units, uint32_t(std::max(maxGroups, 1)));
  if (v->gated) {
    // Every wait of a gated view ends only if all of the view's tiles can run
    // at once (nGroups <= units, and units count every view's tiles); batch
    // orders would query per batch, which the reference's loop does not (a
    // model that orders whole occluders keeps the loop's per-occluder query).
    bool order = v->model != nullptr && !v->byOccluder;
    for (const ViewHost& h : v->views) order |= !h.order.empty();
    if (order) return RZ_ERR_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < nv; ++i)
      if (v->h_views[i].nUnits > v->nGroups) return RZ_ERR_TOO_LARGE;
  }
  const size_t ni = std::max<size_t>(items.size(), 1);
  if (dev_alloc(&v->d_items, ni) != hipSuccess || dev_alloc(&v->d_state, ni) != hipSuccess ||
      dev_alloc(&v->d_done, ni) != hipSuccess || dev_alloc(&v->d_visible, ni) != hipSuccess ||
      dev_alloc(&v->d_recs, std::max<size_t>(quads, 1)) != hipSuccess ||
      dev_alloc(&v->d_featZ, std::max<size_t>(quads, 1)) != hipSuccess ||
      dev_alloc(&v->d_bbox, std::max<size_t>(quads, 1)) != hipSuccess ||
      dev_alloc(&v->d_groupBox, std::max<size_t>((quads + 63) / 64, 1)) != hipSuccess ||
      dev_alloc(&v->d_groupItem, std::max<size_t>((quads + 63) / 64, 1)) != hipSuccess ||
      dev_alloc(&v->d_views, nv) != hipSuccess)
    return RZ_ERR_OUT_OF_MEMORY;
  if (!items.empty()) {
    RZ_HIP(hipMemcpy(v->d_items, items.data(), items.size() * sizeof(ItemDesc), hipMemcpyHostToDevice));
    // Item of each 64-quad group's first quad (the last item starting at or
    // before it), for the setup and features lookups.
    std::vector<uint32_t> groupItem((quads + 63) / 64);
    uint32_t it = 0;
    for (size_t k = 0; k < groupItem.size(); ++k) {
      while (it + 1 < items.size() && items[it + 1].quadBase <= k * 64) ++it;
      groupItem[k] = it;
    }
    if (!groupItem.empty())
      RZ_HIP(hipMemcpy(v->d_groupItem, groupItem.data(), groupItem.size() * 4, hipMemcpyHostToDevice));
  }
  const size_t nbt = std::max<size_t>(v->nBatches, 1);
  if (dev_alloc(&v->d_labels, nbt) != hipSuccess || dev_alloc(&v->d_features, nbt * kFeatures) != hipSuccess)
    return RZ_ERR_OUT_OF_MEMORY;
  RZ_HIP(hipMemset(v->d_labels, 0, nbt * sizeof(uint32_t)));
  RZ_HIP(hipMemset(v->d_features, 0, nbt * kFeatures * sizeof(float)));
  // An explicit batch order in any view switches the step to batch items;
  // views without one keep their batches in occluder order.
  v->batchOrder = v->model != nullptr && !v->byOccluder;
  for (const ViewHost& h : v->views) v->batchOrder |= !h.order.empty();
  // Front-to-back views: the device sorts whole occluders, which neither a
  // model's order nor an explicit batch order can be combined with.
  v->f2b = false;
  for (const ViewHost& h : v->views) v->f2b |= h.eye[3] != 0.0f;
  if (v->f2b && (v->model || v->batchOrder)) return RZ_ERR_INVALID_ARGUMENT;
  if (v->f2b) {
    if (hipHostMalloc(reinterpret_cast<void**>(&v->h_eyes), size_t(nv) * 4 * sizeof(float)) != hipSuccess ||
        dev_alloc(&v->d_eyes, size_t(nv) * 4) != hipSuccess)
      return RZ_ERR_OUT_OF_MEMORY;
  }
  {
    // Per-tile item lists: each unit gets room for all of its view's items
    // (or batches); skipped when that would not fit 32-bit offsets or 4 GB,
    // in which case tiles walk their view's whole range.
    uint64_t total = 0;
    for (uint32_t i = 0; i < nv; ++i) {
      ViewDesc& d = v->h_views[i];
      d.listBase = uint32_t(total);
      total += uint64_t(d.nUnits) * (v->batchOrder ? d.nBatches : d.nItems);
    }
    if (total > 0 && total < (1ull << 30)) {
      const size_t nDraw = std::max<size_t>(v->batchOrder ? v->nBatches : v->nItems, 1);
      if (dev_alloc(&v->d_lists, size_t(total)) != hipSuccess ||
          dev_alloc(&v->d_listLen, std::max<size_t>(v->nUnits, 1)) != hipSuccess ||
          dev_alloc(&v->d_tileBox, nDraw) != hipSuccess)
        return RZ_ERR_OUT_OF_MEMORY;
      // Compacted unit list (view << 20 | tile): fewer than 4096 views of
      // fewer than 2^20 tiles each.
      bool fits = nv < 4096;
      for (uint32_t i = 0; i < nv; ++i) fits &= v->h_views[i].nUnits < (1u << 20);
      if (fits && (dev_alloc(&v->d_unitList, std::max<size_t>(v->nUnits, 1)) != hipSuccess ||
                   dev_alloc(&v->d_unitTmp, std::max<size_t>(v->nUnits, 1)) != hipSuccess ||
                   dev_alloc(&v->d_unitCounts, nv) != hipSuccess || dev_alloc(&v->d_nActive, 1) != hipSuccess))
        return RZ_ERR_OUT_OF_MEMORY;
    }
  }
  // Occluder-granular reorders (learned by occluder, front to back): items
  // and their states are copied in the sorted order after setup.
  const bool itemReorder = (v->model && v->byOccluder) || v->f2b;
  int viewBits = 0;
  while ((1u << viewBits) < nv) ++viewBits;
  v->sortEndBit = 32 + viewBits;  // view index above the 32 key bits
  if (itemReorder) {
    const size_t ni = std::max<size_t>(v->nItems, 1);
    std::vector<uint32_t> iota(v->nItems);
    for (uint32_t i = 0; i < v->nItems; ++i) iota[i] = i;
    if ((v->model && dev_alloc(&v->d_itemScores, ni) != hipSuccess) || dev_alloc(&v->d_itemIota, ni) != hipSuccess ||
        dev_alloc(&v->d_itemOrder, ni) != hipSuccess || dev_alloc(&v->d_ritems, ni) != hipSuccess ||
        dev_alloc(&v->d_rstate, ni) != hipSuccess)
      return RZ_ERR_OUT_OF_MEMORY;
    if (v->nItems) RZ_HIP(hipMemcpy(v->d_itemIota, iota.data(), iota.size() * 4, hipMemcpyHostToDevice));
  }
  if (v->f2b) {
    size_t bytes = 0;
    if (dev_alloc(&v->d_sortKeys, std::max<size_t>(v->nItems, 1)) != hipSuccess ||
        dev_alloc(&v->d_sortKeysOut, std::max<size_t>(v->nItems, 1)) != hipSuccess)
      return RZ_ERR_OUT_OF_MEMORY;
    RZ_HIP(rocprim::radix_sort_pairs(nullptr, bytes, v->d_sortKeys, v->d_sortKeysOut, v->d_itemIota,
                                              v->d_itemOrder, uint32_t(v->nItems), 0u, unsigned(v->sortEndBit), v->stream));
    if (hipMalloc(&v->d_sortTemp, std::max<size_t>(bytes, 16)) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
    v->sortTempBytes = bytes;
  }
  if (v->model) {
    std::vector<uint32_t> iota(v->nBatches);
    for (uint32_t b = 0; b < v->nBatches; ++b) iota[b] = b;
    const size_t nk = std::max<size_t>(std::max(v->nBatches, v->byOccluder ? v->nItems : 0u), 1);
    if (dev_alloc(&v->d_scores, nbt) != hipSuccess || dev_alloc(&v->d_iota, nbt) != hipSuccess ||
        dev_alloc(&v->d_sortKeys, nk) != hipSuccess || dev_alloc(&v->d_sortKeysOut, nk) != hipSuccess ||
        dev_alloc(&v->d_drawn, nbt + 1) != hipSuccess)
      return RZ_ERR_OUT_OF_MEMORY;
    if (v->nBatches) RZ_HIP(hipMemcpy(v->d_iota, iota.data(), iota.size() * 4, hipMemcpyHostToDevice));
    // One device-wide stable radix sort on (view, descending score) keys.
    size_t bytes = 0, itemBytes = 0;
    RZ_HIP(rocprim::radix_sort_pairs(nullptr, bytes, v->d_sortKeys, v->d_sortKeysOut, v->d_iota, v->d_iota,
                                              uint32_t(v->nBatches), 0u, unsigned(v->sortEndBit), v->stream));
    if (v->byOccluder) {
      RZ_HIP(rocprim::radix_sort_pairs(nullptr, itemBytes, v->d_sortKeys, v->d_sortKeysOut, v->d_itemIota,
                                                v->d_itemOrder, uint32_t(v->nItems), 0u, unsigned(v->sortEndBit), v->stream));
      bytes = std::max(bytes, itemBytes);
    }
    if (hipMalloc(&v->d_sortTemp, std::max<size_t>(bytes, 16)) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
    v->sortTempBytes = bytes;
  }
  if (v->batchOrder) {
    std::vector<uint32_t> order(v->nBatches);
    for (uint32_t i = 0; i < nv; ++i) {
      const ViewDesc& d = v->h_views[i];
      const ViewHost& h = v->views[i];
      for (uint32_t k = 0; k < d.nBatches; ++k) {
        const uint32_t b = h.order.empty() ? k : h.order[k];
        if (b >= d.nBatches) return RZ_ERR_INVALID_ARGUMENT;
        order[d.batchBase + k] = d.batchBase + b;
      }
    }
    if (dev_alloc(&v->d_order, nbt) != hipSuccess || dev_alloc(&v->d_bitems, nbt) != hipSuccess ||
        dev_alloc(&v->d_bstate, nbt) != hipSuccess)
      return RZ_ERR_OUT_OF_MEMORY;
    if (v->nBatches) RZ_HIP(hipMemcpy(v->d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice));
  }
  v->dirty = false;
  return RZ_OK;
}

static LearnParams learn_params(rz_views* v) {
  LearnParams lp{};
  lp.items = v->d_items;
  lp.state = v->d_state;
  lp.nItems = v->nItems;
  lp.views = v->d_views;
  lp.nViews = uint32_t(v->views.size());
  lp.featZ = v->d_featZ;
  lp.bbox = v->d_bbox;
  lp.nQuads = v->nQuads;
  lp.cov = v->d_cov;
  lp.labels = v->d_labels;
  lp.features = v->d_features;
  lp.groupItem = v->d_groupItem;
  if (v->model) {
    lp.drawn = v->d_drawn;
    lp.nDrawn = v->d_drawn + v->nBatches;
  }
  return lp;
}

// Parts of a step: the launches before the raster launch, the raster
// launch, and the queries after it.
constexpr int kStepPre = 1, kStepRaster = 2, kStepPost = 4, kStepAll = 7;

// The launches of one step (or of the given parts of it), on v->stream.
static int enqueue_step(rz_views* v, bool capturing, int parts = kStepAll) {
  const uint32_t nv = uint32_t(v->views.size());
  const bool prof = v->profile && !capturing;
  if (parts == kStepRaster) {
    if (v->nItems) RZ_HIP(launch_views(v->rp, v->nGroups, v->stream, v->gated && v->cooperative));
    return RZ_OK;
  }
  if (parts & kStepPre) {
    if (prof) RZ_HIP(hipEventRecord(v->ph[0], v->stream));
    RZ_HIP(launch_views_clear(v->d_hViews, v->d_views, nv, v->maxClear, v->stream));
    PrepParams pp{};
    pp.items = v->d_items;
    pp.state = v->d_state;
    pp.nItems = v->nItems;
    pp.tileW = pp.tileH = 8;
    pp.rcp = v->d_rcp;
    pp.done = v->d_done;
    pp.visible = v->d_visible;
    pp.views = v->d_views;
    pp.queue = v->d_queue;
    if (v->nItems == 0) {
      // Nothing to draw: the prep launch that resets the queue is skipped, and
      // so is the raster launch that reads it.
    } else {
      RZ_HIP(launch_prep(pp, v->stream));
      SetupParams sp{};
      sp.items = v->d_items;
      sp.state = v->d_state;
      sp.nItems = v->nItems;
      sp.nQuads = v->nQuads;
      sp.rcp = v->d_rcp;
      sp.recs = v->d_recs;
      sp.bbox = v->d_bbox;
      sp.groupBox = v->d_groupBox;
      sp.groupItem = v->d_groupItem;
      sp.views = v->d_views;
      sp.featZ = (v->labels || v->model) ? v->d_featZ : nullptr;
      RZ_HIP(launch_setup(sp, v->stream));
      const LearnParams lp = learn_params(v);
      if (lp.nDrawn) RZ_HIP(launch_fill_u32(lp.nDrawn, 1, 0u, v->stream));
      if (v->labels || v->model) RZ_HIP(launch_features(lp, v->nBatches, v->stream));
      if (v->model && v->nBatches) {
        // Score the batches of drawn occluders (the others score -inf), then
        // order each view's batches by descending score; the radix sort is
        // stable, so ties keep occluder order.
        RZ_HIP(launch_fill_f32(v->d_scores, v->nBatches, -std::numeric_limits<float>::infinity(), v->stream));
        MlpParams mp{};
        mp.w = v->model->d_w;
        mp.norm = v->model->d_norm;
        mp.x = v->d_features;
        mp.n = v->nBatches;
        mp.rows = lp.drawn;
        mp.nRows = lp.nDrawn;
        mp.scores = v->d_scores;
        RZ_HIP(launch_mlp_forward(mp, v->stream));
        size_t bytes = v->sortTempBytes;
        if (v->byOccluder) {
          // Whole occluders by their best batch score, stable within a view.
          RZ_HIP(launch_item_scores(lp, v->d_scores, v->d_itemScores, v->stream));
          RZ_HIP(launch_sort_keys(v->d_views, nv, v->d_itemScores, v->nItems, true, v->d_sortKeys, v->stream));
          RZ_HIP(rocprim::radix_sort_pairs(v->d_sortTemp, bytes, v->d_sortKeys, v->d_sortKeysOut,
                                                    v->d_itemIota, v->d_itemOrder, uint32_t(v->nItems), 0u, unsigned(v->sortEndBit),
                                                    v->stream));
          RZ_HIP(launch_item_reorder(lp, v->d_itemOrder, v->d_ritems, v->d_rstate, v->stream));
        } else {
          RZ_HIP(launch_sort_keys(v->d_views, nv, v->d_scores, v->nBatches, false, v->d_sortKeys, v->stream));
          RZ_HIP(rocprim::radix_sort_pairs(v->d_sortTemp, bytes, v->d_sortKeys, v->d_sortKeysOut, v->d_iota,
                                                    v->d_order, uint32_t(v->nBatches), 0u, unsigned(v->sortEndBit), v->stream));
        }
      }
      if (v->f2b) {
        // Front to back (Main.cpp:184-190), stable: ties keep submission order.
        size_t bytes = v->sortTempBytes;
        RZ_HIP(hipMemcpyAsync(v->d_eyes, v->h_eyes, nv * 4 * sizeof(float), hipMemcpyHostToDevice, v->stream));
        RZ_HIP(launch_f2b_keys(v->d_items, v->nItems, v->d_eyes, v->d_sortKeys, v->stream));
        RZ_HIP(rocprim::radix_sort_pairs(v->d_sortTemp, bytes, v->d_sortKeys, v->d_sortKeysOut, v->d_itemIota,
                                                  v->d_itemOrder, uint32_t(v->nItems), 0u, unsigned(v->sortEndBit), v->stream));
        RZ_HIP(launch_item_reorder(lp, v->d_itemOrder, v->d_ritems, v->d_rstate, v->stream));
      }
      if (v->batchOrder) RZ_HIP(launch_batch_items(lp, v->d_order, v->nBatches, v->d_bitems, v->d_bstate, v->stream));
      ViewsParams rp{};
      const bool reordered = ((v->model && v->byOccluder) || v->f2b) && v->nItems;
      rp.items = v->batchOrder ? v->d_bitems : (reordered ? v->d_ritems : v->d_items);
      rp.state = v->batchOrder ? v->d_bstate : (reordered ? v->d_rstate : v->d_state);
      rp.batchItems = v->batchOrder ? 1u : 0u;
      rp.views = v->d_views;
      rp.nViews = nv;
      rp.nUnits = v->nUnits;
      rp.recs = v->d_recs;
      rp.groupBox = v->d_groupBox;
      rp.cov = v->d_cov;
      rp.queue = v->d_queue;
      rp.waves = uint32_t(v->waves);
      rp.tileB = uint32_t(v->planTileB);
      rp.stats = v->stats ? v->d_stats : nullptr;
      if (v->stats)
        RZ_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(v->d_stats), 2 * (3 + kFrameStatWords), 0u, v->stream));
      // Labels are counted by the draw itself (after this step's features
      // have read the previous step's).
      // A model that does not read the previous labels leaves them at 0 (the
      // plan zeroes them) and the draw without label tracking.
      rp.labels = (v->labels || (v->model && v->model->usesLabels)) ? v->d_labels : nullptr;
      if (rp.labels) RZ_HIP(launch_fill_u32(v->d_labels, v->nBatches, 0u, v->stream));
      rp.lists = v->d_lists;
      rp.listLen = v->d_listLen;
      RZ_HIP(launch_views_bin(rp, v->batchOrder ? v->nBatches : v->nItems, v->d_tileBox, v->stream));
      rp.unitList = v->d_unitList;
      rp.nActive = v->d_nActive;
      rp.unitCounts = v->d_unitCounts;
      RZ_HIP(launch_views_compact(rp, v->d_unitTmp, v->stream));
      rp.done = v->gated ? v->d_done : nullptr;
      rp.err = v->d_err;
      rp.timeoutTicks = v->timeoutTicks;
      v->rp = rp;
      if (parts & kStepRaster) {
        if (prof) RZ_HIP(hipEventRecord(v->ph[1], v->stream));
        RZ_HIP(launch_views(rp, v->nGroups, v->stream, v->gated && v->cooperative));
        if (prof) RZ_HIP(hipEventRecord(v->ph[2], v->stream));
      }
    }
  }
  if (!(parts & kStepPost)) return RZ_OK;
  ViewsQueryParams qp{};
  qp.views = v->d_views;
  qp.nViews = nv;
  qp.nQueries = v->nQueries;
  for (const ViewHost& h : v->views) qp.maxPerView = std::max(qp.maxPerView, h.nQueries);
  qp.rcp = v->d_rcp;
  qp.err = v->d_err;
  qp.errOut = v->d_hErr;
  RZ_HIP(launch_views_query(qp, v->stream));
  if (prof) {
    if (v->nItems == 0) {
      RZ_HIP(hipEventRecord(v->ph[1], v->stream));
      RZ_HIP(hipEventRecord(v->ph[2], v->stream));
    }
    RZ_HIP(hipEventRecord(v->ph[3], v->stream));
  }
  v->phRecorded = prof;
  return RZ_OK;
}

extern "C" int rz_views_run(rz_views* v, int useGraph) {
  if (!v) return RZ_ERR_INVALID_ARGUMENT;
  int s = wait_last_step(v);
  if (s != RZ_OK) return s;
  if (v->dirty) {
    s = build_plan(v);
    if (s != RZ_OK) return s;
  }
  for (uint32_t i = 0; i < v->views.size(); ++i) {
    v->h_views[i].mvp = v->views[i].mvp;
    if (v->h_eyes) std::memcpy(v->h_eyes + 4 * i, v->views[i].eye, 4 * sizeof(float));
  }
  if (v->profile) useGraph = 0;
  // Gated steps launch the raster kernel cooperatively (its grid must be
  // co-resident), which is not captured: the launches before it and the
  // queries after it are replayed as two graphs around the direct launch.
  const bool split = useGraph && v->gated && v->cooperative;
  if (split) {
    if (!v->execPre) {
      for (int part : {kStepPre, kStepPost}) {
        RZ_HIP(hipStreamBeginCapture(v->stream, hipStreamCaptureModeThreadLocal));
        s = enqueue_step(v, true, part);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(v->stream, &g);
        if (s != RZ_OK || e != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          return s != RZ_OK ? s : RZ_ERR_DEVICE;
        }
        (part == kStepPre ? v->graphPre : v->graphPost) = g;
        RZ_HIP(hipGraphInstantiate(part == kStepPre ? &v->execPre : &v->execPost, g, nullptr, nullptr, 0));
      }
    }
    RZ_HIP(hipEventRecord(v->t0, v->stream));
    RZ_HIP(hipGraphLaunch(v->execPre, v->stream));
    s = enqueue_step(v, false, kStepRaster);
    if (s != RZ_OK) return s;
    RZ_HIP(hipGraphLaunch(v->execPost, v->stream));
    RZ_HIP(hipEventRecord(v->t1, v->stream));
    v->ran = true;
    return RZ_OK;
  }
  if (useGraph && !v->exec) {
    RZ_HIP(hipStreamBeginCapture(v->stream, hipStreamCaptureModeThreadLocal));
    s = enqueue_step(v, true);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(v->stream, &g);
    if (s != RZ_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return s != RZ_OK ? s : RZ_ERR_DEVICE;
    }
    v->graph = g;
    RZ_HIP(hipGraphInstantiate(&v->exec, v->graph, nullptr, nullptr, 0));
  }
  RZ_HIP(hipEventRecord(v->t0, v->stream));
  if (useGraph) {
    RZ_HIP(hipGraphLaunch(v->exec, v->stream));
  } else {
    s = enqueue_step(v, false);
    if (s != RZ_OK) return s;
  }
  RZ_HIP(hipEventRecord(v->t1, v->stream));
  v->ran = true;
  return RZ_OK;
}

extern "C" int rz_views_sync(rz_views* v, float* ms) {
  if (!v) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  if (ms) {
    *ms = 0.0f;
    if (v->ran) RZ_HIP(hipEventElapsedTime(ms, v->t0, v->t1));
  }
  // Every step ends with the query kernel, which copies the error word to
  // h_err: no copy here.  (Only steps set the word, and before any step
  // both words are 0.)
  const uint32_t err = *static_cast<volatile uint32_t*>(v->h_err);
  if (err) {
    RZ_HIP(hipMemset(v->d_err, 0, sizeof(uint32_t)));
    *static_cast<volatile uint32_t*>(v->h_err) = 0;
    return (err & kErrTimeout) ? RZ_ERR_TIMEOUT : RZ_ERR_DEVICE;
  }
  return RZ_OK;
}

extern "C" int rz_views_read_raw(rz_views* v, uint32_t view, uint16_t* depth, uint16_t* hiz) {
  if (!v || view >= v->views.size()) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  const ViewHost& h = v->views[view];
  const size_t nb = size_t(h.bX) * h.bY;
  if (depth) RZ_HIP(hipMemcpy(depth, h.d_depth, nb * 64 * 2, hipMemcpyDeviceToHost));
  if (hiz) RZ_HIP(hipMemcpy(hiz, h.d_hiz, (nb + 8) * 2, hipMemcpyDeviceToHost));
  return RZ_OK;
}

extern "C" int rz_views_read_queries(rz_views* v, uint32_t view, uint8_t* visible, uint8_t* needsClipping) {
  if (!v || view >= v->views.size()) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  const ViewHost& h = v->views[view];
  if (h.nQueries == 0) return RZ_OK;
  if (visible) RZ_HIP(hipMemcpy(visible, h.d_qvis, h.nQueries, hipMemcpyDeviceToHost));
  if (needsClipping) RZ_HIP(hipMemcpy(needsClipping, h.d_qclip, h.nQueries, hipMemcpyDeviceToHost));
  return RZ_OK;
}

extern "C" int rz_views_read_status(rz_views* v, uint32_t view, uint8_t* status) {
  if (!v || view >= v->views.size() || !status) return RZ_ERR_INVALID_ARGUMENT;
  if (v->dirty || !v->h_views) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  const ViewDesc& d = v->h_views[view];
  std::vector<ItemState> st(d.nItems);
  if (d.nItems)
    RZ_HIP(hipMemcpy(st.data(), v->d_state + d.itemBase, d.nItems * sizeof(ItemState), hipMemcpyDeviceToHost));
  std::vector<uint32_t> done(v->gated ? d.nItems : 0);
  if (v->gated && d.nItems) {
    RZ_HIP(hipMemcpy(done.data(), v->d_done + d.itemBase, d.nItems * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (v->f2b || (v->model && v->byOccluder)) {
      // The launch drew the items in sorted order: done word k is item order[k]'s.
      std::vector<uint32_t> order(d.nItems), byItem(d.nItems);
      RZ_HIP(hipMemcpy(order.data(), v->d_itemOrder + d.itemBase, d.nItems * 4, hipMemcpyDeviceToHost));
      for (uint32_t k = 0; k < d.nItems; ++k) byItem[order[k] - d.itemBase] = done[k];
      done.swap(byItem);
    }
  }
  for (uint32_t i = 0; i < d.nItems; ++i) {
    const uint32_t k = st[i].status;
    // A gated item was drawn (with rasterize<false>) when some tile saw it visible.
    status[i] = uint8_t(k == kStatusTest ? ((done[i] >> 16) != 0u ? 1u : 0u) : k);
  }
  return RZ_OK;
}

extern "C" int rz_views_set_query_gated(rz_views* v, int gated) {
  if (!v) return RZ_ERR_INVALID_ARGUMENT;
  int s = wait_last_step(v);
  if (s != RZ_OK) return s;
  v->gated = gated != 0;
  v->dirty = true;
  return RZ_OK;
}

extern "C" int rz_views_set_wait_timeout(rz_views* v, uint64_t ticks) {
  if (!v) return RZ_ERR_INVALID_ARGUMENT;
  v->timeoutTicks = ticks ? ticks : kGatedTimeoutTicks;
  return RZ_OK;
}

extern "C" int rz_views_enable_labels(rz_views* v, int enable) {
  if (!v) return RZ_ERR_INVALID_ARGUMENT;
  int s = wait_last_step(v);
  if (s != RZ_OK) return s;
  if (enable && !v->labels) {
    for (ViewHost& h : v->views) {
      const size_t n = size_t(h.bX) * h.bY * 64;
      if (!h.d_winner) {
        if (dev_alloc(&h.d_winner, n) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
        RZ_HIP(hipMemset(h.d_winner, 0xFF, n * sizeof(uint32_t)));
      }
    }
  }
  v->labels = enable != 0;
  v->dirty = true;
  return RZ_OK;
}

extern "C" int rz_views_set_model(rz_views* v, rz_mlp* model, int byOccluder) {
  if (!v) return RZ_ERR_INVALID_ARGUMENT;
  int s = wait_last_step(v);
  if (s != RZ_OK) return s;
  if (model && model->usesLabels && !v->labels) {
    s = rz_views_enable_labels(v, 1);  // features read the previous step's labels
    if (s != RZ_OK) return s;
  }
  if (model) (void)mlp_forward_max_groups();  // the occupancy query, before any captured step
  v->model = model;
  v->byOccluder = byOccluder != 0;
  v->dirty = true;
  return RZ_OK;
}

extern "C" int rz_views_read_order(rz_views* v, uint32_t view, uint32_t* order) {
  if (!v || !order || view >= v->views.size() || v->dirty || !v->h_views) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  const ViewDesc& d = v->h_views[view];
  if ((v->model && v->byOccluder) || v->f2b) {
    if (d.nItems) RZ_HIP(hipMemcpy(order, v->d_itemOrder + d.itemBase, d.nItems * 4, hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < d.nItems; ++k) order[k] -= d.itemBase;
    return RZ_OK;
  }
  if (!v->batchOrder) return RZ_ERR_INVALID_ARGUMENT;
  if (d.nBatches) RZ_HIP(hipMemcpy(order, v->d_order + d.batchBase, d.nBatches * 4, hipMemcpyDeviceToHost));
  for (uint32_t k = 0; k < d.nBatches; ++k) order[k] -= d.batchBase;
  return RZ_OK;
}

extern "C" int rz_views_set_eye(rz_views* v, uint32_t view, const float eye[3]) {
  if (!v || view >= v->views.size()) return RZ_ERR_INVALID_ARGUMENT;
  int s = wait_last_step(v);
  if (s != RZ_OK) return s;
  ViewHost& h = v->views[view];
  const bool on = eye != nullptr;
  if (on != (h.eye[3] != 0.0f)) v->dirty = true;  // switching the mode rebuilds the plan
  if (on) {
    h.eye[0] = eye[0];
    h.eye[1] = eye[1];
    h.eye[2] = eye[2];
  }
  h.eye[3] = on ? 1.0f : 0.0f;
  return RZ_OK;
}

extern "C" int rz_views_set_batch_order(rz_views* v, uint32_t view, uint32_t n, const uint32_t* order) {
  if (!v || view >= v->views.size() || (n && !order)) return RZ_ERR_INVALID_ARGUMENT;
  v->views[view].order.assign(order, order + n);
  v->dirty = true;
  return RZ_OK;
}

extern "C" int rz_views_read_labels(rz_views* v, uint32_t view, uint32_t* labels, float* features) {
  if (!v || view >= v->views.size() || v->dirty || !v->h_views) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  const ViewDesc& d = v->h_views[view];
  if (labels && d.nBatches)
    RZ_HIP(hipMemcpy(labels, v->d_labels + d.batchBase, d.nBatches * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (features && d.nBatches)
    RZ_HIP(hipMemcpy(features, v->d_features + size_t(d.batchBase) * kFeatures,
                     size_t(d.nBatches) * kFeatures * sizeof(float), hipMemcpyDeviceToHost));
  return RZ_OK;
}

extern "C" int rz_views_enable_stats(rz_views* v, int enable) {
  if (!v) return RZ_ERR_INVALID_ARGUMENT;
  int s = wait_last_step(v);
  if (s != RZ_OK) return s;
  if (enable && !v->d_stats) {
    if (dev_alloc(&v->d_stats, 3 + kFrameStatWords) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
    RZ_HIP(hipMemset(v->d_stats, 0, (3 + kFrameStatWords) * sizeof(unsigned long long)));
  }
  v->stats = enable != 0;
  v->dirty = true;
  return RZ_OK;
}

extern "C" int rz_views_read_stats(rz_views* v, uint64_t out[3]) {
  if (!v || !out || !v->d_stats) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  RZ_HIP(hipMemcpy(out, v->d_stats, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost));
  return RZ_OK;
}

extern "C" int rz_views_read_tile_stats(rz_views* v, uint64_t out[22]) {
  static_assert(kFrameStatWords == 22, "tile stats layout");
  if (!v || !out || !v->d_stats) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  RZ_HIP(hipMemcpy(out, v->d_stats + 3, kFrameStatWords * sizeof(uint64_t), hipMemcpyDeviceToHost));
  return RZ_OK;
}

extern "C" int rz_views_get_info(rz_views* v, uint32_t info[8]) {
  if (!v || !info) return RZ_ERR_INVALID_ARGUMENT;
  info[0] = uint32_t(v->views.size());
  info[1] = v->nItems;
  info[2] = v->nQuads;
  info[3] = v->nUnits;
  info[4] = v->nQueries;
  info[5] = v->nGroups;
  info[6] = (v->exec || v->execPre) ? 1u : 0u;
  info[7] = v->dirty ? 1u : 0u;
  return RZ_OK;
}

// ---------------------------------------------------------------------------
// Launch configuration and autotuning.
// ---------------------------------------------------------------------------
static std::string views_key(const rz_views* v) {
  uint64_t px = 0;
  for (const ViewHost& h : v->views) px += uint64_t(h.width) * h.height;
  int dev = 0;
  (void)hipGetDevice(&dev);
  return rzi::device_key(dev) + ",views," + std::to_string(v->views.size()) + "," + std::to_string(px);
}

extern "C" int rz_views_get_config(rz_views* v, rz_views_config* config) {
  if (!v || !config) return RZ_ERR_INVALID_ARGUMENT;
  config->tileW = uint32_t(v->tileB);
  config->waves = uint32_t(v->waves);
  config->groupsPerCU = uint32_t(v->groupsPerCU);
  config->ms = -1.0f;
  return RZ_OK;
}

extern "C" int rz_views_set_config(rz_views* v, uint32_t tileW, uint32_t waves, uint32_t groupsPerCU) {
  if (!v || (tileW != 8 && tileW != 16) || (waves != 2 && waves != 4 && waves != 8) || groupsPerCU > 64)
    return RZ_ERR_INVALID_ARGUMENT;
  int s = wait_last_step(v);
  if (s != RZ_OK) return s;
  v->tileB = int(tileW);
  v->waves = int(waves);
  v->groupsPerCU = int(groupsPerCU);
  v->dirty = true;
  return RZ_OK;
}

extern "C" int rz_views_load_config(rz_views* v, const char* cachePath, int* found) {
  if (!v || !cachePath || !found) return RZ_ERR_INVALID_ARGUMENT;
  *found = 0;
  std::string val;
  if (!rzi::tune_lookup(cachePath, views_key(v), val)) return RZ_OK;
  unsigned tileW = 0, waves = 0, gpc = 0;
  if (std::sscanf(val.c_str(), "tileW=%u waves=%u groupsPerCU=%u", &tileW, &waves, &gpc) != 3) return RZ_OK;
  const int s = rz_views_set_config(v, tileW, waves, gpc);
  if (s == RZ_OK) *found = 1;
  return s == RZ_ERR_INVALID_ARGUMENT ? RZ_OK : s;
}

// Everything a step produces, for the autotuner's identity check.
static int step_outputs(rz_views* v, std::vector<uint8_t>& out) {
  out.clear();
  for (uint32_t i = 0; i < v->views.size(); ++i) {
    const ViewHost& h = v->views[i];
    const size_t nb = size_t(h.bX) * h.bY;
    const size_t at = out.size();
    out.resize(at + (nb * 64 + nb + 8) * 2 + 2 * size_t(h.nQueries));
    uint16_t* d = reinterpret_cast<uint16_t*>(out.data() + at);
    int s = rz_views_read_raw(v, i, d, d + nb * 64);
    if (s == RZ_OK && h.nQueries) {
      uint8_t* q = out.data() + at + (nb * 64 + nb + 8) * 2;
      s = rz_views_read_queries(v, i, q, q + h.nQueries);
    }
    if (s != RZ_OK) return s;
    if (v->labels && v->h_views && v->h_views[i].nBatches) {
      const size_t at2 = out.size();
      out.resize(at2 + size_t(v->h_views[i].nBatches) * 4);
      s = rz_views_read_labels(v, i, reinterpret_cast<uint32_t*>(out.data() + at2), nullptr);
      if (s != RZ_OK) return s;
    }
  }
  return RZ_OK;
}

extern "C" int rz_views_tune(rz_views* v, uint32_t reps, const char* cachePath, rz_views_config candidates[12],
                             uint32_t* nCandidates) {
  if (!v || reps == 0) return RZ_ERR_INVALID_ARGUMENT;
  const int oldTile = v->tileB, oldWaves = v->waves, oldGpc = v->groupsPerCU;
  auto restore = [&](int code) {
    (void)rz_views_set_config(v, uint32_t(oldTile), uint32_t(oldWaves), uint32_t(oldGpc));
    return code;
  };
  // Labelled steps draw with 8x8-block tiles only.
  const bool labelled = v->labels || (v->model && v->model->usesLabels);
  std::vector<uint8_t> ref, got;
  rz_views_config all[12];
  uint32_t k = 0;
  int best = -1;
  for (uint32_t tileW : {8u, 16u}) {
    for (uint32_t waves : {2u, 4u, 8u}) {
      int maxGroups = 0, dev = 0, cus = 1;
      RZ_HIP(views_max_groups(int(tileW), int(waves), &maxGroups));
      RZ_HIP(hipGetDevice(&dev));
      RZ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      const uint32_t perCU = uint32_t(std::max(maxGroups / std::max(cus, 1), 1));
      for (uint32_t gpc : {1u, 0u}) {
        rz_views_config& cf = all[k];
        cf.tileW = tileW;
        cf.waves = waves;
        cf.groupsPerCU = gpc;
        cf.ms = -1.0f;
        const uint32_t idx = k++;
        if (gpc == 0 && perCU == 1) continue;  // same launch as one group per CU
        if (tileW == 16 && labelled) continue;
        // Each step is preceded by one with the same inputs, so that every
        // configuration's step starts from the same buffers (and labels).
        int s = rz_views_set_config(v, tileW, waves, gpc);
        float ms = 0.0f;
        for (int warm = 0; warm < 2 && s == RZ_OK; ++warm) {
          s = rz_views_run(v, 0);
          if (s == RZ_OK) s = rz_views_sync(v, &ms);
        }
        if (s == RZ_ERR_TOO_LARGE) continue;  // a gated view's tiles do not fit this grid
        if (s == RZ_OK && v->planTileB != int(tileW)) continue;  // moved to 16x16 tiles: measured as such
        if (s == RZ_OK) s = step_outputs(v, got);
        if (s != RZ_OK) return restore(s);
        if (ref.empty()) ref.swap(got);
        else if (got != ref) return restore(RZ_ERR_NOT_IDENTICAL);
        std::vector<float> t(reps);
        for (uint32_t r = 0; r < reps; ++r) {
          if ((s = rz_views_run(v, 1)) != RZ_OK || (s = rz_views_sync(v, &t[r])) != RZ_OK) return restore(s);
        }
        std::sort(t.begin(), t.end());
        cf.ms = t[reps / 2];
        if (best < 0 || cf.ms < all[best].ms) best = int(idx);
      }
    }
  }
  if (best < 0) return restore(RZ_ERR_DEVICE);
  int s = rz_views_set_config(v, all[best].tileW, all[best].waves, all[best].groupsPerCU);
  if (s != RZ_OK) return restore(s);
  if (cachePath) {
    const std::string val = "tileW=" + std::to_string(all[best].tileW) + " waves=" + std::to_string(all[best].waves) +
                            " groupsPerCU=" + std::to_string(all[best].groupsPerCU);
    if (!rzi::tune_store(cachePath, views_key(v), val)) return RZ_ERR_INVALID_ARGUMENT;
  }
  if (candidates)
    for (uint32_t i = 0; i < k; ++i) candidates[i] = all[i];
  if (nCandidates) *nCandidates = k;
  return RZ_OK;
}

extern "C" int rz_views_profile_enable(rz_views* v, int enable) {
  if (!v) return RZ_ERR_INVALID_ARGUMENT;
  int s = wait_last_step(v);
  if (s != RZ_OK) return s;
  v->profile = enable != 0;
  v->phRecorded = false;
  return RZ_OK;
}

extern "C" int rz_views_profile_read(rz_views* v, float ms[3]) {
  if (!v || !ms) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  if (!v->phRecorded) return RZ_ERR_INVALID_ARGUMENT;
  for (int k = 0; k < 3; ++k) RZ_HIP(hipEventElapsedTime(&ms[k], v->ph[k], v->ph[k + 1]));
  return RZ_OK;
}

extern "C" int rz_views_read_items(rz_views* v, uint32_t view, void* state, void* setup) {
  if (!v || view >= v->views.size() || v->dirty || !v->h_views) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(v->stream));
  const ViewDesc& d = v->h_views[view];
  if (state && d.nItems)
    RZ_HIP(hipMemcpy(state, v->d_state + d.itemBase, size_t(d.nItems) * sizeof(ItemState), hipMemcpyDeviceToHost));
  if (setup && d.nItems) {
    std::vector<ItemDesc> items(d.nItems);
    RZ_HIP(hipMemcpy(items.data(), v->d_items + d.itemBase, items.size() * sizeof(ItemDesc), hipMemcpyDeviceToHost));
    const uint32_t q0 = items.front().quadBase;
    const uint32_t q1 = items.back().quadBase + items.back().quadCount;
    RZ_HIP(hipMemcpy(setup, v->d_recs + q0, size_t(q1 - q0) * sizeof(QuadRec), hipMemcpyDeviceToHost));
  }
  return RZ_OK;
}
