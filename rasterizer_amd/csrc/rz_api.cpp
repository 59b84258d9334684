// C ABI of the MI355X rasterizer (include/rasterizer_amd.h): contexts,
// device buffers, and the enqueue logic around the gfx950 kernels.
#include "rasterizer_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "rz_host.h"
#include "rz_internal.h"
#include "rz_kernels.h"

using namespace rzd;

namespace {

// Host staging ring of item tables: draw d uses slot d % kStagingSlots.  An
// event is recorded after every kDrawsPerEvent-th draw, alternating between
// two events.
constexpr int kStagingSlots = 16;
constexpr uint64_t kDrawsPerEvent = 8;
constexpr int kProfSlots = 256;
constexpr uint64_t kTimeoutTicks = 200000000ull;  // 2 s of the 100 MHz s_memrealtime clock

template <typename T>
hipError_t dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

}  // namespace

namespace rzi {

void report_hip_error(hipError_t error, const char* call, const char* file, int line) {
  static const bool on = std::getenv("RZ_DEBUG") != nullptr;
  if (on) std::fprintf(stderr, "rasterizer_amd: %s failed: %s (%s:%d)\n", call, hipGetErrorString(error), file, line);
}

// Both tables depend only on this host (libm, RCPPS); build them once.
const HostTables& host_tables() {
  static HostTables t;
  static std::once_flag once;
  std::call_once(once, [] {
    if (!rzh::build_coverage_table(t.cov)) t.status = RZ_ERR_UNSUPPORTED_HOST;
    else if (rzh::build_rcp_table(t.rcp) != 0) t.status = RZ_ERR_UNSUPPORTED_HOST;
    if (rzh::build_rsqrt_table(t.rsqrt) != 0) t.rsqrtStatus = RZ_ERR_UNSUPPORTED_HOST;
  });
  return t;
}

}  // namespace rzi

struct rz_context {
  uint32_t width = 0, height = 0, bX = 0, bY = 0;
  int tileW = 8, tilesX = 0, tilesY = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  uint16_t* d_depth = nullptr;
  uint16_t* d_hiz = nullptr;
  uint16_t* d_tileMin = nullptr;  // min hiZ per 8x8-block tile, kept equal to hiZ's
  uint32_t pyrX = 0, pyrY = 0;
  uint64_t* d_cov = nullptr;
  uint32_t* d_rcp = nullptr;
  uint32_t* d_err = nullptr;
  Mvp mvp{};
  // per-draw buffers
  uint32_t capItems = 0, capQuads = 0;
  ItemDesc* d_items = nullptr;
  ItemState* d_state = nullptr;
  uint32_t* d_done = nullptr;
  uint32_t* d_visible = nullptr;
  uint8_t* d_status = nullptr;
  QuadRec* d_recs = nullptr;
  uint64_t* d_bbox = nullptr;
  uint64_t* d_groupBox = nullptr;
  uint32_t* d_groupItem = nullptr;  // per group of 64 quads: the item of its first quad (written by prep)
  // host staging ring for item tables: pinned, coherent host memory that
  // the prep kernel reads through d_hItems (its device mapping)
  ItemDesc* h_items[kStagingSlots] = {};
  const ItemDesc* d_hItems[kStagingSlots] = {};
  uint32_t h_itemsCap[kStagingSlots] = {};
  hipEvent_t ev_items[2] = {};
  uint64_t draws = 0;        // draws staged through the ring
  bool stagingSync = false;  // a draw's staging did not complete: synchronise before the next
  // query staging
  float* d_qbounds = nullptr;
  uint8_t* d_qvis = nullptr;
  uint8_t* d_qclip = nullptr;
  uint32_t capQueries = 0;
  uint32_t* d_rgba = nullptr;
  // pending frame status copy
  uint8_t* h_status = nullptr;
  uint32_t h_statusCap = 0;
  uint8_t* pendingStatusDst = nullptr;
  uint32_t pendingStatusN = 0;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // per-kernel profiling: a ring of event quadruples (prep, setup, frame,
  // end) for draws and pairs for query batches
  bool profiling = false;
  std::vector<hipEvent_t> profEvents;  // kProfSlots * 4, created on first enable
  int profKind[kProfSlots] = {};       // 0 draw, 1 query batch
  int profUsed = 0;
  double profMs[4] = {0, 0, 0, 0};
  uint32_t profCounts[2] = {0, 0};
  uint64_t* d_frameStats = nullptr;  // per-tile frame-kernel counters when enabled
  bool cooperative = true;           // RZ_FRAME_LAUNCH=plain at creation: plain dispatch of the same grid
  int waves = 8;                     // wave64s per frame-kernel tile
  uint64_t timeoutTicks = kTimeoutTicks;  // bound of each cross-tile wait (rz_set_wait_timeout)
};

extern "C" const char* rz_status_string(int s) {
  switch (s) {
    case RZ_OK: return "ok";
    case RZ_ERR_INVALID_ARGUMENT: return "invalid argument";
    case RZ_ERR_OUT_OF_MEMORY: return "out of device or host memory";
    case RZ_ERR_DEVICE: return "HIP runtime or device error";
    case RZ_ERR_TIMEOUT: return "device-side wait timed out";
    case RZ_ERR_UNSUPPORTED_HOST: return "host CPU RCPPS has an unsupported structure";
    case RZ_ERR_TOO_LARGE: return "buffer too large for one co-resident frame launch";
    case RZ_ERR_NOT_IDENTICAL: return "launch configurations disagree (autotuner check)";
    case RZ_ERR_IO: return "file cannot be written or read, or is not this format, version or checksum";
    default: return "unknown status";
  }
}

static void free_context(rz_context* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  hipFree(c->d_depth);
  hipFree(c->d_hiz);
  hipFree(c->d_tileMin);
  hipFree(c->d_cov);
  hipFree(c->d_rcp);
  hipFree(c->d_err);
  hipFree(c->d_items);
  hipFree(c->d_state);
  hipFree(c->d_done);
  hipFree(c->d_visible);
  hipFree(c->d_status);
  hipFree(c->d_recs);
  hipFree(c->d_bbox);
  hipFree(c->d_groupBox);
  hipFree(c->d_groupItem);
  hipFree(c->d_qbounds);
  hipFree(c->d_qvis);
  hipFree(c->d_qclip);
  hipFree(c->d_rgba);
  hipFree(c->d_frameStats);
  for (int i = 0; i < kStagingSlots; ++i)
    if (c->h_items[i]) (void)hipHostFree(c->h_items[i]);
  for (hipEvent_t e : c->ev_items)
    if (e) (void)hipEventDestroy(e);
  if (c->h_status) (void)hipHostFree(c->h_status);
  for (hipEvent_t e : c->profEvents) (void)hipEventDestroy(e);
  if (c->t0) (void)hipEventDestroy(c->t0);
  if (c->t1) (void)hipEventDestroy(c->t1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

static int noalloc_init(int device);

extern "C" int rz_create(uint32_t width, uint32_t height, rz_context** out) {
  if (!out) return RZ_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (width == 0 || height == 0 || (width % 8) || (height % 8) || width > 65528 || height > 65528)
    return RZ_ERR_INVALID_ARGUMENT;
  if (uint64_t(width / 8) * (height / 8) >= (1u << 24)) return RZ_ERR_TOO_LARGE;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RZ_ERR_DEVICE;
  const rzi::HostTables& tabs = rzi::host_tables();
  if (tabs.status != RZ_OK) return tabs.status;

  rz_context* c = new (std::nothrow) rz_context();
  if (!c) return RZ_ERR_OUT_OF_MEMORY;
  c->width = width;
  c->height = height;
  c->bX = width / 8;
  c->bY = height / 8;
  auto fail = [&](int code) {
    free_context(c);
    return code;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return fail(RZ_ERR_DEVICE);
  if (noalloc_init(c->device) != RZ_OK) return fail(RZ_ERR_DEVICE);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(RZ_ERR_DEVICE);
  const size_t nBlocks = size_t(c->bX) * c->bY;
  c->pyrX = (c->bX + 7) / 8;
  c->pyrY = (c->bY + 7) / 8;
  if (dev_alloc(&c->d_depth, nBlocks * 64) != hipSuccess || dev_alloc(&c->d_hiz, nBlocks + 8) != hipSuccess ||
      dev_alloc(&c->d_tileMin, size_t(c->pyrX) * c->pyrY) != hipSuccess ||
      dev_alloc(&c->d_cov, 4096) != hipSuccess || dev_alloc(&c->d_rcp, 4096) != hipSuccess ||
      dev_alloc(&c->d_err, 1) != hipSuccess)
    return fail(RZ_ERR_OUT_OF_MEMORY);
  if (hipMemset(c->d_depth, 0, nBlocks * 64 * sizeof(uint16_t)) != hipSuccess ||
      hipMemset(c->d_hiz, 0, (nBlocks + 8) * sizeof(uint16_t)) != hipSuccess ||
      hipMemset(c->d_tileMin, 0, size_t(c->pyrX) * c->pyrY * sizeof(uint16_t)) != hipSuccess ||
      hipMemset(c->d_err, 0, sizeof(uint32_t)) != hipSuccess ||
      hipMemcpy(c->d_cov, tabs.cov, sizeof tabs.cov, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_rcp, tabs.rcp, sizeof tabs.rcp, hipMemcpyHostToDevice) != hipSuccess)
    return fail(RZ_ERR_DEVICE);

  // Pick the smallest tile whose grid fits co-resident on this device.
  bool placed = false;
  for (int tw : {8, 16}) {
    int maxGroups = 0;
    if (frame_max_groups(tw, c->waves, &maxGroups) != hipSuccess) return fail(RZ_ERR_DEVICE);
    const int tx = int((c->bX + tw - 1) / tw), ty = int((c->bY + tw - 1) / tw);
    if (tx * ty <= maxGroups) {
      c->tileW = tw;
      c->tilesX = tx;
      c->tilesY = ty;
      placed = true;
      break;
    }
  }
  if (!placed) return fail(RZ_ERR_TOO_LARGE);
  for (hipEvent_t& e : c->ev_items)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(RZ_ERR_DEVICE);
  if (hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess) return fail(RZ_ERR_DEVICE);
  if (const char* mode = std::getenv("RZ_FRAME_LAUNCH")) c->cooperative = std::strcmp(mode, "plain") != 0;
  if (const char* cache = rzi::tune_cache_env()) {
    int found = 0;
    (void)rz_load_config(c, cache, &found);  // a stale or unusable entry leaves the default
  }
  // Identity MVP until the caller sets one.
  const float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  rzh::prebake_mvp(width, height, ident, c->mvp.prebaked, c->mvp.raw);
  *out = c;
  return RZ_OK;
}

extern "C" int rz_destroy(rz_context* ctx) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  free_context(ctx);
  return RZ_OK;
}

extern "C" int rz_set_mvp(rz_context* ctx, const float matrix[16]) {
  if (!ctx || !matrix) return RZ_ERR_INVALID_ARGUMENT;
  rzh::prebake_mvp(ctx->width, ctx->height, matrix, ctx->mvp.prebaked, ctx->mvp.raw);
  return RZ_OK;
}

extern "C" int rz_clear(rz_context* ctx) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  // Whole 8-entry groups only (Rasterizer.cpp:115-120): the last
  // (blocks + 8) % 8 padding entries keep their value.
  RZ_HIP(launch_clear(ctx->d_hiz, ((ctx->bX * ctx->bY + 8) / 8) * 8, ctx->d_tileMin, ctx->pyrX * ctx->pyrY,
                      ctx->stream));
  return RZ_OK;
}

namespace {

int occluder_create(rz_context* ctx, const uint32_t* packets, uint32_t packetCount, const float refMin[4],
                    const float refMax[4], const float boundsMin[4], const float boundsMax[4], hipMemcpyKind kind,
                    rz_occluder** out);

}  // namespace

extern "C" int rz_occluder_create(rz_context* ctx, const uint32_t* packets, uint32_t packetCount,
                                  const float refMin[4], const float refMax[4], const float boundsMin[4],
                                  const float boundsMax[4], rz_occluder** out) {
  return occluder_create(ctx, packets, packetCount, refMin, refMax, boundsMin, boundsMax, hipMemcpyHostToDevice, out);
}

extern "C" int rz_occluder_create_device(rz_context* ctx, const uint32_t* d_packets, uint32_t packetCount,
                                         const float refMin[4], const float refMax[4], const float boundsMin[4],
                                         const float boundsMax[4], rz_occluder** out) {
  return occluder_create(ctx, d_packets, packetCount, refMin, refMax, boundsMin, boundsMax, hipMemcpyDeviceToDevice,
                         out);
}

namespace {

int occluder_create(rz_context* ctx, const uint32_t* packets, uint32_t packetCount, const float refMin[4],
                    const float refMax[4], const float boundsMin[4], const float boundsMax[4], hipMemcpyKind kind,
                    rz_occluder** out) {
  if (!ctx || !out || !refMin || !refMax || !boundsMin || !boundsMax || (packetCount % 4) != 0 ||
      (packetCount && !packets))
    return RZ_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  rz_occluder* o = new (std::nothrow) rz_occluder();
  if (!o) return RZ_ERR_OUT_OF_MEMORY;
  o->ctx = ctx;
  o->packetCount = packetCount;
  std::memcpy(o->refMin, refMin, 16);
  std::memcpy(o->refMax, refMax, 16);
  std::memcpy(o->boundsMin, boundsMin, 16);
  std::memcpy(o->boundsMax, boundsMax, 16);
  o->d_packets = nullptr;
  if (packetCount) {
    if (hipMalloc(reinterpret_cast<void**>(&o->d_packets), size_t(packetCount) * 32) != hipSuccess) {
      delete o;
      return RZ_ERR_OUT_OF_MEMORY;
    }
    if (hipMemcpy(o->d_packets, packets, size_t(packetCount) * 32, kind) != hipSuccess) {
      hipFree(o->d_packets);
      delete o;
      return RZ_ERR_DEVICE;
    }
  }
  *out = o;
  return RZ_OK;
}

}  // namespace

extern "C" int rz_occluder_destroy(rz_occluder* o) {
  if (!o) return RZ_ERR_INVALID_ARGUMENT;
  // Any context on this device may still have a draw in flight that reads it.
  (void)hipDeviceSynchronize();
  hipFree(o->d_packets);
  delete o;
  return RZ_OK;
}

// Folds every recorded profiling slot into the running sums.
static int profile_flush(rz_context* c) {
  for (int i = 0; i < c->profUsed; ++i) {
    hipEvent_t* ev = &c->profEvents[size_t(i) * 4];
    const int last = c->profKind[i] == 0 ? 3 : 1;
    RZ_HIP(hipEventSynchronize(ev[last]));
    float ms = 0.0f;
    if (c->profKind[i] == 0) {
      for (int k = 0; k < 3; ++k) {
        RZ_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
        c->profMs[k] += ms;
      }
      c->profCounts[0]++;
    } else {
      RZ_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
      c->profMs[3] += ms;
      c->profCounts[1]++;
    }
  }
  c->profUsed = 0;
  return RZ_OK;
}

// Next profiling slot (or nullptr when profiling is off).
static hipEvent_t* profile_slot(rz_context* c, int kind) {
  if (!c->profiling) return nullptr;
  if (c->profUsed == kProfSlots && profile_flush(c) != RZ_OK) return nullptr;
  c->profKind[c->profUsed] = kind;
  return &c->profEvents[size_t(c->profUsed++) * 4];
}

static int ensure_capacity(rz_context* c, uint32_t nItems, uint32_t nQuads) {
  if (nItems > c->capItems || nQuads > c->capQuads) {
    RZ_HIP(hipStreamSynchronize(c->stream));
  }
  if (nItems > c->capItems) {
    const uint32_t cap = std::max<uint32_t>(nItems, std::max<uint32_t>(64, c->capItems * 2));
    hipFree(c->d_items);
    hipFree(c->d_state);
    hipFree(c->d_done);
    hipFree(c->d_visible);
    hipFree(c->d_status);
    c->capItems = 0;
    if (dev_alloc(&c->d_items, cap) != hipSuccess || dev_alloc(&c->d_state, cap) != hipSuccess ||
        dev_alloc(&c->d_done, cap) != hipSuccess || dev_alloc(&c->d_visible, cap) != hipSuccess ||
        dev_alloc(&c->d_status, cap) != hipSuccess)
      return RZ_ERR_OUT_OF_MEMORY;
    c->capItems = cap;
  }
  if (nQuads > c->capQuads) {
    const uint32_t cap = std::max<uint32_t>(nQuads, std::max<uint32_t>(1024, c->capQuads + c->capQuads / 2));
    hipFree(c->d_recs);
    hipFree(c->d_bbox);
    hipFree(c->d_groupBox);
  hipFree(c->d_groupItem);
    c->capQuads = 0;
    if (dev_alloc(&c->d_recs, cap) != hipSuccess || dev_alloc(&c->d_bbox, cap) != hipSuccess ||
        dev_alloc(&c->d_groupBox, (cap + 63) / 64) != hipSuccess ||
        dev_alloc(&c->d_groupItem, (cap + 63) / 64) != hipSuccess)
      return RZ_ERR_OUT_OF_MEMORY;
    c->capQuads = cap;
  }
  return RZ_OK;
}

static int flush_pending_status(rz_context* c) {
  if (!c->pendingStatusDst) return RZ_OK;
  RZ_HIP(hipStreamSynchronize(c->stream));
  std::memcpy(c->pendingStatusDst, c->h_status, c->pendingStatusN);
  c->pendingStatusDst = nullptr;
  c->pendingStatusN = 0;
  return RZ_OK;
}

// Device buffers one draw uses besides the context's depth, hiZ and tables.
struct DrawBufs {
  ItemDesc* items;
  ItemState* state;
  uint32_t* done;
  uint32_t* visible;
  QuadRec* recs;
  uint64_t* bbox;
  uint64_t* groupBox;
  uint32_t* groupItem;  // optional: prep writes it, setup starts its item search there
};

// Total quads of occ[0..n) (8 per batch of 4 packets); RZ_ERR_INVALID_ARGUMENT
// on a null handle.
static int count_quads(uint32_t n, const rz_occluder* const* occ, uint64_t* totalQuads) {
  uint64_t t = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!occ[i]) return RZ_ERR_INVALID_ARGUMENT;
    t += uint64_t(occ[i]->packetCount / 4) * 8;
  }
  *totalQuads = t;
  return RZ_OK;
}

// Item descriptors of occ[0..n) with their quads numbered in order.
static void fill_items(ItemDesc* items, uint32_t n, const rz_occluder* const* occ, const uint32_t* kinds) {
  uint32_t base = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const rz_occluder* o = occ[i];
    ItemDesc& d = items[i];
    d.packets = o->d_packets;
    d.quadBase = base;
    d.quadCount = (o->packetCount / 4) * 8;
    d.kind = kinds ? kinds[i] : uint32_t(kItemQuery);
    d.view = 0;
    std::memcpy(d.refMin, o->refMin, 16);
    std::memcpy(d.refMax, o->refMax, 16);
    std::memcpy(d.boundsMin, o->boundsMin, 16);
    std::memcpy(d.boundsMax, o->boundsMax, 16);
    base += d.quadCount;
  }
}

// Enqueues the prep, setup and frame kernels of one draw whose item
// descriptors are already in b.items (or enqueued before them), or, with
// hostItems, in mapped host memory from which prep copies them to b.items.
// prof: optional event quadruple (prep, setup, frame, end).
static int enqueue_kernels(rz_context* c, uint32_t n, uint32_t totalQuads, const DrawBufs& b, hipEvent_t* prof,
                           const ItemDesc* hostItems = nullptr) {
  if (prof) RZ_HIP(hipEventRecord(prof[0], c->stream));
  PrepParams pp{};
  pp.items = b.items;
  pp.state = b.state;
  pp.nItems = n;
  pp.width = c->width;
  pp.height = c->height;
  pp.tileW = uint32_t(c->tileW);
  pp.tileH = uint32_t(c->tileW);
  pp.rcp = c->d_rcp;
  pp.done = b.done;
  pp.visible = b.visible;
  pp.mvp = c->mvp;
  pp.hostItems = hostItems;
  pp.itemsOut = b.items;
  pp.groupItem = b.groupItem;
  RZ_HIP(launch_prep(pp, c->stream));
  if (prof) RZ_HIP(hipEventRecord(prof[1], c->stream));

  SetupParams sp{};
  sp.items = b.items;
  sp.state = b.state;
  sp.nItems = n;
  sp.nQuads = totalQuads;
  sp.bX = c->bX;
  sp.bY = c->bY;
  sp.rcp = c->d_rcp;
  sp.recs = b.recs;
  sp.bbox = b.bbox;
  sp.groupBox = b.groupBox;
  sp.groupItem = b.groupItem;
  RZ_HIP(launch_setup(sp, c->stream));
  if (prof) RZ_HIP(hipEventRecord(prof[2], c->stream));

  FrameParams fp{};
  fp.items = b.items;
  fp.state = b.state;
  fp.nItems = n;
  fp.bX = c->bX;
  fp.bY = c->bY;
  fp.tilesX = uint32_t(c->tilesX);
  fp.tilesY = uint32_t(c->tilesY);
  fp.recs = b.recs;
  fp.bbox = b.bbox;
  fp.groupBox = b.groupBox;
  fp.done = b.done;
  fp.visible = b.visible;
  fp.depth = c->d_depth;
  fp.hiz = c->d_hiz;
  fp.cov = c->d_cov;
  fp.err = c->d_err;
  fp.timeoutTicks = c->timeoutTicks;
  fp.tileMin = c->d_tileMin;
  fp.pyrX = c->pyrX;
  fp.stats = c->d_frameStats;
  RZ_HIP(launch_frame(fp, c->tileW, c->stream, c->cooperative, c->waves));
  if (prof) RZ_HIP(hipEventRecord(prof[3], c->stream));
  return RZ_OK;
}

// Enqueues prep, setup and frame kernels for `n` items of the given kinds.
static int enqueue_draw(rz_context* c, uint32_t n, const rz_occluder* const* occ, const uint32_t* kinds,
                        uint8_t* statusOut) {
  if (n == 0) return RZ_OK;
  uint64_t totalQuads = 0;
  int s = count_quads(n, occ, &totalQuads);
  if (s != RZ_OK) return s;
  if (totalQuads >= (1ull << 31)) return RZ_ERR_TOO_LARGE;
  if (statusOut) {
    s = flush_pending_status(c);
    if (s != RZ_OK) return s;
  }
  s = ensure_capacity(c, n, uint32_t(totalQuads));
  if (s != RZ_OK) return s;

  // Host staging: draw d reuses the slot of draw d - kStagingSlots, whose
  // prep must have read it.  The last event recorded at or after that draw
  // follows draw e = (d - kStagingSlots) | 7, between d - 16 and d - 9, and
  // still holds that record: the same event is recorded next after draw
  // e + 16 >= d, and draw d records after this wait.  Waiting for it waits
  // for a frame kernel after draw d - 16's prep.  (An event per draw put a
  // marker between two kernels of every draw.)
  static_assert(kStagingSlots == 2 * kDrawsPerEvent && kDrawsPerEvent == 8, "ring and event spacing");
  const uint64_t d = c->draws++;
  const int slot = int(d % kStagingSlots);
  // Until this draw's kernels and event are enqueued, a failure leaves the
  // ring's events behind its draws: the next draw then synchronises.
  const bool syncAll = c->stagingSync;
  c->stagingSync = true;
  if (syncAll) {
    RZ_HIP(hipStreamSynchronize(c->stream));
  } else if (d >= uint64_t(kStagingSlots)) {
    const uint64_t e = (d - kStagingSlots) | (kDrawsPerEvent - 1);
    RZ_HIP(hipEventSynchronize(c->ev_items[(e / kDrawsPerEvent) & 1]));
  }
  if (c->h_itemsCap[slot] < n) {
    if (c->h_items[slot]) (void)hipHostFree(c->h_items[slot]);
    c->h_items[slot] = nullptr;
    c->d_hItems[slot] = nullptr;
    c->h_itemsCap[slot] = 0;
    const uint32_t cap = std::max<uint32_t>(n, 64);
    // Coherent (fine-grained): the prep kernel reads each descriptor once,
    // over the bus, with nothing of an earlier use of the slot in a cache.
    if (hipHostMalloc(reinterpret_cast<void**>(&c->h_items[slot]), cap * sizeof(ItemDesc),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
      return RZ_ERR_OUT_OF_MEMORY;
    void* mapped = nullptr;
    RZ_HIP(hipHostGetDevicePointer(&mapped, c->h_items[slot], 0));
    c->d_hItems[slot] = static_cast<const ItemDesc*>(mapped);
    c->h_itemsCap[slot] = cap;
  }
  fill_items(c->h_items[slot], n, occ, kinds);
  const DrawBufs bufs{c->d_items, c->d_state,    c->d_done,     c->d_visible,
                      c->d_recs,  c->d_bbox,     c->d_groupBox, c->d_groupItem};
  s = enqueue_kernels(c, n, uint32_t(totalQuads), bufs, profile_slot(c, 0), c->d_hItems[slot]);
  if (s != RZ_OK) return s;
  if (d % kDrawsPerEvent == kDrawsPerEvent - 1)
    RZ_HIP(hipEventRecord(c->ev_items[(d / kDrawsPerEvent) & 1], c->stream));
  c->stagingSync = false;

  if (statusOut) {
    if (c->h_statusCap < n) {
      if (c->h_status) (void)hipHostFree(c->h_status);
      c->h_status = nullptr;
      c->h_statusCap = 0;
      if (hipHostMalloc(reinterpret_cast<void**>(&c->h_status), std::max<uint32_t>(n, 64)) != hipSuccess)
        return RZ_ERR_OUT_OF_MEMORY;
      c->h_statusCap = std::max<uint32_t>(n, 64);
    }
    RZ_HIP(launch_status(c->d_state, c->d_done, n, c->d_status, c->stream));
    RZ_HIP(hipMemcpyAsync(c->h_status, c->d_status, n, hipMemcpyDeviceToHost, c->stream));
    c->pendingStatusDst = statusOut;
    c->pendingStatusN = n;
  }
  return RZ_OK;
}

extern "C" int rz_rasterize(rz_context* ctx, const rz_occluder* occluder, int possiblyNearClipped) {
  if (!ctx || !occluder) return RZ_ERR_INVALID_ARGUMENT;
  const uint32_t kind = possiblyNearClipped ? kItemForceClip : kItemForceNoClip;
  return enqueue_draw(ctx, 1, &occluder, &kind, nullptr);
}

extern "C" int rz_draw_frame(rz_context* ctx, uint32_t n, const rz_occluder* const* occluders, uint8_t* statusOut) {
  if (!ctx || (n && !occluders)) return RZ_ERR_INVALID_ARGUMENT;
  return enqueue_draw(ctx, n, occluders, nullptr, statusOut);
}

// ---------------------------------------------------------------------------
// rz_draw_frame without allocation.  The per-draw buffers (item descriptors
// and states, done and visibility words, quad setup records, quad and group
// boxes) live in `scratch`, a static device array that every GPU the library
// runs on holds for the life of the process.  Its address on each device is
// taken once, with hipGetSymbolAddress, when the first context is created on
// that device; host code uses only that address, never the array's name.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t kNaItemsOff = 0;
constexpr size_t kNaStateOff = kNaItemsOff + align256(sizeof(ItemDesc) * RZ_NOALLOC_MAX_ITEMS);
constexpr size_t kNaDoneOff = kNaStateOff + align256(sizeof(ItemState) * RZ_NOALLOC_MAX_ITEMS);
constexpr size_t kNaVisibleOff = kNaDoneOff + align256(sizeof(uint32_t) * RZ_NOALLOC_MAX_ITEMS);
constexpr size_t kNaRecsOff = kNaVisibleOff + align256(sizeof(uint32_t) * RZ_NOALLOC_MAX_ITEMS);
constexpr size_t kNaBboxOff = kNaRecsOff + align256(sizeof(QuadRec) * size_t(RZ_NOALLOC_MAX_QUADS));
constexpr size_t kNaGroupOff = kNaBboxOff + align256(sizeof(uint64_t) * size_t(RZ_NOALLOC_MAX_QUADS));
constexpr size_t kNaBytes = kNaGroupOff + align256(sizeof(uint64_t) * size_t(RZ_NOALLOC_MAX_QUADS / 64));
constexpr int kNaMaxDevices = 64;

std::mutex g_naMutex;
void* g_naScratch[kNaMaxDevices] = {};  // scratch's address on each device, once known
ItemDesc g_naItems[RZ_NOALLOC_MAX_ITEMS];  // host staging of the item descriptors (guarded by g_naMutex)

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// Takes scratch's address on the current device (idempotent per device).
static int noalloc_init(int device) {
  if (device < 0 || device >= kNaMaxDevices) return RZ_ERR_DEVICE;
  std::lock_guard<std::mutex> lock(g_naMutex);
  if (g_naScratch[device]) return RZ_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return RZ_ERR_DEVICE;
  g_naScratch[device] = ptr;
  return RZ_OK;
}

extern "C" int rz_draw_frame_noalloc(rz_context* ctx, uint32_t n, const rz_occluder* const* occluders,
                                     uint8_t* d_statusOut) {
  if (!ctx || (n && !occluders)) return RZ_ERR_INVALID_ARGUMENT;
  if (n == 0) return RZ_OK;
  if (n > RZ_NOALLOC_MAX_ITEMS) return RZ_ERR_TOO_LARGE;
  uint64_t totalQuads = 0;
  int s = count_quads(n, occluders, &totalQuads);
  if (s != RZ_OK) return s;
  if (totalQuads > RZ_NOALLOC_MAX_QUADS) return RZ_ERR_TOO_LARGE;
  if (ctx->device < 0 || ctx->device >= kNaMaxDevices) return RZ_ERR_DEVICE;
  std::lock_guard<std::mutex> lock(g_naMutex);
  unsigned char* base = static_cast<unsigned char*>(g_naScratch[ctx->device]);
  if (!base) return RZ_ERR_DEVICE;  // set by rz_create on this device
  const DrawBufs bufs{reinterpret_cast<ItemDesc*>(base + kNaItemsOff), reinterpret_cast<ItemState*>(base + kNaStateOff),
                      reinterpret_cast<uint32_t*>(base + kNaDoneOff), reinterpret_cast<uint32_t*>(base + kNaVisibleOff),
                      reinterpret_cast<QuadRec*>(base + kNaRecsOff), reinterpret_cast<uint64_t*>(base + kNaBboxOff),
                      reinterpret_cast<uint64_t*>(base + kNaGroupOff), nullptr};
  fill_items(g_naItems, n, occluders, nullptr);
  // From pageable memory: the copy has consumed g_naItems when it returns.
  RZ_HIP(hipMemcpyAsync(bufs.items, g_naItems, n * sizeof(ItemDesc), hipMemcpyHostToDevice, ctx->stream));
  // The prep kernel writes every state, done and visibility word of items
  // 0..n-1 and the setup kernel every record and box of their quads, so
  // nothing left in scratch by an earlier call is read.
  s = enqueue_kernels(ctx, n, uint32_t(totalQuads), bufs, nullptr);
  if (s != RZ_OK) return s;
  if (d_statusOut) RZ_HIP(launch_status(bufs.state, bufs.done, n, d_statusOut, ctx->stream));
  return RZ_OK;
}

extern "C" int rz_set_wait_timeout(rz_context* ctx, uint64_t ticks) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  ctx->timeoutTicks = ticks ? ticks : kTimeoutTicks;
  return RZ_OK;
}

// Waits for the stream, then reports and clears the device error word.
extern "C" int rz_sync(rz_context* ctx) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipStreamSynchronize(ctx->stream));
  int s = flush_pending_status(ctx);
  if (s != RZ_OK) return s;
  uint32_t err = 0;
  RZ_HIP(hipMemcpy(&err, ctx->d_err, sizeof err, hipMemcpyDeviceToHost));
  if (err) {
    RZ_HIP(hipMemset(ctx->d_err, 0, sizeof(uint32_t)));
    return (err & kErrTimeout) ? RZ_ERR_TIMEOUT : RZ_ERR_DEVICE;
  }
  return RZ_OK;
}

static int ensure_queries(rz_context* c, uint32_t n) {
  if (n <= c->capQueries) return RZ_OK;
  RZ_HIP(hipStreamSynchronize(c->stream));
  hipFree(c->d_qbounds);
  hipFree(c->d_qvis);
  hipFree(c->d_qclip);
  c->capQueries = 0;
  const uint32_t cap = std::max<uint32_t>(n, 1024);
  if (dev_alloc(&c->d_qbounds, size_t(cap) * 8) != hipSuccess || dev_alloc(&c->d_qvis, cap) != hipSuccess ||
      dev_alloc(&c->d_qclip, cap) != hipSuccess)
    return RZ_ERR_OUT_OF_MEMORY;
  c->capQueries = cap;
  return RZ_OK;
}

static QueryParams query_params(rz_context* c, const float* bounds, uint32_t n, uint8_t* vis, uint8_t* clip,
                                uint8_t* code) {
  QueryParams q{};
  q.bounds = bounds;
  q.n = n;
  q.width = c->width;
  q.height = c->height;
  q.bX = c->bX;
  q.depth = c->d_depth;
  q.hiz = c->d_hiz;
  q.tileMin = c->d_tileMin;
  q.bY = c->bY;
  q.pyrX = c->pyrX;
  q.rcp = c->d_rcp;
  q.visible = vis;
  q.needsClipping = clip;
  q.code = code;
  q.mvp = c->mvp;
  return q;
}

extern "C" int rz_query(rz_context* ctx, const float boundsMin[4], const float boundsMax[4], int* visible,
                        int* needsClipping) {
  if (!ctx || !boundsMin || !boundsMax || !visible || !needsClipping) return RZ_ERR_INVALID_ARGUMENT;
  int s = ensure_queries(ctx, 1);
  if (s != RZ_OK) return s;
  float b[8];
  std::memcpy(b, boundsMin, 16);
  std::memcpy(b + 4, boundsMax, 16);
  RZ_HIP(hipMemcpyAsync(ctx->d_qbounds, b, sizeof b, hipMemcpyHostToDevice, ctx->stream));
  QueryParams q = query_params(ctx, ctx->d_qbounds, 1, ctx->d_qvis, nullptr, ctx->d_qclip);
  RZ_HIP(launch_query(q, ctx->stream));
  uint8_t code = 0;
  RZ_HIP(hipMemcpyAsync(&code, ctx->d_qclip, 1, hipMemcpyDeviceToHost, ctx->stream));
  s = rz_sync(ctx);
  if (s != RZ_OK) return s;
  *visible = code & 1;
  if (code & 4) *needsClipping = (code >> 1) & 1;
  return RZ_OK;
}

extern "C" int rz_query2d(rz_context* ctx, uint32_t minX, uint32_t maxX, uint32_t minY, uint32_t maxY,
                          uint32_t maxZ, int* visible) {
  if (!ctx || !visible || minX > maxX || minY > maxY || maxX >= ctx->width || maxY >= ctx->height)
    return RZ_ERR_INVALID_ARGUMENT;
  int s = ensure_queries(ctx, 1);
  if (s != RZ_OK) return s;
  QueryParams q = query_params(ctx, nullptr, 0, ctx->d_qvis, nullptr, nullptr);
  RZ_HIP(launch_query2d(minX, maxX, minY, maxY, maxZ, q, ctx->d_qvis, ctx->stream));
  uint8_t v = 0;
  RZ_HIP(hipMemcpyAsync(&v, ctx->d_qvis, 1, hipMemcpyDeviceToHost, ctx->stream));
  s = rz_sync(ctx);
  if (s != RZ_OK) return s;
  *visible = v;
  return RZ_OK;
}

extern "C" int rz_query_batch(rz_context* ctx, uint32_t n, const float* bounds, uint8_t* visible,
                              uint8_t* needsClipping) {
  if (!ctx || (n && (!bounds || !visible))) return RZ_ERR_INVALID_ARGUMENT;
  if (n == 0) return RZ_OK;
  int s = ensure_queries(ctx, n);
  if (s != RZ_OK) return s;
  RZ_HIP(hipMemcpyAsync(ctx->d_qbounds, bounds, size_t(n) * 32, hipMemcpyHostToDevice, ctx->stream));
  QueryParams q = query_params(ctx, ctx->d_qbounds, n, ctx->d_qvis, ctx->d_qclip, nullptr);
  hipEvent_t* prof = profile_slot(ctx, 1);
  if (prof) RZ_HIP(hipEventRecord(prof[0], ctx->stream));
  RZ_HIP(launch_query(q, ctx->stream));
  if (prof) RZ_HIP(hipEventRecord(prof[1], ctx->stream));
  RZ_HIP(hipMemcpyAsync(visible, ctx->d_qvis, n, hipMemcpyDeviceToHost, ctx->stream));
  if (needsClipping) RZ_HIP(hipMemcpyAsync(needsClipping, ctx->d_qclip, n, hipMemcpyDeviceToHost, ctx->stream));
  return rz_sync(ctx);
}

extern "C" int rz_query_batch_device(rz_context* ctx, uint32_t n, const float* d_bounds, uint8_t* d_visible,
                                     uint8_t* d_needsClipping) {
  if (!ctx || (n && (!d_bounds || !d_visible))) return RZ_ERR_INVALID_ARGUMENT;
  QueryParams q = query_params(ctx, d_bounds, n, d_visible, d_needsClipping, nullptr);
  hipEvent_t* prof = profile_slot(ctx, 1);
  if (prof) RZ_HIP(hipEventRecord(prof[0], ctx->stream));
  RZ_HIP(launch_query(q, ctx->stream));
  if (prof) RZ_HIP(hipEventRecord(prof[1], ctx->stream));
  return RZ_OK;
}

extern "C" int rz_readback_depth(rz_context* ctx, void* rgba) {
  if (!ctx || !rgba) return RZ_ERR_INVALID_ARGUMENT;
  const size_t px = size_t(ctx->width) * ctx->height;
  if (!ctx->d_rgba && dev_alloc(&ctx->d_rgba, px) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
  RZ_HIP(launch_readback(ctx->d_depth, ctx->d_hiz, ctx->width, ctx->height, ctx->bX, ctx->d_rgba, ctx->stream));
  RZ_HIP(hipMemcpyAsync(rgba, ctx->d_rgba, px * 4, hipMemcpyDeviceToHost, ctx->stream));
  return rz_sync(ctx);
}

extern "C" int rz_readback_raw(rz_context* ctx, uint16_t* depth, uint16_t* hiz) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  const size_t nb = size_t(ctx->bX) * ctx->bY;
  if (depth) RZ_HIP(hipMemcpyAsync(depth, ctx->d_depth, nb * 128, hipMemcpyDeviceToHost, ctx->stream));
  if (hiz) RZ_HIP(hipMemcpyAsync(hiz, ctx->d_hiz, (nb + 8) * 2, hipMemcpyDeviceToHost, ctx->stream));
  return rz_sync(ctx);
}

extern "C" int rz_upload_raw(rz_context* ctx, const uint16_t* depth, const uint16_t* hiz) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  const size_t nb = size_t(ctx->bX) * ctx->bY;
  if (depth) RZ_HIP(hipMemcpyAsync(ctx->d_depth, depth, nb * 128, hipMemcpyHostToDevice, ctx->stream));
  if (hiz) {
    RZ_HIP(hipMemcpyAsync(ctx->d_hiz, hiz, (nb + 8) * 2, hipMemcpyHostToDevice, ctx->stream));
    RZ_HIP(launch_tile_min(ctx->d_hiz, ctx->bX, ctx->bY, ctx->d_tileMin, ctx->pyrX, ctx->pyrY, ctx->stream));
  }
  return rz_sync(ctx);
}

extern "C" int rz_get_stream(rz_context* ctx, void** stream) {
  if (!ctx || !stream) return RZ_ERR_INVALID_ARGUMENT;
  *stream = ctx->stream;
  return RZ_OK;
}

extern "C" int rz_get_info(rz_context* ctx, uint32_t info[6]) {
  if (!ctx || !info) return RZ_ERR_INVALID_ARGUMENT;
  info[0] = ctx->width;
  info[1] = ctx->height;
  info[2] = ctx->bX;
  info[3] = ctx->bY;
  info[4] = uint32_t(ctx->tileW);
  info[5] = uint32_t(ctx->tileW);
  return RZ_OK;
}

extern "C" int rz_get_tables(rz_context* ctx, uint64_t* coverage4096, uint32_t* rcp4096) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  if (coverage4096) RZ_HIP(hipMemcpy(coverage4096, ctx->d_cov, 4096 * 8, hipMemcpyDeviceToHost));
  if (rcp4096) RZ_HIP(hipMemcpy(rcp4096, ctx->d_rcp, 4096 * 4, hipMemcpyDeviceToHost));
  return RZ_OK;
}

extern "C" int rz_get_mvp(rz_context* ctx, float prebaked[16], float raw[16]) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  if (prebaked) std::memcpy(prebaked, ctx->mvp.prebaked, 64);
  if (raw) std::memcpy(raw, ctx->mvp.raw, 64);
  return RZ_OK;
}

extern "C" int rz_bake(const float* vertices, uint32_t n, const float refMin[4], const float refMax[4],
                       uint32_t* packets, float boundsMin[4], float boundsMax[4], float center[4]) {
  if (!vertices || !refMin || !refMax || !packets || !boundsMin || !boundsMax || !center)
    return RZ_ERR_INVALID_ARGUMENT;
  return rzh::bake(vertices, n, refMin, refMax, packets, boundsMin, boundsMax, center) == 0 ? RZ_OK
                                                                                             : RZ_ERR_INVALID_ARGUMENT;
}

extern "C" int rz_bake_device(rz_context* ctx, uint32_t n, const uint32_t* vertexOffsets, const float* d_vertices,
                              const float* refMinMax, uint32_t* d_packets, float* d_bounds) {
  if (!ctx || (n && (!vertexOffsets || !d_vertices || !refMinMax || !d_packets || !d_bounds)))
    return RZ_ERR_INVALID_ARGUMENT;
  if (n == 0) return RZ_OK;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t a = vertexOffsets[i], b = vertexOffsets[i + 1];
    if (b <= a || (b - a) % 32 || a % 32) return RZ_ERR_INVALID_ARGUMENT;
  }
  if (vertexOffsets[n] >= (1u << 31)) return RZ_ERR_TOO_LARGE;
  const rzi::HostTables& tabs = rzi::host_tables();
  if (tabs.rsqrtStatus != RZ_OK) return tabs.rsqrtStatus;
  const size_t quads = vertexOffsets[n] / 4;
  BakeParams p{};
  uint32_t* d_offsets = nullptr;
  float* d_ref = nullptr;
  uint32_t* d_rsqrt = nullptr;
  int status = RZ_OK;
  if (dev_alloc(&d_offsets, size_t(n) + 1) != hipSuccess || dev_alloc(&d_ref, size_t(n) * 8) != hipSuccess ||
      dev_alloc(&d_rsqrt, 8192) != hipSuccess || dev_alloc(&p.normals, quads) != hipSuccess ||
      dev_alloc(&p.cluster, quads) != hipSuccess || dev_alloc(&p.order, quads) != hipSuccess) {
    status = RZ_ERR_OUT_OF_MEMORY;
  } else if (hipMemcpyAsync(d_offsets, vertexOffsets, (size_t(n) + 1) * 4, hipMemcpyHostToDevice, ctx->stream) !=
                 hipSuccess ||
             hipMemcpyAsync(d_ref, refMinMax, size_t(n) * 32, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
             hipMemcpyAsync(d_rsqrt, tabs.rsqrt, sizeof tabs.rsqrt, hipMemcpyHostToDevice, ctx->stream) !=
                 hipSuccess) {
    status = RZ_ERR_DEVICE;
  } else {
    p.vertices = d_vertices;
    p.offsets = d_offsets;
    p.ref = d_ref;
    p.packets = d_packets;
    p.bounds = d_bounds;
    p.rsqrt = d_rsqrt;
    if (launch_bake(p, n, ctx->stream) != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
      status = RZ_ERR_DEVICE;
  }
  (void)hipStreamSynchronize(ctx->stream);
  hipFree(d_offsets);
  hipFree(d_ref);
  hipFree(d_rsqrt);
  hipFree(p.normals);
  hipFree(p.cluster);
  hipFree(p.order);
  return status;
}

extern "C" int rz_quad_decompose_device(rz_context* ctx, uint32_t nTriangles, const uint32_t* d_indices,
                                        const float* d_vertices, uint32_t nVertices, uint32_t* d_out,
                                        uint32_t* nOut) {
  if (!ctx || !nOut || (nTriangles && (!d_indices || !d_vertices || !d_out))) return RZ_ERR_INVALID_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(d_vertices) % 16u != 0u) return RZ_ERR_INVALID_ARGUMENT;  // read as float4
  *nOut = 0;
  if (nTriangles == 0) return RZ_OK;
  const rzi::HostTables& tabs = rzi::host_tables();
  if (tabs.rsqrtStatus != RZ_OK) return tabs.rsqrtStatus;
  return quad_decompose_run(ctx->stream, nTriangles, d_indices, d_vertices, nVertices, tabs.rsqrt, d_out, nOut);
}

extern "C" int rz_sah_batches(rz_context* ctx, uint32_t n, const float* d_boxes, uint32_t targetSize,
                              uint32_t splitGranularity, uint32_t* d_indices, uint32_t* batchOffsets,
                              uint32_t* nBatches) {
  if (!ctx || !batchOffsets || !nBatches || (n && (!d_boxes || !d_indices))) return RZ_ERR_INVALID_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(d_boxes) % 16u != 0u) return RZ_ERR_INVALID_ARGUMENT;  // read as float4
  return sah_batches_run(ctx->stream, n, d_boxes, targetSize, splitGranularity, d_indices, batchOffsets, nBatches);
}

namespace {

// Main.cpp:90-128 on the device for nQuads quads (4 xyzw vertices each, in
// HBM), after the arguments are checked: the scene box over d_boxVertices
// (nBox vertices) when given, else over the quads' vertices.
int prepare_scene(rz_context* ctx, uint32_t nQuads, const float* d_vertices, const float* d_boxVertices,
                  uint32_t nBox, uint32_t targetSize, uint32_t splitGranularity, uint32_t* d_packets,
                  float* d_bounds, uint32_t* quadOffsets, uint32_t* nOccluders, float refMinMax[8]) {
  const uint32_t nPad = (nQuads + 7u) & ~7u;  // Main.cpp:90-94
  float* d_boxes = nullptr;
  float* d_sorted = nullptr;
  uint32_t* d_order = nullptr;
  uint32_t* d_keys = nullptr;
  int status = RZ_OK;
  // Scene box as ordered-float keys (min xyzw, max xyzw), then a word that
  // the boxes kernel sets when a vertex is not finite or exceeds 2^40.
  uint32_t keys[9] = {~0u, ~0u, ~0u, ~0u, 0u, 0u, 0u, 0u, 0u};
  if (dev_alloc(&d_boxes, size_t(nPad) * 8) != hipSuccess || dev_alloc(&d_sorted, size_t(nPad) * 16) != hipSuccess ||
      dev_alloc(&d_order, nPad) != hipSuccess || dev_alloc(&d_keys, 9) != hipSuccess) {
    status = RZ_ERR_OUT_OF_MEMORY;
  } else if (hipMemcpyAsync(d_keys, keys, sizeof keys, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
             launch_scene_boxes(d_vertices, nQuads, nPad, d_boxes, d_keys, ctx->stream) != hipSuccess ||
             (d_boxVertices && launch_vertex_keys(d_boxVertices, nBox, d_keys, ctx->stream) != hipSuccess) ||
             hipMemcpyAsync(keys, d_keys, sizeof keys, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
             hipStreamSynchronize(ctx->stream) != hipSuccess) {
    status = RZ_ERR_DEVICE;
  } else if (keys[8] != 0u) {
    status = RZ_ERR_INVALID_ARGUMENT;
  } else {
    for (int l = 0; l < 8; ++l) {
      const uint32_t k = keys[l];
      const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
      std::memcpy(&refMinMax[l], &u, 4);
    }
    uint32_t nOcc = 0;
    status = sah_batches_run(ctx->stream, nPad, d_boxes, targetSize, splitGranularity, d_order, quadOffsets, &nOcc);
    if (status == RZ_OK) {
      if (launch_scene_gather(d_vertices, nQuads, nPad, d_order, d_sorted, ctx->stream) != hipSuccess) {
        status = RZ_ERR_DEVICE;
      } else {
        std::vector<uint32_t> vertexOffsets(size_t(nOcc) + 1);
        std::vector<float> refs(size_t(nOcc) * 8);
        for (uint32_t i = 0; i <= nOcc; ++i) vertexOffsets[i] = quadOffsets[i] * 4u;
        for (uint32_t i = 0; i < nOcc; ++i) std::memcpy(&refs[size_t(i) * 8], refMinMax, 32);
        status = rz_bake_device(ctx, nOcc, vertexOffsets.data(), d_sorted, refs.data(), d_packets, d_bounds);
        if (status == RZ_OK) *nOccluders = nOcc;
      }
    }
  }
  (void)hipStreamSynchronize(ctx->stream);
  hipFree(d_boxes);
  hipFree(d_sorted);
  hipFree(d_order);
  hipFree(d_keys);
  return status;
}

}  // namespace

extern "C" int rz_prepare_scene_device(rz_context* ctx, uint32_t nQuads, const float* d_vertices,
                                       uint32_t targetSize, uint32_t splitGranularity, uint32_t* d_packets,
                                       float* d_bounds, uint32_t* quadOffsets, uint32_t* nOccluders,
                                       float refMinMax[8]) {
  if (!ctx || !quadOffsets || !nOccluders || !refMinMax || (nQuads && (!d_vertices || !d_packets || !d_bounds)))
    return RZ_ERR_INVALID_ARGUMENT;
  // Occluder::bake takes whole 8-quad batches, so every split must fall on one.
  if (splitGranularity == 0 || splitGranularity % 8u != 0u) return RZ_ERR_INVALID_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(d_vertices) % 16u != 0u) return RZ_ERR_INVALID_ARGUMENT;
  quadOffsets[0] = 0;
  *nOccluders = 0;
  if (nQuads == 0) return RZ_OK;
  if (nQuads > (1u << 28)) return RZ_ERR_TOO_LARGE;
  const rzi::HostTables& tabs = rzi::host_tables();
  if (tabs.rsqrtStatus != RZ_OK) return tabs.rsqrtStatus;
  return prepare_scene(ctx, nQuads, d_vertices, nullptr, 0, targetSize, splitGranularity, d_packets, d_bounds,
                       quadOffsets, nOccluders, refMinMax);
}

extern "C" int rz_prepare_mesh_device(rz_context* ctx, uint32_t nTriangles, const uint32_t* d_indices,
                                      const float* d_vertices, uint32_t nVertices, uint32_t targetSize,
                                      uint32_t splitGranularity, uint32_t* d_packets, float* d_bounds,
                                      uint32_t* quadOffsets, uint32_t* nOccluders, float refMinMax[8]) {
  if (!ctx || !quadOffsets || !nOccluders || !refMinMax ||
      (nTriangles && (!d_indices || !d_vertices || !d_packets || !d_bounds)))
    return RZ_ERR_INVALID_ARGUMENT;
  if (splitGranularity == 0 || splitGranularity % 8u != 0u) return RZ_ERR_INVALID_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(d_vertices) % 16u != 0u) return RZ_ERR_INVALID_ARGUMENT;
  quadOffsets[0] = 0;
  *nOccluders = 0;
  if (nTriangles == 0) return RZ_OK;
  if (nTriangles > (1u << 28)) return RZ_ERR_TOO_LARGE;
  const rzi::HostTables& tabs = rzi::host_tables();
  if (tabs.rsqrtStatus != RZ_OK) return tabs.rsqrtStatus;
  uint32_t* d_quadIdx = nullptr;
  float* d_quadVertices = nullptr;
  uint32_t nWords = 0;
  int status = RZ_OK;
  // Main.cpp:86: the quads' indices, then their vertices in quad order.
  if (dev_alloc(&d_quadIdx, size_t(nTriangles) * 4) != hipSuccess) status = RZ_ERR_OUT_OF_MEMORY;
  else
    status = quad_decompose_run(ctx->stream, nTriangles, d_indices, d_vertices, nVertices, tabs.rsqrt, d_quadIdx,
                                &nWords);
  if (status == RZ_OK && nWords > 0) {
    if (dev_alloc(&d_quadVertices, size_t(nWords) * 4) != hipSuccess)
      status = RZ_ERR_OUT_OF_MEMORY;
    else if (launch_quad_gather(d_vertices, d_quadIdx, nWords, d_quadVertices, ctx->stream) != hipSuccess)
      status = RZ_ERR_DEVICE;
    else
      status = prepare_scene(ctx, nWords / 4u, d_quadVertices, d_vertices, nVertices, targetSize, splitGranularity,
                             d_packets, d_bounds, quadOffsets, nOccluders, refMinMax);
  }
  (void)hipStreamSynchronize(ctx->stream);
  hipFree(d_quadIdx);
  hipFree(d_quadVertices);
  return status;
}

extern "C" int rz_timer_start(rz_context* ctx) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipEventRecord(ctx->t0, ctx->stream));
  return RZ_OK;
}

extern "C" int rz_timer_stop(rz_context* ctx, float* ms) {
  if (!ctx || !ms) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipEventRecord(ctx->t1, ctx->stream));
  RZ_HIP(hipEventSynchronize(ctx->t1));
  RZ_HIP(hipEventElapsedTime(ms, ctx->t0, ctx->t1));
  return RZ_OK;
}

extern "C" int rz_profile_enable(rz_context* ctx, int enable) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  if (enable && ctx->profEvents.empty()) {
    ctx->profEvents.resize(size_t(kProfSlots) * 4, nullptr);
    for (hipEvent_t& e : ctx->profEvents) RZ_HIP(hipEventCreate(&e));
  }
  if (enable) {
    int s = profile_flush(ctx);
    if (s != RZ_OK) return s;
    std::fill(ctx->profMs, ctx->profMs + 4, 0.0);
    ctx->profCounts[0] = ctx->profCounts[1] = 0;
  }
  ctx->profiling = enable != 0;
  return RZ_OK;
}

extern "C" int rz_profile_read(rz_context* ctx, double ms[4], uint32_t counts[2]) {
  if (!ctx || !ms || !counts) return RZ_ERR_INVALID_ARGUMENT;
  int s = profile_flush(ctx);
  if (s != RZ_OK) return s;
  std::copy(ctx->profMs, ctx->profMs + 4, ms);
  counts[0] = ctx->profCounts[0];
  counts[1] = ctx->profCounts[1];
  return RZ_OK;
}

extern "C" int rz_device_alloc(rz_context* ctx, size_t bytes, const void* init, void** devptr) {
  if (!ctx || !devptr || bytes == 0) return RZ_ERR_INVALID_ARGUMENT;
  *devptr = nullptr;
  if (hipMalloc(devptr, bytes) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
  const hipError_t e = init ? hipMemcpy(*devptr, init, bytes, hipMemcpyHostToDevice) : hipMemset(*devptr, 0, bytes);
  if (e != hipSuccess) {
    hipFree(*devptr);
    *devptr = nullptr;
    return RZ_ERR_DEVICE;
  }
  return RZ_OK;
}

extern "C" int rz_device_read(rz_context* ctx, const void* devptr, size_t bytes, void* dst) {
  if (!ctx || !devptr || !dst) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipMemcpyAsync(dst, devptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  return rz_sync(ctx);
}

extern "C" int rz_device_free(void* devptr) {
  if (!devptr) return RZ_ERR_INVALID_ARGUMENT;
  RZ_HIP(hipDeviceSynchronize());
  RZ_HIP(hipFree(devptr));
  return RZ_OK;
}

extern "C" int rz_frame_stats_enable(rz_context* ctx, int enable) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  if (enable && !ctx->d_frameStats) {
    const size_t n = size_t(ctx->tilesX) * ctx->tilesY * kFrameStatWords;
    if (dev_alloc(&ctx->d_frameStats, n) != hipSuccess) return RZ_ERR_OUT_OF_MEMORY;
    RZ_HIP(hipMemset(ctx->d_frameStats, 0, n * sizeof(uint64_t)));
  } else if (!enable && ctx->d_frameStats) {
    RZ_HIP(hipStreamSynchronize(ctx->stream));
    hipFree(ctx->d_frameStats);
    ctx->d_frameStats = nullptr;
  }
  return RZ_OK;
}

extern "C" int rz_frame_stats_read(rz_context* ctx, uint64_t* out, uint32_t* count) {
  if (!ctx || !count) return RZ_ERR_INVALID_ARGUMENT;
  const uint32_t tiles = uint32_t(ctx->tilesX * ctx->tilesY);
  if (!ctx->d_frameStats) {
    *count = 0;
    return RZ_OK;
  }
  if (out && *count >= tiles) {
    RZ_HIP(hipMemcpyAsync(out, ctx->d_frameStats, size_t(tiles) * kFrameStatWords * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, ctx->stream));
    int s = rz_sync(ctx);
    if (s != RZ_OK) return s;
  }
  *count = tiles;
  return RZ_OK;
}

// ---------------------------------------------------------------------------
// Launch configuration and autotuning.
// ---------------------------------------------------------------------------
static std::string frame_key(const rz_context* c) {
  return rzi::device_key(c->device) + ",frame," + std::to_string(c->width) + "x" + std::to_string(c->height);
}

// Switches the frame kernel to (tileW, waves) if its tile grid can be
// co-resident; reallocates the per-tile counters when the tile count changes.
static int apply_config(rz_context* c, int tw, int waves) {
  if ((tw != 8 && tw != 16) || (waves != 2 && waves != 4 && waves != 8)) return RZ_ERR_INVALID_ARGUMENT;
  int maxGroups = 0;
  RZ_HIP(frame_max_groups(tw, waves, &maxGroups));
  const int tx = int((c->bX + tw - 1) / tw), ty = int((c->bY + tw - 1) / tw);
  if (tx * ty > maxGroups) return RZ_ERR_TOO_LARGE;
  RZ_HIP(hipStreamSynchronize(c->stream));
  const bool stats = c->d_frameStats != nullptr;
  if (stats) {
    const int s = rz_frame_stats_enable(c, 0);
    if (s != RZ_OK) return s;
  }
  c->tileW = tw;
  c->waves = waves;
  c->tilesX = tx;
  c->tilesY = ty;
  return stats ? rz_frame_stats_enable(c, 1) : RZ_OK;
}

extern "C" int rz_get_config(rz_context* ctx, rz_frame_config* config) {
  if (!ctx || !config) return RZ_ERR_INVALID_ARGUMENT;
  config->tileW = uint32_t(ctx->tileW);
  config->waves = uint32_t(ctx->waves);
  config->ms = -1.0f;
  return RZ_OK;
}

extern "C" int rz_set_config(rz_context* ctx, uint32_t tileW, uint32_t waves) {
  if (!ctx) return RZ_ERR_INVALID_ARGUMENT;
  return apply_config(ctx, int(tileW), int(waves));
}

extern "C" int rz_load_config(rz_context* ctx, const char* cachePath, int* found) {
  if (!ctx || !cachePath || !found) return RZ_ERR_INVALID_ARGUMENT;
  *found = 0;
  std::string v;
  if (!rzi::tune_lookup(cachePath, frame_key(ctx), v)) return RZ_OK;
  int tw = 0, waves = 0;
  if (std::sscanf(v.c_str(), "tileW=%d waves=%d", &tw, &waves) != 2) return RZ_OK;
  const int s = apply_config(ctx, tw, waves);
  if (s == RZ_OK) *found = 1;
  return (s == RZ_ERR_TOO_LARGE || s == RZ_ERR_INVALID_ARGUMENT) ? RZ_OK : s;
}

extern "C" int rz_tune_frame(rz_context* ctx, uint32_t n, const rz_occluder* const* occ, uint32_t reps,
                             const char* cachePath, rz_frame_config candidates[6], uint32_t* nCandidates) {
  if (!ctx || (n && !occ) || reps == 0) return RZ_ERR_INVALID_ARGUMENT;
  rz_context* c = ctx;
  const int oldTw = c->tileW, oldWaves = c->waves;
  auto restore = [&](int code) {
    (void)apply_config(c, oldTw, oldWaves);
    return code;
  };
  const size_t nb = size_t(c->bX) * c->bY;
  std::vector<uint16_t> d(nb * 64), h(nb + 8), refD, refH;
  std::vector<uint8_t> st(std::max<uint32_t>(n, 1)), refS;
  rz_frame_config all[6];
  uint32_t k = 0;
  int best = -1;
  for (int tw : {8, 16})
    for (int waves : {2, 4, 8}) {
      rz_frame_config& cf = all[k];
      cf.tileW = uint32_t(tw);
      cf.waves = uint32_t(waves);
      cf.ms = -1.0f;
      const uint32_t idx = k++;
      int s = apply_config(c, tw, waves);
      if (s == RZ_ERR_TOO_LARGE) continue;
      if (s != RZ_OK) return restore(s);
      if ((s = rz_clear(c)) != RZ_OK || (s = rz_draw_frame(c, n, occ, st.data())) != RZ_OK ||
          (s = rz_sync(c)) != RZ_OK || (s = rz_readback_raw(c, d.data(), h.data())) != RZ_OK)
        return restore(s);
      if (refD.empty()) {
        refD = d;
        refH = h;
        refS = st;
      } else if (d != refD || h != refH || st != refS) {
        return restore(RZ_ERR_NOT_IDENTICAL);
      }
      std::vector<float> t(reps);
      for (uint32_t r = 0; r < reps; ++r) {
        if ((s = rz_clear(c)) != RZ_OK || (s = rz_timer_start(c)) != RZ_OK ||
            (s = rz_draw_frame(c, n, occ, nullptr)) != RZ_OK || (s = rz_timer_stop(c, &t[r])) != RZ_OK)
          return restore(s);
      }
      std::sort(t.begin(), t.end());
      cf.ms = t[reps / 2];
      if (best < 0 || cf.ms < all[best].ms) best = int(idx);
    }
  if (best < 0) return restore(RZ_ERR_TOO_LARGE);
  int s = apply_config(c, int(all[best].tileW), int(all[best].waves));
  if (s != RZ_OK) return restore(s);
  if (cachePath) {
    const std::string v = "tileW=" + std::to_string(all[best].tileW) + " waves=" + std::to_string(all[best].waves);
    if (!rzi::tune_store(cachePath, frame_key(c), v)) return RZ_ERR_INVALID_ARGUMENT;
  }
  if (candidates)
    for (uint32_t i = 0; i < k; ++i) candidates[i] = all[i];
  if (nCandidates) *nCandidates = k;
  return RZ_OK;
}
