// On-disk formats: baked occluder sets (Occluder::m_vertexData packets with
// their reference boxes and bounds, Occluder.cpp:153-176) and frame dumps
// (raw depth, hiZ and query bits).  Host code only.
//
// Both files are little-endian:
//   magic[8]  "RZOCC\0\0\0" or "RZFRM\0\0\0"
//   u32       format version (1)
//   u32       reserved (0)
//   u64       payload bytes
//   payload
//   u64       FNV-1a 64 of the payload
// Occluder-set payload: u32 count, then per occluder u32 packetCount,
// f32 refMin[4], refMax[4], boundsMin[4], boundsMax[4], u32 words[packetCount * 8].
// Frame payload: u32 width, height, nQueries, then u16 depth[bX * bY * 64]
// (block-linear, as rz_readback_raw), u16 hiz[bX * bY + 8], u8 visible[n],
// u8 needsClipping[n].
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rasterizer_amd.h"

namespace {

constexpr char kMagicOcc[8] = {'R', 'Z', 'O', 'C', 'C', 0, 0, 0};
constexpr char kMagicFrame[8] = {'R', 'Z', 'F', 'R', 'M', 0, 0, 0};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 8 + 4 + 4 + 8;

uint64_t fnv1a(const uint8_t* p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

struct Writer {
  std::vector<uint8_t> b;
  void raw(const void* p, size_t n) {
    const uint8_t* c = static_cast<const uint8_t*>(p);
    b.insert(b.end(), c, c + n);
  }
  void u32(uint32_t v) { raw(&v, 4); }
};

// Writes magic, header, payload and checksum; RZ_ERR_IO on any failure.
int write_file(const char* path, const char (&magic)[8], const std::vector<uint8_t>& payload) {
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return RZ_ERR_IO;
  const uint32_t version = kVersion, reserved = 0;
  const uint64_t size = payload.size(), sum = fnv1a(payload.data(), payload.size());
  bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&version, 4, 1, f) == 1 &&
            std::fwrite(&reserved, 4, 1, f) == 1 && std::fwrite(&size, 8, 1, f) == 1 &&
            (payload.empty() || std::fwrite(payload.data(), 1, payload.size(), f) == payload.size()) &&
            std::fwrite(&sum, 8, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? RZ_OK : RZ_ERR_IO;
}

// Reads and checks a whole file: magic, version, payload size, checksum.
int read_file(const char* path, const char (&magic)[8], std::vector<uint8_t>& payload) {
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return RZ_ERR_IO;
  std::vector<uint8_t> all;
  uint8_t buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) all.insert(all.end(), buf, buf + got);
  const bool err = std::ferror(f) != 0;
  std::fclose(f);
  if (err || all.size() < kHeader + 8 || std::memcmp(all.data(), magic, 8) != 0) return RZ_ERR_IO;
  uint32_t version;
  uint64_t size, sum;
  std::memcpy(&version, all.data() + 8, 4);
  std::memcpy(&size, all.data() + 16, 8);
  if (version != kVersion || size != all.size() - kHeader - 8) return RZ_ERR_IO;
  std::memcpy(&sum, all.data() + kHeader + size, 8);
  if (fnv1a(all.data() + kHeader, size) != sum) return RZ_ERR_IO;
  payload.assign(all.begin() + kHeader, all.begin() + kHeader + size);
  return RZ_OK;
}

struct Reader {
  const std::vector<uint8_t>& b;
  size_t at = 0;
  bool ok = true;
  void raw(void* p, size_t n) {
    if (!ok || b.size() - at < n) {
      ok = false;
      return;
    }
    std::memcpy(p, b.data() + at, n);
    at += n;
  }
  uint32_t u32() {
    uint32_t v = 0;
    raw(&v, 4);
    return v;
  }
};

}  // namespace

extern "C" int rz_write_occluders(const char* path, uint32_t n, const rz_occluder_data* occluders) {
  if (!path || (n && !occluders)) return RZ_ERR_INVALID_ARGUMENT;
  Writer w;
  w.u32(n);
  for (uint32_t i = 0; i < n; ++i) {
    const rz_occluder_data& o = occluders[i];
    if ((o.packetCount % 4) || (o.packetCount && !o.packets)) return RZ_ERR_INVALID_ARGUMENT;
    w.u32(o.packetCount);
    w.raw(o.refMin, 16);
    w.raw(o.refMax, 16);
    w.raw(o.boundsMin, 16);
    w.raw(o.boundsMax, 16);
    w.raw(o.packets, size_t(o.packetCount) * 32);
  }
  return write_file(path, kMagicOcc, w.b);
}

extern "C" int rz_read_occluders(const char* path, uint32_t* n, uint32_t* totalPackets, rz_occluder_data* occluders,
                                 uint32_t* packetWords) {
  if (!path || !n || !totalPackets || (occluders && !packetWords)) return RZ_ERR_INVALID_ARGUMENT;
  std::vector<uint8_t> payload;
  int s = read_file(path, kMagicOcc, payload);
  if (s != RZ_OK) return s;
  // First pass: sizes and structure.
  Reader r{payload};
  const uint32_t count = r.u32();
  uint64_t packets = 0;
  for (uint32_t i = 0; i < count && r.ok; ++i) {
    const uint32_t pc = r.u32();
    float skip[16];
    r.raw(skip, 64);
    if (!r.ok || pc % 4 || uint64_t(pc) * 32 > payload.size() - r.at) return RZ_ERR_IO;
    r.at += size_t(pc) * 32;
    packets += pc;
  }
  if (!r.ok || r.at != payload.size() || packets > 0xFFFFFFFFull) return RZ_ERR_IO;
  if (!occluders) {
    *n = count;
    *totalPackets = uint32_t(packets);
    return RZ_OK;
  }
  if (*n < count || *totalPackets < packets) return RZ_ERR_INVALID_ARGUMENT;
  Reader q{payload};
  q.u32();
  uint32_t* words = packetWords;
  for (uint32_t i = 0; i < count; ++i) {
    rz_occluder_data& o = occluders[i];
    o.packetCount = q.u32();
    q.raw(o.refMin, 16);
    q.raw(o.refMax, 16);
    q.raw(o.boundsMin, 16);
    q.raw(o.boundsMax, 16);
    q.raw(words, size_t(o.packetCount) * 32);
    o.packets = words;
    words += size_t(o.packetCount) * 8;
  }
  *n = count;
  *totalPackets = uint32_t(packets);
  return RZ_OK;
}

extern "C" int rz_write_frame(const char* path, uint32_t width, uint32_t height, const uint16_t* depth,
                              const uint16_t* hiz, uint32_t nQueries, const uint8_t* visible,
                              const uint8_t* needsClipping) {
  if (!path || !depth || !hiz || width == 0 || height == 0 || width % 8 || height % 8 ||
      (nQueries && (!visible || !needsClipping)))
    return RZ_ERR_INVALID_ARGUMENT;
  const size_t nb = size_t(width / 8) * (height / 8);
  Writer w;
  w.u32(width);
  w.u32(height);
  w.u32(nQueries);
  w.raw(depth, nb * 64 * 2);
  w.raw(hiz, (nb + 8) * 2);
  w.raw(visible, nQueries);
  w.raw(needsClipping, nQueries);
  return write_file(path, kMagicFrame, w.b);
}

extern "C" int rz_read_frame(const char* path, uint32_t* width, uint32_t* height, uint32_t* nQueries, uint16_t* depth,
                             uint16_t* hiz, uint8_t* visible, uint8_t* needsClipping) {
  if (!path || !width || !height || !nQueries) return RZ_ERR_INVALID_ARGUMENT;
  std::vector<uint8_t> payload;
  int s = read_file(path, kMagicFrame, payload);
  if (s != RZ_OK) return s;
  Reader r{payload};
  const uint32_t w = r.u32(), h = r.u32(), nq = r.u32();
  if (!r.ok || w == 0 || h == 0 || w % 8 || h % 8) return RZ_ERR_IO;
  // The buffer sizes rz_create accepts; larger headers are damaged files, and
  // bounding them first keeps the size arithmetic below far from wrapping.
  if (w > 65528 || h > 65528 || uint64_t(w / 8) * (h / 8) >= (1u << 24)) return RZ_ERR_IO;
  const size_t nb = size_t(w / 8) * (h / 8);
  if (payload.size() != 12 + nb * 128 + (nb + 8) * 2 + size_t(nq) * 2) return RZ_ERR_IO;
  *width = w;
  *height = h;
  *nQueries = nq;
  if (depth) std::memcpy(depth, payload.data() + 12, nb * 128);
  if (hiz) std::memcpy(hiz, payload.data() + 12 + nb * 128, (nb + 8) * 2);
  const uint8_t* q = payload.data() + 12 + nb * 128 + (nb + 8) * 2;
  if (visible) std::memcpy(visible, q, nq);
  if (needsClipping) std::memcpy(needsClipping, q + nq, nq);
  return RZ_OK;
}
