// Definitions shared by the library's C ABI translation units (not installed).
#pragma once

#include <stdint.h>

#include "rasterizer_amd.h"

// A baked occluder's device copy: Occluder::m_vertexData (Occluder.h:17-18)
// plus its reference box and bounds.
struct rz_occluder {
  rz_context* ctx;
  uint32_t* d_packets;
  uint32_t packetCount;
  float refMin[4], refMax[4], boundsMin[4], boundsMax[4];
};

namespace rzi {

// The coverage table and this host's RCPPS table, built once per process.
struct HostTables {
  int status = RZ_OK;
  uint64_t cov[4096];
  uint32_t rcp[4096];
};
const HostTables& host_tables();

}  // namespace rzi
