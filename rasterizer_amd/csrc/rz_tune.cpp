// Persisted autotuning cache: one text line per tuned shape,
// "<key>\t<value>", where the key names the device (architecture and CU
// count) and the workload shape, and the value the chosen launch
// configuration.  Files are replaced whole (write to a temporary, rename), so
// a reader never sees a half-written file.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rz_internal.h"

namespace rzi {

std::string device_key(int device) {
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
  std::string arch = prop.gcnArchName;
  const size_t colon = arch.find(':');
  if (colon != std::string::npos) arch.resize(colon);  // drop feature flags (":sramecc+:xnack-")
  std::ostringstream k;
  k << arch << ",cu" << prop.multiProcessorCount;
  return k.str();
}

bool tune_lookup(const char* path, const std::string& key, std::string& value) {
  if (!path || !*path) return false;
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  bool found = false;
  while (std::getline(in, line)) {
    const size_t tab = line.find('\t');
    if (tab == std::string::npos || line.compare(0, tab, key) != 0 || tab != key.size()) continue;
    value = line.substr(tab + 1);
    found = true;  // the last entry for a key wins
  }
  return found;
}

bool tune_store(const char* path, const std::string& key, const std::string& value) {
  if (!path || !*path) return false;
  std::vector<std::string> keep;
  {
    std::ifstream in(path);
    std::string line;
    while (in && std::getline(in, line)) {
      const size_t tab = line.find('\t');
      if (tab == std::string::npos) continue;
      if (tab == key.size() && line.compare(0, tab, key) == 0) continue;
      keep.push_back(line);
    }
  }
  keep.push_back(key + "\t" + value);
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream out(tmp, std::ios::trunc);
    if (!out) return false;
    for (const std::string& l : keep) out << l << '\n';
    if (!out.good()) return false;
  }
  return std::rename(tmp.c_str(), path) == 0;
}

}  // namespace rzi
