// Launch helpers shared by the library's HIP translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <set>
#include <utility>

namespace sbi {

// Let `fn` use up to `bytes` of dynamic LDS (gfx950: 160 KB per CU).  The attribute belongs to the
// kernel on the current device, so it is set once per (kernel, device), under a lock.
inline void allow_lds(const void* fn, int bytes = 160 * 1024) {
  static std::mutex mu;
  static std::set<std::pair<const void*, int>> done;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> lock(mu);
  if (done.insert({fn, dev}).second)
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace sbi
