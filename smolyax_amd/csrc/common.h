// Shared host-side helpers: thread-local error message, HIP status checks.
#pragma once

#include <string>

namespace sbi {

void set_error(const std::string& msg);

}  // namespace sbi

// Return SBI_ERR_HIP with the HIP error text recorded when a HIP runtime call fails (used by the
// HIP translation units only).
#define SBI_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      sbi::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                 \
      return SBI_ERR_HIP;                                                                \
    }                                                                                    \
  } while (0)
