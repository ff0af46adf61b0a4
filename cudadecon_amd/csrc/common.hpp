// Shared host/device helpers for libcudaDecon (MI355X / gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include "common_dev.hpp"
#include "host_common.hpp"

#include <cstdint>
#include <stdexcept>
#include <string>

namespace cdn {

// Host-side HIP error check: throws; the C ABI layer turns exceptions into a
// 0 return code plus a last-error string (never exit() across the ABI).  A
// failed call also leaves the runtime's per-thread last error set (an
// out-of-memory hipMalloc, say), which the next launch check would report as
// its own; it is cleared before the throw.
#define CDN_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      (void)hipGetLastError();                                                          \
      throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_) +     \
                               " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    }                                                                                   \
  } while (0)

}  // namespace cdn
