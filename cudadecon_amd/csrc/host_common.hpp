// Host-only helpers shared by the library, the command-line caller and the
// CPU-sanitizer build of the host code (no HIP headers here).
#pragma once

#include <stdexcept>
#include <string>

namespace cdn {

// Every failure is an exception; the C ABI layer turns it into a 0 return
// code plus a last-error string (never exit() across the ABI).
inline void fail(const std::string& msg) { throw std::runtime_error(msg); }

}  // namespace cdn
