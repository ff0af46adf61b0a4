// RL-BA run that allocates nothing (noalloc.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace cdn {

// Size of the static device scratch array every GPU the library runs on holds.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;

// n_iters RL-BA iterations of the stack d_raw (device, preprocessed, x
// fastest) into d_out (device), with all other device memory taken from the
// static scratch array; built-in FFT lengths only.  Throws on bad input or a
// stack too large for the array.  Synchronises `st` before returning.
void noalloc_run(const float* d_raw, float* d_out, int nx, int ny, int nz, const float* otf, int otf_width,
                 int otf_height, int otf_depth, const float kscale[3], int n_iters, hipStream_t st);

}  // namespace cdn
