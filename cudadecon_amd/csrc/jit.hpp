// Run-time compilation of the FFT pass kernels for 2,3,5,7-smooth lengths
// outside the precompiled set (hipRTC, gfx950).  The kernels are the same
// templates as the precompiled ones (fft_kernels_dev.hpp).
#pragma once

#include <hip/hip_runtime.h>

namespace cdn {

enum JitKernel { JK_X_R2C = 0, JK_X_C2R, JK_X_LRCORE, JK_X_UPDATE, JK_COL_F, JK_COL_I, JK_Z_FILTER, JK_COUNT };

// n is 2,3,5,7-smooth and within the lengths the kernels are sized for
bool jit_length_ok(int n);
// compile the kernels of length n (host only, cached); throws on failure
void jit_prepare(int n);
// launch kernel k of length n on the current device (loads its module once)
void jit_launch(int n, JitKernel k, unsigned grid, unsigned block, hipStream_t st, void** args);

}  // namespace cdn
