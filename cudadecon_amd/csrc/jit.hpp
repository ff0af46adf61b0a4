// Run-time compilation of the FFT pass kernels for 2,3,5,7-smooth lengths
// outside the precompiled set (hipRTC, gfx950).  The kernels are the same
// templates as the precompiled ones (fft_kernels_dev.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace cdn {

enum JitKernel {
  JK_X_R2C = 0,
  JK_X_C2R,
  JK_X_LRCORE,
  JK_X_UPDATE,
  JK_COL_F,
  JK_COL_I,
  JK_Z_FILTER,
  JK_X_R2C_RAGGED,
  JK_COL_RAGGED_F,
  JK_COL_RAGGED_I,
  JK_Z_RAGGED,
  JK_COL_RAGGED_FP,
  JK_COL_RAGGED_IP,
  JK_Z_RAGGED_P,
  JK_Z_FILTER_RA,
  JK_Z_RAGGED_RA,
  JK_Z_RAGGED_P_RA,
  JK_COUNT
};

// Launch-shape choices for the kernels of one length (0 / -1: the built-in
// rule): columns per y/z tile, threads per x-pass workgroup, non-temporal
// loads in the column FFT, the column passes' workgroup-to-XCD renumbering
// (0 none, 1 tiles narrower than a line, 2 every tile), threads per column
// line, the z filter's minimum waves per SIMD (a register budget), OTF table
// rows the z filter stages in LDS per tile (-1: none).  Picked by the
// autotuner (tune.cpp).
struct FftVariant {
  int cols = 0;
  int xthreads = 0;
  int col_nt_load = -1;
  int xcd_remap = -1;
  int col_tpt = 0;
  int zf_waves = 0;
  int zf_rows = 0;
  bool is_default() const {
    return cols == 0 && xthreads == 0 && col_nt_load < 0 && xcd_remap < 0 && col_tpt == 0 && zf_waves == 0 &&
           zf_rows == 0;
  }
  // -D options that build this variant for length n
  std::string defines(int n) const;
  bool operator==(const FftVariant& o) const {
    return cols == o.cols && xthreads == o.xthreads && col_nt_load == o.col_nt_load && xcd_remap == o.xcd_remap &&
           col_tpt == o.col_tpt && zf_waves == o.zf_waves && zf_rows == o.zf_rows;
  }
};

// cache directory ($XDG_CACHE_HOME or $HOME/.cache, /cudadecon_amd); "" if none
std::string jit_cache_dir();

// n is 2,3,5,7-smooth and within the lengths the kernels are sized for
bool jit_length_ok(int n);
// compile the kernels of length n (host only, cached); throws on failure
void jit_prepare(int n, const FftVariant& v = FftVariant());
// launch kernel k of length n on the current device (loads its module once)
void jit_launch(int n, JitKernel k, unsigned grid, unsigned block, hipStream_t st, void** args,
                const FftVariant& v = FftVariant());

}  // namespace cdn
