// Types shared by the FFT pass kernels and their host launchers.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace cdn {

// How the x-forward pass obtains its real input Y (one RL-BA iteration's
// first kernel, src/RL-Biggs-Andrews.cpp:296-321).
enum XInMode : int {
  XIN_PLAIN = 0,  // Y = X (read X)
  XIN_PRED = 1,   // Y = max(X + lambda (X - Xp), 0); Y is also stored
  XIN_SEED = 2,   // Y read from the Y buffer (warm-start seed)
};

struct RaOtf {
  const float2* tz;   // z-interpolated radial table, (nr, nz) kz fastest
  const float4* tz2;  // rows r and r + 1 side by side, (nr - 1, nz): one load per lookup
  int nr;             // radial samples (nxotf); valid iff kr < nr - 1
};

struct Otf3d {
  const float2* raw;  // (nzotf, nyotf, nxotf) complex, x fastest
  int nxotf, nyotf, nzotf;
};

struct OtfParams {
  int ra;  // 1: rotationally averaged (RaOtf), 0: 3-D table (Otf3d)
  RaOtf ra_tab;
  Otf3d tab3;
  float kxscale, kyscale, kzscale;
  float inv_n;  // 1 / (nx ny nz): the filter's normalisation
};

// One stack's share of a column pass that spans stacks of different shapes
// (same transform length): its spectrum, its column geometry, and the first
// work item (tile) of the launch that belongs to it.  Items of stack s are
// item0 .. item0 + n_outer ntiles - 1, slab-major.
struct ColSeg {
  float2* S;
  long ncols, line_stride, outer_stride, ntiles, item0;
};
// The same for the z filter (slabs are ky rows; the stack's own OTF).
struct ZSeg {
  float2* S;
  long ncols, line_stride, outer_stride, ntiles, ny, item0;
  OtfParams otf;
};

}  // namespace cdn
