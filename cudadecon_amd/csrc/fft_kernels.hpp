// Host-side launch interface of the 3-D FFT pass kernels (fft_kernels.hip).
//
// Volume layout: real (nz, ny, nx), x fastest, no padding.  Half spectrum:
// (nz, ny, nx/2+1) complex64 with row pitch `pitch` >= nx/2+1 (padded to a
// multiple of 16 so column tiles start on 128-byte boundaries).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_types.hpp"
#include "jit.hpp"

namespace cdn {

// Length is compiled in, or can be compiled at run time (2,3,5,7-smooth).
bool fft_length_supported(int n);
// Length is among the lengths built into the library.
bool fft_length_precompiled(int n);
// Compile the kernels of a length that is not precompiled (no-op otherwise);
// throws with the compiler log on failure.  Needs no GPU.
void fft_prepare_length(int n);
// Twiddle table for a supported length (host, FP64-derived float2 values).
int fft_twiddle_count(int n);
void fft_twiddles(int n, float2* out);

// R2C along x of row pairs with the prediction fused in; writes S.  Rows
// come in stacks of stack_rows (even; 0: one stack), stack s using the
// factor lam[4 s]; or, with row_starts (device, nstk ascending even rows,
// the first 0), in stacks of different heights.
void launch_x_r2c(int nx, hipStream_t st, const float* X, const float* Xp, float* Y, const double* lam,
                  int mode, float2* S, long rows, long pitch, const float2* tw, long stack_rows = 0,
                  const long* row_starts = nullptr, int nstk = 0);
// plain C2R along x (ignores imag of kx = 0 and kx = nx/2).
void launch_x_c2r(int nx, hipStream_t st, const float2* S, float* out, long rows, long pitch,
                  const float2* tw);
// C2R along x -> LRcore with raw -> R2C along x, in place on S.
void launch_x_lrcore(int nx, hipStream_t st, float2* S, const float* raw, long rows, long pitch,
                     const float2* tw, float eps);
// C2R along x -> X = max(C Y, 0), G_new = X - Y written over G; if dots:
// per-block partial sums of G_new.G_old and G_old.G_old (FP32 products,
// FP64 sums) into part[2*block..].  Returns the number of blocks launched.
int launch_x_update(int nx, hipStream_t st, const float2* S, const float* Y, float* Xout, float* G,
                    double* part, int dots, long rows, long pitch, const float2* tw);
int x_update_blocks(int nx, long rows);
// rows per x-update workgroup of the launch shape in use for nx
long x_update_rows_per_block(int nx);
// upper bound of x_update_blocks over every launch shape (one row pair per block)
int x_update_blocks_max(long rows);

// Tuned launch shapes: a length with a variant runs that variant's run-time
// compiled kernels (results are bitwise those of the default shapes).
void fft_set_variant(int n, const FftVariant& v);
// One column tile of all `ncols` columns of length `len` when they would
// need two built-in tiles, the second a quarter full or less, and the tile
// fits (v: the variant); the column launchers use it unless the length has
// a tuned variant.
bool fft_balanced_cols(int len, long ncols, FftVariant* v);
// compile and load the kernels those launches will need (no GPU work)
void fft_prepare_cols(int len, long ncols);
bool fft_get_variant(int n, FftVariant* v);
// While one exists on this thread, every launch uses the built-in kernels of
// its length: no tuned or single-tile variant, whose run-time compiled code
// would have to be loaded (column results are bitwise the same).
struct BuiltinShapesOnly {
  BuiltinShapesOnly();
  ~BuiltinShapesOnly();
};
void fft_clear_variants();

// Column FFT (forward or inverse) over `len` points at element stride
// line_stride, for columns [0, ncols) of each of n_outer slabs.
void launch_col_fft(int len, int inverse, hipStream_t st, float2* S, long ncols, long line_stride,
                    long outer_stride, long n_outer, const float2* tw);
// z forward -> x H (or conj H) / N -> z inverse; slab index = ky, for
// nstacks spectra stack_stride complex elements apart.
void launch_z_filter(int nz, hipStream_t st, float2* S, long ncols, long line_stride, long outer_stride,
                     long ny, int conj, const OtfParams& otf, const float2* tw, int nstacks = 1,
                     long stack_stride = 0);

// Column passes spanning stacks of different shapes: one launch for every
// stack whose transform length is `len`, segs (device) one per stack with
// items numbered by tiles of col_tile_cols(len) columns, `total` items in
// all.  grid_cap > 0 bounds the workgroups, which then walk the items.
long col_tile_cols(int len);
// compile and load the run-time kernels the ragged launches of a length not
// built in will use (no GPU work; call before any stream capture)
void fft_prepare_ragged(int len);
void launch_col_ragged(int len, int inverse, hipStream_t st, const ColSeg* segs, int nseg, long total,
                       const float2* tw, long grid_cap = 0);
// ra: every segment's OTF is rotationally averaged (the kernel kind the
// one-stack z filter uses for such a table)
void launch_z_ragged(int nz, hipStream_t st, const ZSeg* segs, int nseg, long total, int conj, const float2* tw,
                     long grid_cap = 0, bool ra = false);

}  // namespace cdn
