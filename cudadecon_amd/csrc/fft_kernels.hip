// Host launchers of the FFT pass kernels: precompiled lengths dispatch to
// template instantiations; any other 2,3,5,7-smooth length runs the same
// kernels compiled at run time (jit.cpp).
#include "common.hpp"
#include "fft_kernels_dev.hpp"
#include "fft_kernels.hpp"
#include "jit.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace cdn {

// ---------------------------------------------------------------------------
// host side: length dispatch
// ---------------------------------------------------------------------------
#define CDN_FFT_LENGTHS(X) \
  X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(810) X(1024) X(2048)

bool fft_length_precompiled(int n) {
  switch (n) {
#define CDN_CASE(N) \
  case N:           \
    return true;
    CDN_FFT_LENGTHS(CDN_CASE)
#undef CDN_CASE
    default:
      return false;
  }
}

bool fft_length_supported(int n) { return fft_length_precompiled(n) || jit_length_ok(n); }

void fft_prepare_length(int n) {
  if (!fft_length_precompiled(n)) jit_prepare(n);
}

static Plan host_plan(int n) {
  const Plan P = make_plan(n);
  if (!P.ok || n < 1) fail("FFT length " + std::to_string(n) + " is not 2,3,5,7-smooth");
  return P;
}

int fft_twiddle_count(int n) {
  const Plan P = host_plan(n);
  return P.twsize > 0 ? P.twsize : 1;
}

void fft_twiddles(int n, float2* out) {
  const Plan P = host_plan(n);
  out[0] = make_float2(1.f, 0.f);
  for (int s = 0; s < P.ns; ++s) {
    const int R = P.r[s], lp = P.lp[s];
    if (lp == 1) continue;
    for (int k = 0; k < lp; ++k)
      for (int t = 1; t < R; ++t) {
        // exp(-2 pi i t k / (lp R)) in long double, reduced by integer modulus
        const long M = (long)lp * R;
        const long m = ((long)t * k) % M;
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)M;
        out[P.twoff[s] + k * (R - 1) + (t - 1)] = make_float2((float)cosl(a), (float)sinl(a));
      }
  }
}

static inline unsigned grid1(long n) {
  if (n <= 0) return 0;
  if (n > 0x7fffffffL) fail("grid too large");
  return (unsigned)n;
}

// ---------------------------------------------------------------------------
// Tuned launch shapes per length (set by the autotuner).  A length with a
// variant always runs that variant's run-time compiled kernels.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_vmu;
std::map<int, FftVariant> g_variants;
thread_local int t_builtin_only = 0;  // > 0: no run-time compiled variant
}  // namespace

BuiltinShapesOnly::BuiltinShapesOnly() { ++t_builtin_only; }
BuiltinShapesOnly::~BuiltinShapesOnly() { --t_builtin_only; }

void fft_set_variant(int n, const FftVariant& v) {
  std::lock_guard<std::mutex> lk(g_vmu);
  if (v.is_default())
    g_variants.erase(n);
  else
    g_variants[n] = v;
}
bool fft_get_variant(int n, FftVariant* v) {
  if (t_builtin_only > 0) return false;
  std::lock_guard<std::mutex> lk(g_vmu);
  auto it = g_variants.find(n);
  if (it == g_variants.end()) return false;
  if (v) *v = it->second;
  return true;
}
void fft_clear_variants() {
  std::lock_guard<std::mutex> lk(g_vmu);
  g_variants.clear();
}

#define CDN_LAUNCH(kernel, N, grid, block, ...) \
  hipLaunchKernelGGL(kernel<N>, dim3(grid), dim3(block), 0, st, __VA_ARGS__)

void launch_x_r2c(int nx, hipStream_t st, const float* X, const float* Xp, float* Y, const double* lam, int mode,
                  float2* S, long rows, long pitch, const float2* tw, long stack_rows, const long* row_starts,
                  int nstk) {
  FftVariant tv;
  const bool tuned = fft_get_variant(nx, &tv);
  const XShape xs = x_shape(nx, tuned ? tv.xthreads : -1, XK_R2C);
  const unsigned grid = grid1(((rows + 1) / 2 + xs.NL - 1) / xs.NL);
  if (stack_rows <= 0) stack_rows = rows;
  if (stack_rows != rows && stack_rows % 2) fail("launch_x_r2c: rows per stack must be even");
  const bool ragged = row_starts != nullptr;
  if (ragged && nstk < 1) fail("launch_x_r2c: no stacks");
  const JitKernel jk = ragged ? JK_X_R2C_RAGGED : JK_X_R2C;
  void* args[] = {&X, &Xp, &Y, &lam, &mode, &S, &rows, &pitch, &stack_rows, &row_starts, &nstk, &tw};
  if (tuned) {
    jit_launch(nx, jk, grid, xs.THREADS, st, args, tv);
  } else {
    switch (nx) {
#define CDN_CASE(N)                                                                                            \
  case N:                                                                                                      \
    if (ragged)                                                                                                \
      hipLaunchKernelGGL((k_x_r2c<N, true>), dim3(grid), dim3(xs.THREADS), 0, st, X, Xp, Y, lam, mode, S, rows, \
                         pitch, stack_rows, row_starts, nstk, tw);                                             \
    else                                                                                                       \
      hipLaunchKernelGGL((k_x_r2c<N, false>), dim3(grid), dim3(xs.THREADS), 0, st, X, Xp, Y, lam, mode, S,      \
                         rows, pitch, stack_rows, row_starts, nstk, tw);                                       \
    break;
      CDN_FFT_LENGTHS(CDN_CASE)
#undef CDN_CASE
      default:
        jit_launch(nx, jk, grid, xs.THREADS, st, args);
    }
  }
  CDN_HIP(hipGetLastError());
}

void launch_x_c2r(int nx, hipStream_t st, const float2* S, float* out, long rows, long pitch, const float2* tw) {
  FftVariant tv;
  const bool tuned = fft_get_variant(nx, &tv);
  const XShape xs = x_shape(nx, tuned ? tv.xthreads : -1);
  const unsigned grid = grid1(((rows + 1) / 2 + xs.NL - 1) / xs.NL);
  void* args[] = {&S, &out, &rows, &pitch, &tw};
  if (tuned) {
    jit_launch(nx, JK_X_C2R, grid, xs.THREADS, st, args, tv);
  } else {
    switch (nx) {
#define CDN_CASE(N)                                                      \
  case N:                                                                \
    CDN_LAUNCH(k_x_c2r, N, grid, xs.THREADS, S, out, rows, pitch, tw); \
    break;
      CDN_FFT_LENGTHS(CDN_CASE)
#undef CDN_CASE
      default:
        jit_launch(nx, JK_X_C2R, grid, xs.THREADS, st, args);
    }
  }
  CDN_HIP(hipGetLastError());
}

void launch_x_lrcore(int nx, hipStream_t st, float2* S, const float* raw, long rows, long pitch, const float2* tw,
                     float eps) {
  FftVariant tv;
  const bool tuned = fft_get_variant(nx, &tv);
  const XShape xs = x_shape(nx, tuned ? tv.xthreads : -1, XK_LR);
  const unsigned grid = grid1(((rows + 1) / 2 + xs.NL - 1) / xs.NL);
  void* args[] = {&S, &raw, &rows, &pitch, &tw, &eps};
  if (tuned) {
    jit_launch(nx, JK_X_LRCORE, grid, xs.THREADS, st, args, tv);
  } else {
    switch (nx) {
#define CDN_CASE(N)                                                              \
  case N:                                                                        \
    CDN_LAUNCH(k_x_lrcore, N, grid, xs.THREADS, S, raw, rows, pitch, tw, eps); \
    break;
      CDN_FFT_LENGTHS(CDN_CASE)
#undef CDN_CASE
      default:
        jit_launch(nx, JK_X_LRCORE, grid, xs.THREADS, st, args);
    }
  }
  CDN_HIP(hipGetLastError());
}

int x_update_blocks(int nx, long rows) {
  FftVariant tv;
  const bool tuned = fft_get_variant(nx, &tv);
  const XShape xs = x_shape(nx, tuned ? tv.xthreads : -1);
  return (int)(((rows + 1) / 2 + xs.NL - 1) / xs.NL);
}

int x_update_blocks_max(long rows) { return (int)((rows + 1) / 2); }

long x_update_rows_per_block(int nx) {
  FftVariant tv;
  const bool tuned = fft_get_variant(nx, &tv);
  return 2L * x_shape(nx, tuned ? tv.xthreads : -1).NL;
}

int launch_x_update(int nx, hipStream_t st, const float2* S, const float* Y, float* Xout, float* G, double* part,
                    int dots, long rows, long pitch, const float2* tw) {
  FftVariant tv;
  const bool tuned = fft_get_variant(nx, &tv);
  const XShape xs = x_shape(nx, tuned ? tv.xthreads : -1);
  const int nb = (int)(((rows + 1) / 2 + xs.NL - 1) / xs.NL);
  void* args[] = {&S, &Y, &Xout, &G, &part, &dots, &rows, &pitch, &tw};
  if (tuned) {
    jit_launch(nx, JK_X_UPDATE, grid1(nb), xs.THREADS, st, args, tv);
  } else {
    switch (nx) {
#define CDN_CASE(N)                                                                                        \
  case N:                                                                                                  \
    CDN_LAUNCH(k_x_update, N, grid1(nb), xs.THREADS, S, Y, Xout, G, part, dots, rows, pitch, tw);        \
    break;
      CDN_FFT_LENGTHS(CDN_CASE)
#undef CDN_CASE
      default:
        jit_launch(nx, JK_X_UPDATE, grid1(nb), xs.THREADS, st, args);
    }
  }
  CDN_HIP(hipGetLastError());
  return nb;
}

// Column tiles: when a slab's columns need two built-in tiles, the second
// one a quarter or less full (65 columns of a 128 x 128 plane in tiles of
// 64), one tile holding all of them instead, if it fits (threads <= 1024,
// LDS within the built-in budget class): a run-time compiled variant,
// bitwise the same results.  (Measured: 8% on stacks of 64 x 128 x 128; a
// wider single tile, 129 columns for 3 built-in tiles, was slower.)
bool fft_balanced_cols(int len, long ncols, FftVariant* v) {
  if (t_builtin_only > 0) return false;
  const CShape d = c_shape(len);
  if (ncols <= d.L || ncols >= 2 * d.L || 4 * (ncols - d.L) > d.L) return false;
  const long budget = len >= 768 ? 160L * 1024 : 80L * 1024;
  if (ncols * c_shape(len, 1).TPT > 1024 || (long)len * ncols * 8 > budget) return false;
  if (v) {
    *v = FftVariant();
    v->cols = (int)ncols;
  }
  return true;
}

// tuned variant of the length, else one tile of all columns, else built in
static bool col_variant(int len, long ncols, FftVariant* v) {
  if (fft_get_variant(len, v)) return true;
  return fft_balanced_cols(len, ncols, v);
}

void fft_prepare_cols(int len, long ncols) {
  FftVariant v;
  if (fft_get_variant(len, &v) || !fft_balanced_cols(len, ncols, &v)) return;
  // compile and load (grid 0: no launch) before any stream capture
  for (JitKernel k : {JK_COL_F, JK_COL_I, JK_Z_FILTER, JK_Z_FILTER_RA}) jit_launch(len, k, 0, 1, nullptr, nullptr, v);
}

void launch_col_fft(int len, int inverse, hipStream_t st, float2* S, long ncols, long line_stride,
                    long outer_stride, long n_outer, const float2* tw) {
  FftVariant tv;
  const bool tuned = col_variant(len, ncols, &tv);
  const CShape cs = c_shape(len, tuned ? tv.cols : -1, tuned ? tv.col_tpt : -1);
  long ntiles = (ncols + cs.L - 1) / cs.L;
  const unsigned grid = grid1(ntiles * n_outer);
  void* args[] = {&S, &ncols, &line_stride, &outer_stride, &ntiles, &tw};
  if (tuned) {
    jit_launch(len, inverse ? JK_COL_I : JK_COL_F, grid, cs.THREADS, st, args, tv);
  } else {
    switch (len) {
#define CDN_CASE(N)                                                                                            \
  case N:                                                                                                      \
    if (inverse)                                                                                               \
      hipLaunchKernelGGL((k_col<N, true>), dim3(grid), dim3(cs.THREADS), 0, st, S, ncols, line_stride,         \
                         outer_stride, ntiles, tw);                                                            \
    else                                                                                                       \
      hipLaunchKernelGGL((k_col<N, false>), dim3(grid), dim3(cs.THREADS), 0, st, S, ncols, line_stride,        \
                         outer_stride, ntiles, tw);                                                            \
    break;
      CDN_FFT_LENGTHS(CDN_CASE)
#undef CDN_CASE
      default:
        jit_launch(len, inverse ? JK_COL_I : JK_COL_F, grid, cs.THREADS, st, args);
    }
  }
  CDN_HIP(hipGetLastError());
}

void launch_z_filter(int nz, hipStream_t st, float2* S, long ncols, long line_stride, long outer_stride, long ny,
                     int conj, const OtfParams& otf, const float2* tw, int nstacks, long stack_stride) {
  FftVariant tv;
  const bool tuned = col_variant(nz, ncols, &tv);
  const CShape cs = c_shape(nz, tuned ? tv.cols : -1, tuned ? tv.col_tpt : -1);
  long ntiles = (ncols + cs.L - 1) / cs.L;
  if (nstacks < 1) fail("launch_z_filter: no stacks");
  const unsigned grid = grid1(ntiles * ny * nstacks);
  OtfParams op = otf;
  void* args[] = {&S, &ncols, &line_stride, &outer_stride, &ntiles, &ny, &stack_stride, &conj, &op, &tw};
  // the averaged-table kind for a rotationally averaged OTF
  const JitKernel jk = op.ra ? JK_Z_FILTER_RA : JK_Z_FILTER;
  if (tuned) {
    jit_launch(nz, jk, grid, cs.THREADS, st, args, tv);
  } else {
    switch (nz) {
#define CDN_CASE(N)                                                                                          \
  case N:                                                                                                    \
    if (op.ra)                                                                                               \
      hipLaunchKernelGGL((k_z_filter<N, ZF_RA>), dim3(grid), dim3(cs.THREADS), 0, st, S, ncols, line_stride, \
                         outer_stride, ntiles, ny, stack_stride, conj, op, tw);                              \
    else                                                                                                     \
      hipLaunchKernelGGL((k_z_filter<N, ZF_ANY>), dim3(grid), dim3(cs.THREADS), 0, st, S, ncols,             \
                         line_stride, outer_stride, ntiles, ny, stack_stride, conj, op, tw);                 \
    break;
      CDN_FFT_LENGTHS(CDN_CASE)
#undef CDN_CASE
      default:
        jit_launch(nz, jk, grid, cs.THREADS, st, args);
    }
  }
  CDN_HIP(hipGetLastError());
}

long col_tile_cols(int len) { return c_shape(len).L; }

void fft_prepare_ragged(int len) {
  if (fft_length_precompiled(len)) return;
  for (JitKernel k : {JK_X_R2C_RAGGED, JK_COL_RAGGED_F, JK_COL_RAGGED_I, JK_Z_RAGGED, JK_COL_RAGGED_FP,
                      JK_COL_RAGGED_IP, JK_Z_RAGGED_P, JK_Z_RAGGED_RA, JK_Z_RAGGED_P_RA, JK_X_C2R, JK_X_LRCORE,
                      JK_X_UPDATE})
    jit_launch(len, k, 0, 1, nullptr, nullptr);
}

// grid of a launch over `total` items: one workgroup per item, or at most
// `cap` workgroups (a multiple of 8) that walk the items
static unsigned ragged_grid(long total, long cap) {
  if (cap > 0 && cap < total) return grid1(std::max(8L, cap / 8 * 8));
  return grid1(total);
}

// the kernel instance for (inverse, persistent)
template <int N>
static void col_ragged_n(bool inv, bool pers, unsigned grid, unsigned threads, hipStream_t st, const ColSeg* segs,
                         int nseg, long total, const float2* tw) {
  if (inv && pers)
    hipLaunchKernelGGL((k_col_ragged<N, true, true>), dim3(grid), dim3(threads), 0, st, segs, nseg, total, tw);
  else if (inv)
    hipLaunchKernelGGL((k_col_ragged<N, true, false>), dim3(grid), dim3(threads), 0, st, segs, nseg, total, tw);
  else if (pers)
    hipLaunchKernelGGL((k_col_ragged<N, false, true>), dim3(grid), dim3(threads), 0, st, segs, nseg, total, tw);
  else
    hipLaunchKernelGGL((k_col_ragged<N, false, false>), dim3(grid), dim3(threads), 0, st, segs, nseg, total, tw);
}

void launch_col_ragged(int len, int inverse, hipStream_t st, const ColSeg* segs, int nseg, long total,
                       const float2* tw, long grid_cap) {
  if (total <= 0) return;
  const CShape cs = c_shape(len);
  const unsigned grid = ragged_grid(total, grid_cap);
  const bool pers = grid < total;
  void* args[] = {&segs, &nseg, &total, &tw};
  switch (len) {
#define CDN_CASE(N)                                                                      \
  case N:                                                                                \
    col_ragged_n<N>(inverse != 0, pers, grid, cs.THREADS, st, segs, nseg, total, tw); \
    break;
    CDN_FFT_LENGTHS(CDN_CASE)
#undef CDN_CASE
    default:
      jit_launch(len, pers ? (inverse ? JK_COL_RAGGED_IP : JK_COL_RAGGED_FP) : (inverse ? JK_COL_RAGGED_I : JK_COL_RAGGED_F),
                 grid, cs.THREADS, st, args);
  }
  CDN_HIP(hipGetLastError());
}

// the kernel instance for (persistent, kind)
template <int N>
static void z_ragged_n(bool pers, bool ra, unsigned grid, unsigned threads, hipStream_t st, const ZSeg* segs,
                       int nseg, long total, int conj, const float2* tw) {
  if (pers && ra)
    hipLaunchKernelGGL((k_z_ragged<N, true, ZF_RA>), dim3(grid), dim3(threads), 0, st, segs, nseg, total, conj, tw);
  else if (pers)
    hipLaunchKernelGGL((k_z_ragged<N, true, ZF_ANY>), dim3(grid), dim3(threads), 0, st, segs, nseg, total, conj, tw);
  else if (ra)
    hipLaunchKernelGGL((k_z_ragged<N, false, ZF_RA>), dim3(grid), dim3(threads), 0, st, segs, nseg, total, conj, tw);
  else
    hipLaunchKernelGGL((k_z_ragged<N, false, ZF_ANY>), dim3(grid), dim3(threads), 0, st, segs, nseg, total, conj,
                       tw);
}

void launch_z_ragged(int nz, hipStream_t st, const ZSeg* segs, int nseg, long total, int conj, const float2* tw,
                     long grid_cap, bool ra) {
  if (total <= 0) return;
  const CShape cs = c_shape(nz);
  const unsigned grid = ragged_grid(total, grid_cap);
  const bool pers = grid < total;
  void* args[] = {&segs, &nseg, &total, &conj, &tw};
  switch (nz) {
#define CDN_CASE(N)                                                                    \
  case N:                                                                              \
    z_ragged_n<N>(pers, ra, grid, cs.THREADS, st, segs, nseg, total, conj, tw);     \
    break;
    CDN_FFT_LENGTHS(CDN_CASE)
#undef CDN_CASE
    default:
      jit_launch(nz, pers ? (ra ? JK_Z_RAGGED_P_RA : JK_Z_RAGGED_P) : (ra ? JK_Z_RAGGED_RA : JK_Z_RAGGED), grid,
                 cs.THREADS, st, args);
  }
  CDN_HIP(hipGetLastError());
}

}  // namespace cdn
