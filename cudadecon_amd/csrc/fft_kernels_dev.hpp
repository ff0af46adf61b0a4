// FFT pass kernels of the RL-Biggs-Andrews iteration on gfx950 (device code).
//
// One RL-BA iteration = 9 streaming passes over HBM:
//   x-R2C (+prediction)  y-fwd  z-fwd.H.z-inv  y-inv  x-C2R.LRcore.x-R2C
//   y-fwd  z-fwd.conjH.z-inv  y-inv  x-C2R.update.dots
// replacing the reference's cufftExecR2C/C2R calls (src/RLgpuImpl.cu:183,218)
// and the elementwise kernels at src/RLgpuImpl.cu:331-350,383-508,571-666.
//
// x passes: two real rows are packed as z = a + i b, transformed as one
// complex line, and split into two half spectra (and the reverse for C2R,
// where the imaginary parts of the kx = 0 and kx = nx/2 bins are ignored,
// i.e. numpy irfft semantics).
// y/z passes: a workgroup owns a tile of L adjacent kx columns of one slab,
// so global loads are L*8 bytes contiguous.
//
// Compiled twice: by hipcc for the precompiled lengths (fft_kernels.hip) and
// by hipRTC at run time for any other 2,3,5,7-smooth length (jit.cpp).
#pragma once

#include "fft_core.hpp"
#include "fft_types.hpp"

namespace cdn {

// Non-temporal policy per kernel family (measured at 256 x 1024 x 1024):
// x passes and the z filter stream both ways; plain y/z column transforms
// keep cached loads and stream their stores.
#ifndef CDN_NT_X
#define CDN_NT_X 1
#endif
#ifndef CDN_NT_Z
#define CDN_NT_Z 1
#endif
#ifndef CDN_NT_COL_LOAD
#define CDN_NT_COL_LOAD 0
#endif
#ifndef CDN_NT_COL_STORE
#define CDN_NT_COL_STORE 1
#endif
constexpr bool NTX = CDN_NT_X != 0, NTZ = CDN_NT_Z != 0;
constexpr bool NTCL = CDN_NT_COL_LOAD != 0, NTCS = CDN_NT_COL_STORE != 0;

// ---------------------------------------------------------------------------
// Per-length launch configuration
// ---------------------------------------------------------------------------
// Build-time tile overrides (pairs "length, value"): columns per y/z tile and
// threads per x-pass workgroup.  Used to sweep launch shapes.
#ifndef CDN_COLS_OVERRIDE
#define CDN_COLS_OVERRIDE 0, 0
#endif
#ifndef CDN_XTHREADS_OVERRIDE
#define CDN_XTHREADS_OVERRIDE 0, 0
#endif
constexpr int kColsOverride[] = {CDN_COLS_OVERRIDE};
constexpr int kXThreadsOverride[] = {CDN_XTHREADS_OVERRIDE};
template <int K>
constexpr int lookup_override(const int (&tab)[K], int N) {
  for (int i = 0; i + 1 < K; i += 2)
    if (tab[i] == N) return tab[i + 1];
  return 0;
}

// threads per line for the all-LDS fallback: largest divisor of N not above
// N/16 (about 16 elements per thread)
constexpr int tpt_for(int N) {
  int t = N / 16;
  if (t < 1) t = 1;
  for (int d = t; d >= 1; --d)
    if (N % d == 0) return d;
  return 1;
}

// threads per line: elements per thread E = radix lcm (doubled up to 8 when
// it is small) makes the plan exact, so lines stay in registers between the
// first and last stages; lengths with a radix lcm above 32 use the fallback
constexpr int tpt_col(int N) {
  const Plan P = make_plan(N);
  if (P.lcm <= 32) {
    int E = P.lcm;
    while (E < 8 && N % (E * 2) == 0) E *= 2;
    return N / E;
  }
  return tpt_for(N);
}

#ifndef CDN_XTPT_OVERRIDE
#define CDN_XTPT_OVERRIDE 0, 0
#endif
constexpr int kXTptOverride[] = {CDN_XTPT_OVERRIDE};

// x passes: lines of up to 64 threads split/merge the packed row pair with
// wavefront shuffles; longer lines (N >= 2048, measured faster with 16
// elements per thread) use an LDS exchange.  Short lines (16 threads or
// fewer by the radix rule) get the smallest divisor of N from 32 up, so
// that a wavefront's row loads and stores cover whole 128-byte lines
// instead of 32-byte pieces of several rows (x passes 2-8x faster on stacks
// 64 to 384 wide); the LRcore pass (lr), two transforms per line, measured
// best from 16 up.  CDN_XTPT_RULE 1 keeps the rule to powers of two.
//
// Rule 3 adds: lines whose plan still leaves more than 16 elements per
// thread (radix lcm 20 to 30: 768, 810) held them in up to 256 registers,
// one or two waves per SIMD; they take the smallest thread count with at
// most CDN_XEMAX elements per thread and the all-LDS transform.  The LRcore
// pass (two transforms per line) does so only above 24 elements per thread,
// to at most CDN_XEMAX_LR: at 24 (384, 768) the register-resident transform
// measured faster (x_lrcore 0.102 vs 0.120 ms at 64 x 768 x 768).  At
// 128 x 512 x 810 the x passes went from 0.25 / 0.47 / 0.39 ms (R2C / LRcore
// / update) to 0.18 / 0.26 / 0.24 (caps 6 and 15); at 64 x 768 x 768 R2C and
// update from 0.117 / 0.132 to 0.106 / 0.125.  With the all-LDS transform's
// stage pairs fused (CDN_FUSE_MAX), caps 5 and 9 (810: 162 threads per line
// for R2C and the update, 90 for LRcore, against 135 and 54) measured
// 0.172 / 0.200 / 0.223 ms against 0.181 / 0.196 / 0.234 for 6 and 9,
// 0.171 / 0.221 / 0.220 for 5 and 5, 0.172 / 0.210 / 0.222 for 5 and 15;
// 768 was the same under all of them (iteration 1.389-1.394 ms at
// 128 x 768 x 768).
#ifndef CDN_XTPT_RULE
#define CDN_XTPT_RULE 3
#endif
constexpr int tpt_x_short(int N, bool lr) {
  int t = tpt_col(N);
  const int want = lr ? 16 : 32;
  if (t >= want || t > 16 || N < 2 * want) return t;
  if (CDN_XTPT_RULE == 1 && (N & (N - 1)) != 0) return t;
  // the smallest divisor of N from `want` up (N itself is one)
  for (int d = want; d <= N / 2; ++d)
    if (N % d == 0) return d;
  return t;
}
// Kinds of x pass with their own line shape: the inverse and update passes,
// the LRcore pass (two transforms per line), the R2C pass with the prediction.
enum XKind : int { XK_PLAIN = 0, XK_LR = 1, XK_R2C = 2 };
// Threads per line of the R2C pass where they differ from the other x
// passes (pairs length, threads).  More threads per line, fewer elements
// each (all-LDS transform): 1024 points with 256 threads took the R2C pass
// from 0.742 to 0.712 ms at 256 x 1024 x 1024 (the same shape made LRcore
// slower, 0.60 -> 0.76 ms, and the update no faster, so they keep theirs);
// 512 with 128 from 0.098-0.100 to 0.094-0.096 ms at 128 x 512 x 512 (three
// alternating runs each; 64 threads: 0.098).  2048-point lines keep the 128
// of the other passes: 256 or 512 threads measured the same (5.4-6.2 ms
// either way at 512 x 2048 x 2048), and 64 threads (32 elements each) gave
// 4.77-5.16 ms against 4.85-5.17 in alternating processes, where the x
// passes there vary by 7% between processes of one build.
#ifndef CDN_R2C_TPT
#define CDN_R2C_TPT 512, 128, 1024, 256
#endif
constexpr int kR2cTpt[] = {CDN_R2C_TPT};
// Rule 3's element caps (elements per thread) for the plain / R2C passes and
// for the LRcore pass.
#ifndef CDN_XEMAX
#define CDN_XEMAX 5
#endif
#ifndef CDN_XEMAX_LR
#define CDN_XEMAX_LR 9
#endif
constexpr int tpt_x(int N, int kind = XK_PLAIN) {
  if (lookup_override(kXTptOverride, N) > 0) return lookup_override(kXTptOverride, N);
  if (kind == XK_R2C && lookup_override(kR2cTpt, N) > 0) return lookup_override(kR2cTpt, N);
  const bool lr = kind == XK_LR;
  const int t = tpt_x_short(N, lr);
  if (CDN_XTPT_RULE < 3 || N / t <= (lr ? 24 : 16)) return t;
  const int emax = lr ? CDN_XEMAX_LR : CDN_XEMAX;
  for (int d = t; d <= N && d <= 1024; ++d)
    if (N % d == 0 && N / d <= emax) return d;
  return t;
}

// columns per tile for y/z passes: fill the LDS budget, at most 64 columns
// (32 above length 64) and 1024 threads, power of two (cols > 0 forces a
// width).  Lengths 96 and 128 with 32 instead of 64 columns (smaller
// workgroups fill the register-limited wave slots: the length-128 z filter
// runs 5 waves per SIMD, which 512-thread workgroups leave at 4), ms per
// launch: z filter 0.062 -> 0.050 at 128 x 384 x 512, 0.110 -> 0.102 at
// 128 x 512 x 810, 0.043 -> 0.038-0.040 at 96 x 384 x 384; 16 stacks of
// 96 x 128 x 256 0.974 -> 0.922 ms per iteration, 64 of 64 x 128 x 128
// 1.38-1.39 -> 1.36; no length-128 y pass measured slower.
constexpr int lc_for(int N, int TPT, int cols = -1) {
  if (cols < 0) cols = lookup_override(kColsOverride, N);
  if (cols > 0) return cols;
  const int budget = (N >= 768) ? 131072 : 65536;
  int L = budget / (N * 8);
  if (L > 64) L = 64;
  if (N > 64 && L > 32) L = 32;
  while (L > 1 && L * TPT > 1024) L /= 2;
  int p = 1;
  while (p * 2 <= L) p *= 2;
  return p;
}

// Launch shapes, evaluated at compile time for the kernels and at run time
// by the host launchers (one definition for both).
struct XShape {
  int TPT, E, NL, PAD, LS, THREADS, EH;
  bool SHUFFLE;
};
// xthreads > 0 forces the workgroup size (the build-time override otherwise)
constexpr XShape x_shape(int N, int xthreads = -1, int kind = XK_PLAIN) {
  XShape s{};
  s.TPT = tpt_x(N, kind);
  s.E = N / s.TPT;
  s.SHUFFLE = plan_exact(N, s.TPT) && s.TPT <= 64 && (64 % s.TPT) == 0 && (s.E % 2) == 0;
  const int ov = xthreads >= 0 ? xthreads : lookup_override(kXThreadsOverride, N);
  // LRcore lines whose thread count neither divides nor is a multiple of 64
  // straddle wavefronts: two lines per workgroup leave fewer lanes idle
  // (810: 90 threads per line, 0.200-0.203 -> 0.187-0.190 ms at
  // 128 x 512 x 810 in three alternations; five lines 0.217-0.220)
  const bool straddle = kind == XK_LR && s.TPT % 64 != 0 && 64 % s.TPT != 0;
  const int xt = ov > 0 ? ov : straddle ? 256 : 128;
  const int nl0 = (xt / s.TPT) < 1 ? 1 : (xt / s.TPT);
  s.NL = (nl0 * N * 8 > 65536) ? ((65536 / (N * 8)) < 1 ? 1 : 65536 / (N * 8)) : nl0;
  s.PAD = s.TPT < 64 ? s.TPT : 0;  // spreads lines over LDS banks
  s.LS = N + s.PAD;
  s.THREADS = s.TPT * s.NL;
  s.EH = (N / 2 + s.TPT) / s.TPT;  // half-spectrum bins per thread (LDS split path)
  return s;
}
struct CShape {
  int TPT, L, THREADS, E;
};
#ifndef CDN_CTPT_OVERRIDE
#define CDN_CTPT_OVERRIDE 0, 0
#endif
constexpr int kCTptOverride[] = {CDN_CTPT_OVERRIDE};
// cols, tpt > 0: a run-time compiled variant's shape (what its -D overrides
// give the kernel)
constexpr CShape c_shape(int N, int cols = -1, int tpt = -1) {
  CShape s{};
  s.TPT = tpt > 0 ? tpt : lookup_override(kCTptOverride, N) > 0 ? lookup_override(kCTptOverride, N) : tpt_col(N);
  s.L = lc_for(N, s.TPT, cols);
  s.THREADS = s.TPT * s.L;
  s.E = N / s.TPT;
  return s;
}

template <int N, int K = XK_PLAIN> struct XCfg {
  static constexpr XShape S = x_shape(N, -1, K);
  static constexpr int TPT = S.TPT, E = S.E, NL = S.NL, PAD = S.PAD, LS = S.LS, THREADS = S.THREADS, EH = S.EH;
  static constexpr bool SHUFFLE = S.SHUFFLE;
};

// Column passes exchange between FFT stages in halves (real, then imaginary
// parts) when the plan is exact: half the LDS per workgroup.
#ifndef CDN_SPLIT_LDS
#define CDN_SPLIT_LDS 1
#endif
template <int N> struct CCfg {
  static constexpr CShape S = c_shape(N);
  static constexpr int TPT = S.TPT, L = S.L, THREADS = S.THREADS, E = S.E;
  static constexpr bool HALF = CDN_SPLIT_LDS != 0 && plan_exact(N, S.TPT);
  static constexpr int LDS = HALF ? (N * L + 1) / 2 : N * L;  // float2 elements
};

// ---------------------------------------------------------------------------
// x passes.  Thread p of a line holds v[m] = z[p + m TPT] with z = a + i b
// for the two real rows r0, r0+1 packed into one complex line.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float load_y(int mode, const float* __restrict__ X, const float* __restrict__ Xp,
                                        float* __restrict__ Y, float lam, long idx) {
  if (mode == XIN_PRED) {
    const float x = X[idx], xp = Xp[idx];
    float y = x + lam * (x - xp);
    y = y > 0.f ? y : 0.f;
    Y[idx] = y;
    return y;
  }
  if (mode == XIN_SEED) return Y[idx];
  return X[idx];
}

__device__ __forceinline__ float2 shfl2(float2 v, int src) {
  return make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// Half spectra A (row r0) and B (row r0+1) from the packed transform Z:
// A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / (2i).
template <int N, int K = XK_PLAIN>
__device__ __forceinline__ void split_store(float2 (&v)[(XCfg<N, K>::E)], float2* line, int p,
                                            float2* __restrict__ S, long r0, long pitch, bool v0, bool v1) {
  using C = XCfg<N, K>;
  constexpr int E = C::E, TPT = C::TPT;
  auto put = [&](int k, float2 zk, float2 zn) {
    const float2 A = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
    const float2 D = make_float2(0.5f * (zk.x - zn.x), 0.5f * (zk.y + zn.y));
    if (v0) st_stream<NTX>(S + r0 * pitch + k, A);
    if (v1) st_stream<NTX>(S + (r0 + 1) * pitch + k, make_float2(D.y, -D.x));
  };
  if constexpr (C::SHUFFLE) {
    // element N-k of k = p + m TPT sits in lane (TPT-p) % TPT of this line,
    // slot E-1-m (p > 0) or in this lane, slot (E-m) % E (p = 0)
    const int lane = threadIdx.x & 63;
    const int src = lane - p + ((TPT - p) % TPT);
#pragma unroll
    for (int m = 0; m <= E / 2; ++m) {
      const float2 t = shfl2(v[E - 1 - m], src);
      const float2 zn = (p == 0) ? v[(E - m) % E] : t;
      if (m < E / 2 || p == 0) put(p + m * TPT, v[m], zn);
    }
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) line[p + m * TPT] = v[m];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < C::EH; ++e) {
      const int k = p + e * TPT;
      if (k <= N / 2) put(k, line[k], line[k == 0 ? 0 : N - k]);
    }
  }
}

// Packed line Z = A + i B (Hermitian extension of both rows; imaginary parts
// of the kx = 0 and kx = N/2 bins ignored) into v.
template <int N, int K = XK_PLAIN>
__device__ __forceinline__ void merge_load(float2 (&v)[(XCfg<N, K>::E)], float2* line, int p,
                                           const float2* __restrict__ S, long r0, long pitch, bool v0, bool v1) {
  using C = XCfg<N, K>;
  constexpr int E = C::E, TPT = C::TPT;
  auto ld = [&](int k, float2& A, float2& B) {
    A = v0 ? ld_stream<NTX>(S + r0 * pitch + k) : make_float2(0.f, 0.f);
    B = v1 ? ld_stream<NTX>(S + (r0 + 1) * pitch + k) : make_float2(0.f, 0.f);
    if (k == 0 || 2 * k == N) {
      A.y = 0.f;
      B.y = 0.f;
    }
  };
  if constexpr (C::SHUFFLE) {
    float2 A[E / 2 + 1], B[E / 2 + 1];
#pragma unroll
    for (int m = 0; m < E / 2; ++m) ld(p + m * TPT, A[m], B[m]);
    A[E / 2] = make_float2(0.f, 0.f);
    B[E / 2] = make_float2(0.f, 0.f);
    if (p == 0) ld(N / 2, A[E / 2], B[E / 2]);
#pragma unroll
    for (int m = 0; m < E / 2; ++m) v[m] = make_float2(A[m].x - B[m].y, A[m].y + B[m].x);
    const int lane = threadIdx.x & 63;
    const int src = lane - p + ((TPT - p) % TPT);
#pragma unroll
    for (int m = E / 2; m < E; ++m) {
      // Z[k] for k > N/2 is conj(A[N-k]) + i conj(B[N-k])
      float2 a = shfl2(A[E - 1 - m], src), b = shfl2(B[E - 1 - m], src);
      if (p == 0) {
        a = A[E - m];
        b = B[E - m];
      }
      if (p == 0 && m == E / 2)
        v[m] = make_float2(a.x - b.y, a.y + b.x);  // the kx = N/2 bin itself
      else
        v[m] = make_float2(a.x + b.y, b.x - a.y);
    }
  } else {
#pragma unroll
    for (int e = 0; e < C::EH; ++e) {
      const int k = p + e * TPT;
      if (k <= N / 2) {
        float2 A, B;
        ld(k, A, B);
        line[k] = make_float2(A.x - B.y, A.y + B.x);
        if (k != 0 && 2 * k != N) line[N - k] = make_float2(A.x + B.y, B.x - A.y);
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = line[p + m * TPT];
    __syncthreads();
  }
}

// index of the last of n ascending starts that is <= x (starts[0] <= x)
__device__ __forceinline__ int find_start(const long* __restrict__ starts, int n, long x) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// RAGGED: the rows are those of nstk stacks of different heights, stack s
// starting at row row_starts[s] (even); otherwise stacks of stack_rows rows.
template <int N, bool RAGGED = false>
__global__ void __launch_bounds__((XCfg<N, XK_R2C>::THREADS))
    k_x_r2c(const float* __restrict__ X, const float* __restrict__ Xp, float* __restrict__ Y,
            const double* __restrict__ lam, int mode, float2* __restrict__ S, long rows, long pitch,
            long stack_rows, const long* __restrict__ row_starts, int nstk, const float2* __restrict__ tw) {
  using C = XCfg<N, XK_R2C>;
  __shared__ float2 lds[C::NL * C::LS];
  const int tid = threadIdx.x;
  const int l = tid / C::TPT, p = tid % C::TPT;
  const long r0 = 2 * ((long)blockIdx.x * C::NL + l);
  const bool v0 = r0 < rows, v1 = r0 + 1 < rows;
  float2* line = lds + l * C::LS;
  // the factor of this row pair's stack
  long stk = 0;
  if (mode == XIN_PRED && v0) {
    if constexpr (RAGGED)
      stk = find_start(row_starts, nstk, r0);
    else
      stk = r0 / stack_rows;
  }
  const float lamf = (mode == XIN_PRED) ? (float)lam[4 * stk] : 0.f;
  float2 v[C::E];
  // all loads are issued before any store
  const float* src = (mode == XIN_SEED) ? Y : X;
  float a[C::E], b[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int i = p + m * C::TPT;
    a[m] = v0 ? ld_stream<NTX>(src + r0 * N + i) : 0.f;
    b[m] = v1 ? ld_stream<NTX>(src + (r0 + 1) * N + i) : 0.f;
  }
  if (mode == XIN_PRED) {
    // updatePrediction_kernel (src/RLgpuImpl.cu:658-666)
    float ap[C::E], bp[C::E];
#pragma unroll
    for (int m = 0; m < C::E; ++m) {
      const int i = p + m * C::TPT;
      ap[m] = v0 ? ld_stream<NTX>(Xp + r0 * N + i) : 0.f;
      bp[m] = v1 ? ld_stream<NTX>(Xp + (r0 + 1) * N + i) : 0.f;
    }
#pragma unroll
    for (int m = 0; m < C::E; ++m) {
      const int i = p + m * C::TPT;
      float ya = a[m] + lamf * (a[m] - ap[m]);
      float yb = b[m] + lamf * (b[m] - bp[m]);
      a[m] = ya > 0.f ? ya : 0.f;
      b[m] = yb > 0.f ? yb : 0.f;
      if (v0) st_stream<NTX>(Y + r0 * N + i, a[m]);
      if (v1) st_stream<NTX>(Y + (r0 + 1) * N + i, b[m]);
    }
  }
#pragma unroll
  for (int m = 0; m < C::E; ++m) v[m] = make_float2(a[m], b[m]);
  fft_line<N, C::TPT, 1, false>(v, line, p, tw);
  __syncthreads();
  split_store<N, XK_R2C>(v, line, p, S, r0, pitch, v0, v1);
}

template <int N>
__global__ void __launch_bounds__(XCfg<N>::THREADS)
    k_x_c2r(const float2* __restrict__ S, float* __restrict__ out, long rows, long pitch,
            const float2* __restrict__ tw) {
  using C = XCfg<N>;
  __shared__ float2 lds[C::NL * C::LS];
  const int tid = threadIdx.x;
  const int l = tid / C::TPT, p = tid % C::TPT;
  const long r0 = 2 * ((long)blockIdx.x * C::NL + l);
  const bool v0 = r0 < rows, v1 = r0 + 1 < rows;
  float2* line = lds + l * C::LS;
  float2 v[C::E];
  merge_load<N>(v, line, p, S, r0, pitch, v0, v1);
  fft_line<N, C::TPT, 1, true>(v, line, p, tw);
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int i = p + m * C::TPT;
    if (v0) out[r0 * N + i] = v[m].x;
    if (v1) out[(r0 + 1) * N + i] = v[m].y;
  }
}

// LRcore_kernel (src/RLgpuImpl.cu:416-434)
__device__ __forceinline__ float lrcore(float c, float raw, float eps) {
  c = fabsf(c) > 0.f ? c : eps;
  c = raw / c + eps;
  c = c > 10.f ? 10.f : c;
  c = c < -10.f ? -10.f : c;
  return c;
}

template <int N>
__global__ void __launch_bounds__((XCfg<N, XK_LR>::THREADS))
    k_x_lrcore(float2* __restrict__ S, const float* __restrict__ raw, long rows, long pitch,
               const float2* __restrict__ tw, float eps) {
  using C = XCfg<N, XK_LR>;
  __shared__ float2 lds[C::NL * C::LS];
  const int tid = threadIdx.x;
  const int l = tid / C::TPT, p = tid % C::TPT;
  const long r0 = 2 * ((long)blockIdx.x * C::NL + l);
  const bool v0 = r0 < rows, v1 = r0 + 1 < rows;
  float2* line = lds + l * C::LS;
  float2 v[C::E];
  merge_load<N, XK_LR>(v, line, p, S, r0, pitch, v0, v1);
  fft_line<N, C::TPT, 1, true>(v, line, p, tw);
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int i = p + m * C::TPT;
    const float ra = v0 ? ld_stream<NTX>(raw + r0 * N + i) : 0.f;
    const float rb = v1 ? ld_stream<NTX>(raw + (r0 + 1) * N + i) : 0.f;
    v[m] = make_float2(lrcore(v[m].x, ra, eps), lrcore(v[m].y, rb, eps));
  }
  __syncthreads();
  fft_line<N, C::TPT, 1, false>(v, line, p, tw);
  __syncthreads();
  split_store<N, XK_LR>(v, line, p, S, r0, pitch, v0, v1);
}

// Deterministic workgroup sum of two doubles; out[0], out[1] written by
// thread 0.  Wavefront shuffles only over full wavefronts: with a partial
// last wavefront the first one gathers the entries (workgroups of more than
// 64 threads) or thread 0 sums them alone (64 or fewer).
template <int THREADS>
__device__ __forceinline__ void block_sum2(double a, double b, double (*red)[THREADS], double* out) {
  const int tid = threadIdx.x;
  if constexpr (THREADS % 64 == 0) {
    a = wave_sum(a);
    b = wave_sum(b);
    if ((tid & 63) == 0) {
      red[0][tid >> 6] = a;
      red[1][tid >> 6] = b;
    }
    __syncthreads();
    if (tid == 0) {
      double sa = 0.0, sb = 0.0;
      for (int w = 0; w < THREADS / 64; ++w) {
        sa += red[0][w];
        sb += red[1][w];
      }
      out[0] = sa;
      out[1] = sb;
    }
  } else if constexpr (THREADS > 64) {
    // a partial last wavefront: lane t of the first (full) wavefront sums
    // entries t, t + 64, ... in that order, then a wavefront sum
    red[0][tid] = a;
    red[1][tid] = b;
    __syncthreads();
    if (tid < 64) {
      double sa = 0.0, sb = 0.0;
#pragma unroll
      for (int w = 0; w < (THREADS + 63) / 64; ++w) {
        if (tid + 64 * w < THREADS) {
          sa += red[0][tid + 64 * w];
          sb += red[1][tid + 64 * w];
        }
      }
      sa = wave_sum(sa);
      sb = wave_sum(sb);
      if (tid == 0) {
        out[0] = sa;
        out[1] = sb;
      }
    }
  } else {
    red[0][tid] = a;
    red[1][tid] = b;
    __syncthreads();
    if (tid == 0) {
      double sa = 0.0, sb = 0.0;
      for (int w = 0; w < THREADS; ++w) {
        sa += red[0][w];
        sb += red[1][w];
      }
      out[0] = sa;
      out[1] = sb;
    }
  }
}

template <int N>
__global__ void __launch_bounds__(XCfg<N>::THREADS)
    k_x_update(const float2* __restrict__ S, const float* __restrict__ Y, float* __restrict__ Xout,
               float* G, double* __restrict__ part, int dots, long rows, long pitch,
               const float2* __restrict__ tw) {
  using C = XCfg<N>;
  __shared__ float2 lds[C::NL * C::LS];
  __shared__ double red[2][C::THREADS];
  const int tid = threadIdx.x;
  const int l = tid / C::TPT, p = tid % C::TPT;
  const long r0 = 2 * ((long)blockIdx.x * C::NL + l);
  const bool v0 = r0 < rows, v1 = r0 + 1 < rows;
  float2* line = lds + l * C::LS;
  float2 v[C::E];
  merge_load<N>(v, line, p, S, r0, pitch, v0, v1);
  fft_line<N, C::TPT, 1, true>(v, line, p, tw);
  // all loads are issued before any store (G is read, then overwritten)
  float ya[C::E], yb[C::E], ga[C::E], gb[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int i = p + m * C::TPT;
    ya[m] = v0 ? ld_stream<NTX>(Y + r0 * N + i) : 0.f;
    yb[m] = v1 ? ld_stream<NTX>(Y + (r0 + 1) * N + i) : 0.f;
    ga[m] = (dots && v0) ? ld_stream<NTX>(G + r0 * N + i) : 0.f;
    gb[m] = (dots && v1) ? ld_stream<NTX>(G + (r0 + 1) * N + i) : 0.f;
  }
  double num = 0.0, den = 0.0;
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int i = p + m * C::TPT;
    // currEstimate_kernel and currPrevDiff_kernel (src/RLgpuImpl.cu:462-508)
    float xa = v[m].x * ya[m], xb = v[m].y * yb[m];
    xa = xa > 0.f ? xa : 0.f;
    xb = xb > 0.f ? xb : 0.f;
    const float gna = xa - ya[m], gnb = xb - yb[m];
    if (dots) {
      // innerProduct_kernel: FP32 products, FP64 sums (src/RLgpuImpl.cu:583-592)
      if (v0) {
        num += (double)(gna * ga[m]);
        den += (double)(ga[m] * ga[m]);
      }
      if (v1) {
        num += (double)(gnb * gb[m]);
        den += (double)(gb[m] * gb[m]);
      }
    }
    if (v0) {
      st_stream<NTX>(Xout + r0 * N + i, xa);
      st_stream<NTX>(G + r0 * N + i, gna);
    }
    if (v1) {
      st_stream<NTX>(Xout + (r0 + 1) * N + i, xb);
      st_stream<NTX>(G + (r0 + 1) * N + i, gnb);
    }
  }
  if (dots) block_sum2(num, den, red, part + 2 * (long)blockIdx.x);
}

// Workgroups are dispatched round-robin over the 8 XCDs (separate L2s).
// Renumber them so that consecutive logical tiles, which may share 128-byte
// lines when a tile row is narrower than a line, run on the same XCD.
#ifndef CDN_XCD_REMAP
#define CDN_XCD_REMAP 1
#endif
template <int L>
__device__ __forceinline__ long xcd_logical_block(long b, long G) {
  // 1: only tiles narrower than a 128-byte line (they share lines with a
  // neighbour); 2: every tile (neighbouring tiles also share pages)
  if constexpr (CDN_XCD_REMAP == 0 || (CDN_XCD_REMAP == 1 && L * 8 >= 128)) return b;
  const long q = G / 8, r = G % 8, x = b % 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + b / 8;
}

// ---------------------------------------------------------------------------
// y / z column passes: thread (l, p) holds column l's elements p + m TPT
// ---------------------------------------------------------------------------
// the transform of column l of a tile (LDS as declared by the kernel)
template <int N, bool INV>
__device__ __forceinline__ void col_line(float2 (&v)[CCfg<N>::E], float2* lds, int l, int p,
                                         const float2* __restrict__ tw) {
  using C = CCfg<N>;
  if constexpr (C::HALF) {
    fft_line_half<N, C::TPT, C::L, INV>(v, reinterpret_cast<float*>(lds) + l, p, tw);
  } else {
    fft_line<N, C::TPT, C::L, INV>(v, lds + l, p, tw);
  }
}

// tile lb (slab-major) of a column pass; lds as declared by the kernel
template <int N, bool INV>
__device__ __forceinline__ void col_tile(float2* lds, float2* __restrict__ S, long ncols, long line_stride,
                                         long outer_stride, long ntiles, long lb, const float2* __restrict__ tw) {
  using C = CCfg<N>;
  const int tid = threadIdx.x;
  const int l = tid % C::L, p = tid / C::L;
  const long outer = lb / ntiles, tile = lb - outer * ntiles;
  const long c = tile * C::L + l;
  const bool valid = c < ncols;
  float2* base = S + outer * outer_stride + c;
  float2 v[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m)
    v[m] = valid ? ld_stream<NTCL>(base + (long)(p + m * C::TPT) * line_stride) : make_float2(0.f, 0.f);
  col_line<N, INV>(v, lds, l, p, tw);
  if (valid) {
#pragma unroll
    for (int m = 0; m < C::E; ++m) st_stream<NTCS>(base + (long)(p + m * C::TPT) * line_stride, v[m]);
  }
}

template <int N, bool INV>
__global__ void __launch_bounds__(CCfg<N>::THREADS)
    k_col(float2* __restrict__ S, long ncols, long line_stride, long outer_stride, long ntiles,
          const float2* __restrict__ tw) {
  __shared__ float2 lds[CCfg<N>::LDS];
  col_tile<N, INV>(lds, S, ncols, line_stride, outer_stride, ntiles,
                   xcd_logical_block<CCfg<N>::L>(blockIdx.x, gridDim.x), tw);
}

// segment holding work item `it`: the last whose item0 <= it (segs[0].item0 = 0)
template <class Seg>
__device__ __forceinline__ int find_seg(const Seg* __restrict__ segs, int n, long it) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].item0 <= it)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// The column transform of every stack whose columns have length N, stacks
// of different shapes in one launch (ColSeg per stack).  One workgroup per
// item; or (PERSIST) a persistent grid of gridDim.x workgroups, a multiple of
// 8 so an item runs on the XCD the block order assumes, workgroup b taking
// items b, b + gridDim.x, ...  (The loop keeps more values live: 112 against
// 61 VGPRs at length 1024, so it is an option, not the default.)
template <int N, bool INV, bool PERSIST>
__global__ void __launch_bounds__(CCfg<N>::THREADS)
    k_col_ragged(const ColSeg* __restrict__ segs, int nseg, long total, const float2* __restrict__ tw) {
  __shared__ float2 lds[CCfg<N>::LDS];
  long it = blockIdx.x;
  do {
    const long lb = xcd_logical_block<CCfg<N>::L>(it, total);
    const ColSeg& g = segs[find_seg(segs, nseg, lb)];
    col_tile<N, INV>(lds, g.S, g.ncols, g.line_stride, g.outer_stride, g.ntiles, lb - g.item0, tw);
    if constexpr (!PERSIST) break;
    __syncthreads();  // the next item's exchanges reuse lds
    it += gridDim.x;
  } while (it < total);
}

// Kinds of z filter: the OTF table chosen at run time (3-D or rotationally
// averaged), or built for the rotationally averaged table only (the common
// case; without the 3-D interpolation's registers).
enum ZfKind : int { ZF_ANY = 0, ZF_RA = 1 };

// Waves per SIMD the z filter is compiled for (1: the compiler's choice).
// Length 256 (32-column tiles of 512 threads, 32 KB of LDS) at 6 waves, three
// workgroups per CU instead of two, for 80 registers and 15 spilled dwords:
// 4% faster at 256 x 1024 x 1024; the other lengths spill far more and
// measured slower (64: 0.018 -> 0.023 ms at 64 x 256 x 256).  Length 512
// (16-column tiles of 512 threads) of the averaged-table kind fits 6 waves
// in 76 registers without spilling.
#ifndef CDN_ZF_WAVES
#define CDN_ZF_WAVES 0
#endif
constexpr int zf_waves(int N, int kind = ZF_ANY) {
  if (CDN_ZF_WAVES > 0) return CDN_ZF_WAVES;
  const CShape c = c_shape(N);
  const bool half = CDN_SPLIT_LDS != 0 && plan_exact(N, c.TPT);
  if (N == 256 && c.THREADS == 512 && half && N * c.L * 4 <= 32768) return 6;
  if (N == 512 && kind == ZF_RA && c.THREADS == 512 && c.E == 16 && half) return 6;
  return 1;
}
// Rows of the rotationally averaged OTF table staged in LDS per tile.  A
// tile is one ky and L adjacent kx, so its radii span a few table rows.
// Measured: z length 128 (128 x 512 x 810) 0.117 -> 0.104 ms; 64 no change;
// 256 (256 x 1024 x 1024, an OTF support ending near 70% of Nyquist): the
// averaged-table kind 0.493 -> 0.457 ms (1 spilled dword, 44 KB of LDS),
// the kind with both tables 0.508 -> 0.527 (15 spilled dwords); 512 slower
// (one workgroup per CU fewer).  0 disables.
#ifndef CDN_ZF_ROWS
#define CDN_ZF_ROWS 6
#endif
// scaled signed ky of slab row `outer` (k > n>>1 -> k - n)
__device__ __forceinline__ float zf_ky(long outer, long ny, float kyscale) {
  const int kyi = (outer > (ny >> 1)) ? (int)(outer - ny) : (int)outer;
  return (float)kyi * kyscale;
}

// per-length override (n, rows; rows < 0: none), for run-time variants
#ifndef CDN_ZF_ROWS_OVERRIDE
#define CDN_ZF_ROWS_OVERRIDE 0, 0
#endif
constexpr int kZfRowsOverride[] = {CDN_ZF_ROWS_OVERRIDE};
constexpr int zf_rows(int N, int kind = ZF_ANY) {
  const int o = lookup_override(kZfRowsOverride, N);
  return o > 0 ? o : o < 0 ? 0 : (N == 128 || (N == 256 && kind == ZF_RA)) ? CDN_ZF_ROWS : 0;
}
template <int N, int KIND> struct ZRows {
  static constexpr int R = zf_rows(N, KIND);
};

// filter_kernel / filterConj_kernel (src/RLgpuImpl.cu:331-350) with the OTF
// value computed on the fly as in dev_otfinterpolate (src/RLgpuImpl.cu:228-329)
template <int N, int KIND = ZF_ANY>
__global__ void __launch_bounds__(CCfg<N>::THREADS) __attribute__((amdgpu_waves_per_eu(zf_waves(N, KIND))))
    k_z_filter(float2* __restrict__ S, long ncols, long line_stride, long outer_stride, long ntiles, long ny,
               long stack_stride, int conj, OtfParams otf, const float2* __restrict__ tw) {
  using C = CCfg<N>;
  constexpr int RS = ZRows<N, KIND>::R;
  constexpr int PER = RS > 0 ? (RS * N + C::THREADS - 1) / C::THREADS : 1;
  constexpr bool ZRA = KIND == ZF_RA;
  __shared__ float2 lds[C::LDS];
  __shared__ float2 otf_rows[RS > 0 ? RS * N : 1];
  const long lb = xcd_logical_block<C::L>(blockIdx.x, gridDim.x);
#include "zf_tile_body.inc"
}

// The z filter of every stack whose z length is N, stacks of different shapes
// and OTFs in one launch (ZSeg per stack); items and PERSIST as k_col_ragged.
// KIND as k_z_filter's (ZF_RA: every segment's table is rotationally
// averaged), so that each stack runs the code its one-stack engine runs.
template <int N, bool PERSIST, int KIND = ZF_ANY>
__global__ void __launch_bounds__(CCfg<N>::THREADS) __attribute__((amdgpu_waves_per_eu(zf_waves(N, KIND))))
    k_z_ragged(const ZSeg* __restrict__ segs, int nseg, long total, int conj, const float2* __restrict__ tw) {
  using C = CCfg<N>;
  constexpr int RS = ZRows<N, KIND>::R;
  constexpr int PER = RS > 0 ? (RS * N + C::THREADS - 1) / C::THREADS : 1;
  __shared__ float2 lds[C::LDS];
  __shared__ float2 otf_rows[RS > 0 ? RS * N : 1];
  long it = blockIdx.x;
  do {
    const long gi = xcd_logical_block<C::L>(it, total);
    const ZSeg* g = segs + find_seg(segs, nseg, gi);
    float2* __restrict__ S = g->S;
    const long ncols = g->ncols, line_stride = g->line_stride, outer_stride = g->outer_stride;
    const long ntiles = g->ntiles, ny = g->ny, stack_stride = 0, lb = gi - g->item0;
    const OtfParams& otf = g->otf;
    constexpr bool ZRA = KIND == ZF_RA;
    {
#include "zf_tile_body.inc"
    }
    if constexpr (!PERSIST) break;
    __syncthreads();  // the next item's exchanges and staged rows reuse lds
    it += gridDim.x;
  } while (it < total);
}

}  // namespace cdn
