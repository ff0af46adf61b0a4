// Device FFT core: radix-2/3/4/5/7/8/16 butterflies and an LDS-resident
// Stockham autosort transform, fully specialised at compile time on the
// length N (all 2,3,5,7-smooth lengths the library instantiates).
//
// Convention: forward X[k] = sum_n x[n] exp(-2 pi i n k / N); inverse uses
// exp(+...) and is unnormalised (the 1/N of the reference's filterGPU,
// src/RLgpuImpl.cu:176-177, is folded into the OTF multiply).
//
// Stockham stage s with radix R and Lp = product of earlier radices, for each
// butterfly j in [0, N/R), k = j mod Lp:
//   y_t = a[j + t N/R] * exp(-+2 pi i t k / (Lp R)),  Y_q = DFT_R(y)_q,
//   b[(j - k) R + k + q Lp] = Y_q.
// Reads and writes of one stage are separated by a workgroup barrier, so one
// LDS buffer is used in place; each thread keeps its butterfly outputs in
// registers across the barrier.
#pragma once

#include "common_dev.hpp"

namespace cdn {

// ---------------------------------------------------------------------------
// In-register DFTs.  INV selects the sign of the exponent (+ for inverse).
// ---------------------------------------------------------------------------
template <int R, bool INV> struct Dft;

template <bool INV> struct Dft<2, INV> {
  static __device__ __forceinline__ void run(float2* x) {
    float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};

template <bool INV> struct Dft<4, INV> {
  static __device__ __forceinline__ void run(float2* x) {
    float2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    float2 c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
    // forward: X1 = b - i d, X3 = b + i d
    float2 jd = INV ? cmuli<1>(d) : cmuli<-1>(d);
    x[0] = cadd(a, c);
    x[2] = csub(a, c);
    x[1] = cadd(b, jd);
    x[3] = csub(b, jd);
  }
};

template <bool INV> struct Dft<8, INV> {
  static __device__ __forceinline__ void run(float2* x) {
    constexpr float h = 0.7071067811865476f;
    float2 a[4], b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      a[t] = cadd(x[t], x[t + 4]);
      b[t] = csub(x[t], x[t + 4]);
    }
    // b_t *= W8^{+-t}
    if (!INV) {
      b[1] = make_float2(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));
      b[2] = cmuli<-1>(b[2]);
      b[3] = make_float2(h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y));
    } else {
      b[1] = make_float2(h * (b[1].x - b[1].y), h * (b[1].x + b[1].y));
      b[2] = cmuli<1>(b[2]);
      b[3] = make_float2(-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y));
    }
    Dft<4, INV>::run(a);
    Dft<4, INV>::run(b);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      x[2 * q] = a[q];
      x[2 * q + 1] = b[q];
    }
  }
};

template <bool INV> struct Dft<16, INV> {
  static __device__ __forceinline__ void run(float2* x) {
    // 16 = 2 x 8, decimation in frequency: a_t = x_t + x_{t+8}, b_t = (x_t - x_{t+8}) W16^t
    constexpr float c1 = 0.9238795325112867f, s1 = 0.3826834323650898f, h = 0.7071067811865476f;
    const float tw[8][2] = {{1.f, 0.f}, {c1, s1}, {h, h}, {s1, c1}, {0.f, 1.f}, {-s1, c1}, {-h, h}, {-c1, s1}};
    float2 a[8], b[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      a[t] = cadd(x[t], x[t + 8]);
      float2 d = csub(x[t], x[t + 8]);
      // W16^t = cos - i sin (forward), cos + i sin (inverse)
      float2 w = make_float2(tw[t][0], INV ? tw[t][1] : -tw[t][1]);
      b[t] = (t == 0) ? d : cmul(d, w);
    }
    Dft<8, INV>::run(a);
    Dft<8, INV>::run(b);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      x[2 * q] = a[q];
      x[2 * q + 1] = b[q];
    }
  }
};

// Odd prime R: X_q = x0 + sum_t a_t cos(2pi tq/R) -+ i sum_t b_t sin(2pi tq/R),
// a_t = x_t + x_{R-t}, b_t = x_t - x_{R-t}, t = 1..(R-1)/2.
template <int R> struct OddConst;
template <> struct OddConst<3> {
  static constexpr float c[1] = {-0.5f};
  static constexpr float s[1] = {0.8660254037844387f};
};
template <> struct OddConst<5> {
  static constexpr float c[2] = {0.30901699437494745f, -0.8090169943749473f};
  static constexpr float s[2] = {0.9510565162951535f, 0.5877852522924732f};
};
template <> struct OddConst<7> {
  static constexpr float c[3] = {0.6234898018587336f, -0.22252093395631434f, -0.900968867902419f};
  static constexpr float s[3] = {0.7818314824680298f, 0.9749279121818236f, 0.43388373911755823f};
};

template <int R, bool INV> struct DftOdd {
  static __device__ __forceinline__ void run(float2* x) {
    constexpr int H = (R - 1) / 2;
    float2 a[H], b[H];
#pragma unroll
    for (int t = 1; t <= H; ++t) {
      a[t - 1] = cadd(x[t], x[R - t]);
      b[t - 1] = csub(x[t], x[R - t]);
    }
    float2 x0 = x[0];
    float2 s0 = x0;
#pragma unroll
    for (int t = 0; t < H; ++t) s0 = cadd(s0, a[t]);
    x[0] = s0;
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      float2 re = x0, im = make_float2(0.f, 0.f);
#pragma unroll
      for (int t = 1; t <= H; ++t) {
        // cos(2 pi t q / R) and sin(2 pi t q / R) via symmetric index m = tq mod R
        const int m = (t * q) % R;
        const int mm = m <= H ? m : R - m;  // cos index
        const float cs = OddConst<R>::c[mm - 1];
        const float sn = m <= H ? OddConst<R>::s[mm - 1] : -OddConst<R>::s[mm - 1];
        re = make_float2(re.x + cs * a[t - 1].x, re.y + cs * a[t - 1].y);
        im = make_float2(im.x + sn * b[t - 1].x, im.y + sn * b[t - 1].y);
      }
      // forward: X_q = re - i im, X_{R-q} = re + i im
      float2 iim = cmuli<1>(im);
      if (!INV) {
        x[q] = csub(re, iim);
        x[R - q] = cadd(re, iim);
      } else {
        x[q] = cadd(re, iim);
        x[R - q] = csub(re, iim);
      }
    }
  }
};
template <bool INV> struct Dft<3, INV> : DftOdd<3, INV> {};
template <bool INV> struct Dft<5, INV> : DftOdd<5, INV> {};
template <bool INV> struct Dft<7, INV> : DftOdd<7, INV> {};

// ---------------------------------------------------------------------------
// Compile-time radix plan.
// ---------------------------------------------------------------------------
struct Plan {
  int ns = 0;
  int r[24] = {};
  int lp[24] = {};
  int twoff[24] = {};
  int twsize = 0;
  int lcm = 1;
  bool ok = false;
};

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }

constexpr Plan make_plan(int N) {
  Plan P{};
  int n = N;
  int ns = 0;
  int two = 0;
  while (n % 2 == 0) {
    n /= 2;
    ++two;
  }
  // powers of two: radix 16 first (few stages, lcm 16); mixed lengths keep
  // the power-of-two radices at most 8 so the radix lcm stays small
  const int maxr = (n == 1) ? 16 : 8;
  while (two >= 4 && maxr == 16) { P.r[ns++] = 16; two -= 4; }
  while (two >= 3) { P.r[ns++] = 8; two -= 3; }
  if (two == 2) P.r[ns++] = 4;
  if (two == 1) P.r[ns++] = 2;
  while (n % 3 == 0) { P.r[ns++] = 3; n /= 3; }
  while (n % 5 == 0) { P.r[ns++] = 5; n /= 5; }
  while (n % 7 == 0) { P.r[ns++] = 7; n /= 7; }
  P.ns = ns;
  P.ok = (n == 1);
  int lp = 1, off = 0, l = 1;
  for (int s = 0; s < ns; ++s) {
    P.lp[s] = lp;
    P.twoff[s] = off;
    if (lp > 1) off += lp * (P.r[s] - 1);
    lp *= P.r[s];
    l = l / gcd_c(l, P.r[s]) * P.r[s];
  }
  P.twsize = off;
  P.lcm = l;
  return P;
}

// True when every stage's butterflies split evenly over TPT threads, so
// thread p owns elements p + m TPT before the first and after the last stage.
constexpr bool plan_exact(int N, int TPT) {
  const Plan P = make_plan(N);
  if (P.ns < 2) return false;
  for (int s = 0; s < P.ns; ++s)
    if ((N / P.r[s]) % TPT != 0) return false;
  return true;
}

// ---------------------------------------------------------------------------
// One Stockham stage over a line held in LDS at element stride SE.
// Threads p in [0, TPT) of the line share the N/R butterflies cyclically.
// tw: this stage's twiddles, tw[k (R-1) + t - 1] = exp(-2 pi i t k / (Lp R)).
// ---------------------------------------------------------------------------
template <int N, int R, int LP, int TPT, int SE, bool INV>
__device__ __forceinline__ void stockham_stage(float2* line, int p, const float2* __restrict__ tw) {
  constexpr int B = N / R;
  constexpr int U = (B + TPT - 1) / TPT;
  constexpr bool EXACT = (B % TPT) == 0;
  float2 v[U][R];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = p + u * TPT;
    if (EXACT || j < B) {
      float2 x[R];
#pragma unroll
      for (int t = 0; t < R; ++t) x[t] = line[(j + t * B) * SE];
      if constexpr (LP > 1) {
        const int k = j % LP;
        const float2* w = tw + k * (R - 1);
#pragma unroll
        for (int t = 1; t < R; ++t) x[t] = INV ? cmulc(x[t], w[t - 1]) : cmul(x[t], w[t - 1]);
      }
      Dft<R, INV>::run(x);
#pragma unroll
      for (int q = 0; q < R; ++q) v[u][q] = x[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = p + u * TPT;
    if (EXACT || j < B) {
      const int k = j % LP;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) line[(base + q * LP) * SE] = v[u][q];
    }
  }
  __syncthreads();
}

// Two consecutive stages (radices R1 then R2, the first at span LP) in one
// pass over LDS.  Super-butterfly j < N/(R1 R2) loads the R1 R2 elements
// j + u N/(R1 R2); they are the inputs of the first stage's butterflies
// j + m N/(R1 R2) (m < R2, twiddle index j mod LP), whose outputs are the
// inputs of the second stage's butterflies (j - j mod LP) R1 + j mod LP + q LP
// (q < R1, twiddle index j mod LP + q LP).  The same operations in the same
// order as the two stages run one after the other, with one LDS round trip
// and two barriers fewer; results agree to FP32 rounding (the compiler may
// pick other products to fuse into multiply-adds in the longer block).
template <int N, int R1, int R2, int LP, int TPT, int SE, bool INV>
__device__ __forceinline__ void stockham_stage2(float2* line, int p, const float2* __restrict__ tw1,
                                                const float2* __restrict__ tw2) {
  constexpr int B = N / (R1 * R2);
  constexpr int U = (B + TPT - 1) / TPT;
  constexpr bool EXACT = (B % TPT) == 0;
  float2 v[U][R1 * R2];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = p + u * TPT;
    if (EXACT || j < B) {
      const int k1 = j % LP;
      float2 y[R2][R1];
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        float2 x[R1];
#pragma unroll
        for (int t = 0; t < R1; ++t) x[t] = line[(j + (m + t * R2) * B) * SE];
        if constexpr (LP > 1) {
          const float2* w = tw1 + k1 * (R1 - 1);
#pragma unroll
          for (int t = 1; t < R1; ++t) x[t] = INV ? cmulc(x[t], w[t - 1]) : cmul(x[t], w[t - 1]);
        }
        Dft<R1, INV>::run(x);
#pragma unroll
        for (int q = 0; q < R1; ++q) y[m][q] = x[q];
      }
#pragma unroll
      for (int q = 0; q < R1; ++q) {
        float2 z[R2];
#pragma unroll
        for (int t = 0; t < R2; ++t) z[t] = y[t][q];
        const float2* w = tw2 + (k1 + q * LP) * (R2 - 1);
#pragma unroll
        for (int t = 1; t < R2; ++t) z[t] = INV ? cmulc(z[t], w[t - 1]) : cmul(z[t], w[t - 1]);
        Dft<R2, INV>::run(z);
#pragma unroll
        for (int q2 = 0; q2 < R2; ++q2) v[u][q * R2 + q2] = z[q2];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = p + u * TPT;
    if (EXACT || j < B) {
      const int k1 = j % LP;
      const int base = (j - k1) * R1 * R2 + k1;
#pragma unroll
      for (int q = 0; q < R1; ++q)
#pragma unroll
        for (int q2 = 0; q2 < R2; ++q2) line[(base + q * LP + q2 * LP * R1) * SE] = v[u][q * R2 + q2];
    }
  }
  __syncthreads();
}

// Stage pairs whose radix product is at most this run as one pass
// (stockham_stage2); 0 runs every stage on its own.  810-point lines
// (stages 2, 3, 3, 3, 3, 5 -> (2, 3), (3, 3), (3, 5)): the LRcore pass
// 0.283 -> 0.208 ms at 128 x 512 x 810 with the same launch shape; a limit
// of 9 ((2, 3), (3, 3), 3, 5) measured 0.208 against 0.200-0.206 for 16.
// Powers of two pair nothing: every pair of their stages holds a radix 16.
#ifndef CDN_FUSE_MAX
#define CDN_FUSE_MAX 16
#endif

template <int N, int TPT, int SE, bool INV, int S = 0>
__device__ __forceinline__ void fft_lds(float2* line, int p, const float2* __restrict__ tw) {
  constexpr Plan P = make_plan(N);
  static_assert(P.ok, "FFT length must be 2,3,5,7-smooth");
  if constexpr (S + 1 < P.ns && P.r[S] * P.r[S + 1] <= CDN_FUSE_MAX) {
    stockham_stage2<N, P.r[S], P.r[S + 1], P.lp[S], TPT, SE, INV>(line, p, tw + P.twoff[S], tw + P.twoff[S + 1]);
    fft_lds<N, TPT, SE, INV, S + 2>(line, p, tw);
  } else if constexpr (S < P.ns) {
    stockham_stage<N, P.r[S], P.lp[S], TPT, SE, INV>(line, p, tw + P.twoff[S]);
    fft_lds<N, TPT, SE, INV, S + 1>(line, p, tw);
  }
}

// ---------------------------------------------------------------------------
// Register-resident variant (exact plans): the first stage reads the
// thread's registers, the last stage writes them; LDS only carries the
// exchanges between stages.
// ---------------------------------------------------------------------------
template <int N, int TPT, int SE, bool INV, int S>
__device__ __forceinline__ void reg_stage(float2 (&v)[N / TPT], float2* line, int p, const float2* __restrict__ tw) {
  constexpr Plan P = make_plan(N);
  constexpr int R = P.r[S], LP = P.lp[S], B = N / R, U = B / TPT;
  constexpr bool FIRST = (S == 0), LAST = (S == P.ns - 1);
  float2 w[U][R];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = p + u * TPT;
    float2 x[R];
#pragma unroll
    for (int t = 0; t < R; ++t) x[t] = FIRST ? v[u + t * U] : line[(j + t * B) * SE];
    if constexpr (LP > 1) {
      const int k = j % LP;
      const float2* wt = tw + P.twoff[S] + k * (R - 1);
#pragma unroll
      for (int t = 1; t < R; ++t) x[t] = INV ? cmulc(x[t], wt[t - 1]) : cmul(x[t], wt[t - 1]);
    }
    Dft<R, INV>::run(x);
#pragma unroll
    for (int q = 0; q < R; ++q) w[u][q] = x[q];
  }
  if constexpr (LAST) {
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < R; ++q) v[u + q * U] = w[u][q];
  } else {
    if constexpr (!FIRST) __syncthreads();  // this stage's reads precede its writes
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = p + u * TPT;
      const int k = j % LP;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) line[(base + q * LP) * SE] = w[u][q];
    }
    __syncthreads();
    reg_stage<N, TPT, SE, INV, S + 1>(v, line, p, tw);
  }
}

// ---------------------------------------------------------------------------
// Register-resident variant with the exchanges in halves: real parts through
// LDS, then imaginary parts, so a line needs N floats of LDS instead of N
// complex values (more workgroups per CU) for two more barriers per
// exchange.  `in` holds stage S's inputs, in[u R + t] = element j + t B with
// j = p + u TPT; each exchange reads the next stage's inputs straight into
// registers.  Same arithmetic, same order as reg_stage.
// ---------------------------------------------------------------------------
template <int N, int TPT, int SE, bool INV, int S>
__device__ __forceinline__ void reg_stage_half(float2 (&in)[N / TPT], float2 (&v)[N / TPT], float* line, int p,
                                               const float2* __restrict__ tw) {
  constexpr Plan P = make_plan(N);
  constexpr int R = P.r[S], LP = P.lp[S], B = N / R, U = B / TPT;
  constexpr bool FIRST = (S == 0), LAST = (S == P.ns - 1);
  float2 w[U * R];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = p + u * TPT;
    float2 x[R];
#pragma unroll
    for (int t = 0; t < R; ++t) x[t] = in[u * R + t];
    if constexpr (LP > 1) {
      const int k = j % LP;
      const float2* wt = tw + P.twoff[S] + k * (R - 1);
#pragma unroll
      for (int t = 1; t < R; ++t) x[t] = INV ? cmulc(x[t], wt[t - 1]) : cmul(x[t], wt[t - 1]);
    }
    Dft<R, INV>::run(x);
#pragma unroll
    for (int q = 0; q < R; ++q) w[u * R + q] = x[q];
  }
  if constexpr (LAST) {
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < R; ++q) v[u + q * U] = w[u * R + q];
  } else {
    constexpr int R2 = P.r[S + 1], B2 = N / R2, U2 = B2 / TPT;
    float2 nx[N / TPT];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      // the previous reads of the line precede these writes
      if (!FIRST || half == 1) __syncthreads();
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = p + u * TPT;
        const int k = j % LP;
        const int base = (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) line[(base + q * LP) * SE] = half ? w[u * R + q].y : w[u * R + q].x;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < U2; ++u) {
        const int j = p + u * TPT;
#pragma unroll
        for (int t = 0; t < R2; ++t) {
          const float f = line[(j + t * B2) * SE];
          if (half)
            nx[u * R2 + t].y = f;
          else
            nx[u * R2 + t].x = f;
        }
      }
    }
    reg_stage_half<N, TPT, SE, INV, S + 1>(nx, v, line, p, tw);
  }
}

// fft_line (below) for exact plans with the exchanges in halves; `line` is
// the float LDS line (N floats at stride SE).  Same contract as fft_line.
template <int N, int TPT, int SE, bool INV>
__device__ __forceinline__ void fft_line_half(float2 (&v)[N / TPT], float* line, int p,
                                              const float2* __restrict__ tw) {
  static_assert(plan_exact(N, TPT), "the half exchange needs an exact plan");
  constexpr Plan P = make_plan(N);
  constexpr int R = P.r[0], U = N / R / TPT;
  float2 in[N / TPT];
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int t = 0; t < R; ++t) in[u * R + t] = v[u + t * U];
  reg_stage_half<N, TPT, SE, INV, 0>(in, v, line, p, tw);
}

// FFT of one line whose thread p holds elements p + m TPT in v[m] (m < N/TPT)
// on entry and on exit.  Contract: the caller has made the LDS line free
// (a barrier after its last use) and must barrier again before reusing it.
template <int N, int TPT, int SE, bool INV>
__device__ __forceinline__ void fft_line(float2 (&v)[N / TPT], float2* line, int p, const float2* __restrict__ tw) {
  static_assert(N % TPT == 0, "threads per line must divide the length");
  constexpr int E = N / TPT;
  if constexpr (plan_exact(N, TPT)) {
    reg_stage<N, TPT, SE, INV, 0>(v, line, p, tw);
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) line[(p + m * TPT) * SE] = v[m];
    __syncthreads();
    fft_lds<N, TPT, SE, INV>(line, p, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = line[(p + m * TPT) * SE];
  }
}

}  // namespace cdn
