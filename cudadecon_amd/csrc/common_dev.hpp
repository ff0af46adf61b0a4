// Device-side helpers shared by all kernels (also compiled at run time by
// hipRTC for FFT lengths outside the precompiled set).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace cdn {

// ---- complex helpers (float2 = re, im) ----
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by +i (s = +1) or -i (s = -1)
template <int S> __device__ __forceinline__ float2 cmuli(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Streaming global accesses: NT selects non-temporal instructions for data
// that a pass reads or writes exactly once (measured per kernel family).
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
template <bool NT> __device__ __forceinline__ float2 ld_stream(const float2* p) {
  if constexpr (NT) {
    const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
    return make_float2(v.x, v.y);
  } else {
    return *p;
  }
}
template <bool NT> __device__ __forceinline__ float ld_stream(const float* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT> __device__ __forceinline__ void st_stream(float2* p, float2 v) {
  if constexpr (NT) {
    f2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
  } else {
    *p = v;
  }
}
template <bool NT> __device__ __forceinline__ void st_stream(float* p, float v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Sum over a wavefront of 64 lanes (doubles), result valid in every lane.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace cdn
