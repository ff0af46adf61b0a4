// Launch wrappers for the preprocessing, reduction and geometry kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cdn {

void u16_to_f32_crop(hipStream_t st, const uint16_t* in, int nx, int ny, int nz, float* out, int onx, int ony,
                     int onz);
void apodize(hipStream_t st, float* img, int nx, int ny, int nz, int napodize);
void bgsub(hipStream_t st, float* img, long n, float bg);
// uint16 nz x ny x nx -> FP32 max(v - bg, 0) (u16_to_f32_crop without a
// crop, then bgsub, in one pass)
void u16_to_f32_bgsub(hipStream_t st, const uint16_t* in, int nx, int ny, int nz, float* out, float bg);
void rescale(hipStream_t st, float* img, long n, float s);
// d_part: >= 2048 doubles, d_acc: >= 4 doubles, d_mean: 1 float.  Synchronises.
double mean_above_background(hipStream_t st, const float* img, long n, double* d_part, double* d_acc, float* d_mean);
void zblend(hipStream_t st, float* img, int nx, int ny, int nz, int nZblend);
void fill(hipStream_t st, float* p, long n, float v);
// sum, min, max of n floats (FP64 sum); d_part holds >= 3 * 1024 doubles.  Synchronises.
void stats(hipStream_t st, const float* img, long n, double* d_part, double out[3]);
// sum (a-b)^2 and sum b^2 of n floats (FP64); d_part holds >= 2 * 1024 doubles.  Synchronises.
void diff_norm(hipStream_t st, const float* a, const float* b, long n, double* d_part, double out[2]);
// number of NaN / Inf values among n floats; d_part holds >= 2 * 1024
// doubles.  Synchronises.
long count_nonfinite(hipStream_t st, const float* img, long n, double* d_part);
// camera correction of a uint16 stack with 3 parameter planes (a, b, offset)
void camcor(hipStream_t st, const uint16_t* raw, const float* cam, int nx, int ny, int nz, uint16_t* out);
void add_scalar(hipStream_t st, float* p, long n, float c);
// Wiener step on a transformed half spectrum with the kz-interpolated RA table
void wiener_multiply(hipStream_t st, float2* S, int nxh, int ny, int nz, long pitch, const float2* tz, int nr,
                     float kxscale, float kyscale, float dkx, float dky, float rcut, float w2, float inv_n);
// bytes must be a multiple of 16
void stream_copy(hipStream_t st, const void* in, void* out, size_t bytes);
void copy_f32(hipStream_t st, const float* in, float* out, size_t n);
// acceleration factor of nstacks stacks, nb FP64 partial pairs each
// (stack s: part[2 (s nb ..)]) -> lam[4 s] = factor, lam[4 s + 1..2] = sums.
// The partials are used up: chunk sums are written over some of them.
void accel_factor(hipStream_t st, double* part, int nb, double eps, double* lam, int nstacks = 1);
// the same for stacks of different sizes: stack s sums the ranges[2 s + 1]
// partial pairs from pair ranges[2 s] (ranges: device, 2 ints per stack)
void accel_factor_ragged(hipStream_t st, double* part, const int* ranges, int nstacks, double eps, double* lam);
void deskew(hipStream_t st, const float* in, int nx, int ny, int nz, double factor, float* out, int nxOut,
            int extraShift, float padVal);
void rotate(hipStream_t st, const float* in, int nx, int ny, int nz, const float* rot, float* out, int nx_out,
            int nz_out);
void affine(hipStream_t st, const float* in, int nx, int ny, int nz, float* out, const float* mat, int ra, float dx,
            float dy, float dz);

}  // namespace cdn
