// Launchers of the warm-start network kernels (net_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace cdn {

typedef unsigned short bf16_t;  // bfloat16 bits

// one resolution level of a batch: nb samples of nz x ny x nx voxels
struct Vol {
  int nb, nz, ny, nx;
};

// 3x3x3 convolution (cin in {16,32,64}); relu => bias + ReLU epilogue
void net_conv(hipStream_t st, int cin, int cout, bool relu, const bf16_t* in, const bf16_t* wf, const float* bias,
              bf16_t* out, Vol v);
// first layer from the FP32 stack * inv_s[sample] (1 -> 16 channels, bias +
// ReLU): mode 0 writes e1 (out16), 1 writes pool(e1) (out16, half size),
// 2 / 3 also apply up(d2) and the head hp and write y / the RL seed (out32)
void net_first(hipStream_t st, int mode, const float* raw, const float* inv_s, const bf16_t* wf, const float* bias,
               bf16_t* out16, const bf16_t* d2, const float* hp, float* out32, Vol v);
// 32 -> 16 convolution (bias + ReLU) of up(coarse) + skip, the sum formed
// on load (inference: the decoder's a2 is never stored)
// enc2 (16 -> 32, ReLU, bias) writing out and also pooled = pool(out) (2x2x2
// mean, as net_pool computes it); dimensions even
void net_conv_pool_16_32(hipStream_t st, const bf16_t* in, const bf16_t* wf, const float* bias, bf16_t* out,
                         bf16_t* pooled, Vol v);
void net_conv_upadd_32_16(hipStream_t st, const bf16_t* coarse, const bf16_t* skip, const bf16_t* wf,
                          const float* bias, bf16_t* out, Vol v);
void net_pool(hipStream_t st, int c, const bf16_t* in, bf16_t* out, Vol fine);
void net_upadd(hipStream_t st, int c, const bf16_t* coarse, const bf16_t* skip, bf16_t* out, Vol fine);
int net_head_blocks(Vol v);
// mode 0: normalised output; 1: RL seed max(y,0)*s; 2: training (dL/dy + partials)
void net_head(hipStream_t st, int mode, const float* raw, const float* inv_s, const bf16_t* e1, const bf16_t* d2,
              const float* hp, const float* tgt, float* out, float* part, Vol v);
void net_head_reduce(hipStream_t st, const float* part, int blocks, double inv_total, float* grad, long head_off,
                     float* loss);
void net_grad_enc(hipStream_t st, int c, const bf16_t* e, const bf16_t* skip, const float* dout, const float* wh,
                  const bf16_t* dP, bf16_t* dpre, Vol v);
void net_grad_dec(hipStream_t st, int c, const bf16_t* d, const bf16_t* fine, const float* dout, const float* wh,
                  bf16_t* dpre, Vol coarse);
void net_chan_sum(hipStream_t st, int c, const bf16_t* t, long nvox, int blocks, float* part);
void net_wgrad(hipStream_t st, int cin, int cout, const bf16_t* in, const bf16_t* g, Vol v, int G, float* part);
void net_wgrad_in1(hipStream_t st, const float* raw, const float* inv_s, const bf16_t* g, Vol v, int G, float* part);
void net_reduce_rows(hipStream_t st, const float* part, int rows, long cols, float* out);
int net_sum_blocks(int nb);
void net_inv_scale(hipStream_t st, const float* raw, Vol v, double* part, float* inv_s);
// packed operand size (bf16 elements) of a layer as the convolution reads it
long net_packed_weights(int cin, int cout);
void net_pack(hipStream_t st, const float* W, int cin, int cout, int flip, bf16_t* out);
void net_adam(hipStream_t st, float* p, const float* g, float* m, float* v, long n, float lr_t, float b1, float b2,
              float eps, float c2);

}  // namespace cdn
