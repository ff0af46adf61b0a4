// Warm-start network on one GPU: parameters (FP32 master + packed BF16 MFMA
// operands), activations of one batch geometry, forward, backward, Adam, a
// device-resident training set and the training loop.
//
// Architecture (u = stack / mean(stack) per sample, channels-last BF16):
//   e1 = relu(conv3(u; 1->16))                    full resolution
//   e2 = relu(conv3(pool(e1); 16->32))            1/2
//   e3 = relu(conv3(pool(e2); 32->64))            1/4
//   d3 = relu(conv3(e3; 64->32))                  1/4
//   d2 = relu(conv3(up(d3) + e2; 32->16))         1/2
//   y  = u + bh + sum_c Wh[c] (up(d2) + e1)[c]    full resolution (1x1x1 head)
// The RL seed is max(y, 0) * mean(stack); training minimises the mean
// squared error of y against target / mean(stack).
#pragma once

#include "decon.hpp"
#include "net_kernels.hpp"

#include <cstdint>

namespace cdn {

constexpr int NET_CONVS = 5;
constexpr int NET_CIN[NET_CONVS] = {1, 16, 32, 64, 32};
constexpr int NET_COUT[NET_CONVS] = {16, 32, 64, 32, 16};

struct WarmNet {
  long w_off[NET_CONVS] = {}, b_off[NET_CONVS] = {}, head_off = 0, nparam = 0;
  DevBuf P, Gr, M, V;                     // FP32 parameters, gradient, Adam moments
  DevBuf wf[NET_CONVS], wb[NET_CONVS];    // packed BF16 forward / adjoint operands
  int adam_t = 0;
  // current batch geometry and its buffers
  Vol L1{0, 0, 0, 0}, L2{0, 0, 0, 0}, L3{0, 0, 0, 0};
  DevBuf inv_s, sum_part;
  DevBuf e1, p1, e2, p2, e3, d3, a2, d2;
  DevBuf dout, g5, gA2, g4, gE3, g3, gP2, g2, gP1, g1;
  DevBuf hpart, cpart, wpart, loss;
  // training set (FP32 stacks and targets of one patch shape)
  long ds_count = 0;
  int ds_nz = 0, ds_ny = 0, ds_nx = 0;
  DevBuf ds_raw, ds_tgt, b_raw, b_tgt;

  WarmNet();
  static long num_params();
  void init(uint64_t seed, hipStream_t st);   // He-normal weights, zero biases, zero head
  void set_params(const float* host, hipStream_t st);
  void get_params(float* host, hipStream_t st) const;
  void pack(hipStream_t st);
  void geometry(int nb, int nz, int ny, int nx, bool training);
  // activations of `raw` (device, nb samples of the current geometry); with
  // keep_e1 = false (inference) the full-size e1 is not stored: the first
  // layer writes only its pooled output and head() recomputes e1
  void forward(hipStream_t st, const float* raw, bool keep_e1 = true);
  // mode 0: normalised output, 1: RL seed (both into out, device)
  void head(hipStream_t st, const float* raw, int mode, float* out);
  bool e1_kept = false;
  // loss (device float at loss_slot) and the full gradient into Gr
  void backward(hipStream_t st, const float* raw, const float* tgt, float* loss_slot);
  void adam(hipStream_t st, float lr);
  // copy a BF16 activation / gradient (or FP32 dout) to host as float; returns count
  long fetch(int which, float* host, hipStream_t st);
  void add_samples(const float* raw, const float* tgt, long count, int nz, int ny, int nx, hipStream_t st);
  // `steps` Adam steps over random batches of the training set; losses (host,
  // optional) gets one value per step
  void train(int steps, int batch, float lr, uint64_t seed, hipStream_t st, float* losses);
};

}  // namespace cdn
