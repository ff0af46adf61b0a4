// Host orchestration of the warm-start network (architecture in net.hpp).
#include "net.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace cdn {
namespace {

constexpr int WG_CHUNKS = 64;       // weight-gradient partial rows (3x3x3 layers, x27 taps)
constexpr int WG_CHUNKS_IN1 = 512;  // first layer
constexpr int CS_BLOCKS = 256;      // bias-gradient partial rows

struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() {  // Box-Muller
    const double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

long vox(const Vol& v) { return (long)v.nb * v.nz * v.ny * v.nx; }

}  // namespace

WarmNet::WarmNet() {
  long o = 0;
  for (int l = 0; l < NET_CONVS; ++l) {
    w_off[l] = o;
    o += 27L * NET_CIN[l] * NET_COUT[l];
    b_off[l] = o;
    o += NET_COUT[l];
  }
  head_off = o;
  nparam = o + 17;
}

long WarmNet::num_params() { return WarmNet().nparam; }

void WarmNet::init(uint64_t seed, hipStream_t st) {
  std::vector<float> h((size_t)nparam, 0.f);
  SplitMix rng{seed};
  for (int l = 0; l < NET_CONVS; ++l) {
    const double sd = std::sqrt(2.0 / (27.0 * NET_CIN[l]));
    for (long i = 0; i < 27L * NET_CIN[l] * NET_COUT[l]; ++i) h[(size_t)(w_off[l] + i)] = (float)(sd * rng.normal());
  }
  set_params(h.data(), st);
}

void WarmNet::set_params(const float* host, hipStream_t st) {
  const size_t b = sizeof(float) * (size_t)nparam;
  P.ensure(b);
  Gr.ensure(b);
  M.ensure(b);
  V.ensure(b);
  CDN_HIP(hipMemcpyAsync(P.p, host, b, hipMemcpyHostToDevice, st));
  CDN_HIP(hipMemsetAsync(Gr.p, 0, b, st));
  CDN_HIP(hipMemsetAsync(M.p, 0, b, st));
  CDN_HIP(hipMemsetAsync(V.p, 0, b, st));
  adam_t = 0;
  pack(st);
  CDN_HIP(hipStreamSynchronize(st));
}

void WarmNet::get_params(float* host, hipStream_t st) const {
  CDN_HIP(hipMemcpyAsync(host, P.p, sizeof(float) * (size_t)nparam, hipMemcpyDeviceToHost, st));
  CDN_HIP(hipStreamSynchronize(st));
}

void WarmNet::pack(hipStream_t st) {
  const float* p = P.as<float>();
  for (int l = 0; l < NET_CONVS; ++l) {
    const int ci = NET_CIN[l], co = NET_COUT[l];
    wf[l].ensure(sizeof(bf16_t) * (size_t)net_packed_weights(ci, co));
    net_pack(st, p + w_off[l], ci, co, 0, wf[l].as<bf16_t>());
    if (ci > 1) {
      wb[l].ensure(sizeof(bf16_t) * (size_t)net_packed_weights(co, ci));
      net_pack(st, p + w_off[l], ci, co, 1, wb[l].as<bf16_t>());
    }
  }
}

void WarmNet::geometry(int nb, int nz, int ny, int nx, bool training) {
  if (nb < 1 || nz < 4 || ny < 4 || nx < 4 || nz % 4 || ny % 4 || nx % 4)
    fail("warm-start network: every dimension must be a positive multiple of 4 (got " + std::to_string(nz) + " x " +
         std::to_string(ny) + " x " + std::to_string(nx) + ")");
  L1 = Vol{nb, nz, ny, nx};
  L2 = Vol{nb, nz / 2, ny / 2, nx / 2};
  L3 = Vol{nb, nz / 4, ny / 4, nx / 4};
  const size_t n1 = (size_t)vox(L1), n2 = (size_t)vox(L2), n3 = (size_t)vox(L3), h = sizeof(bf16_t);
  inv_s.ensure(sizeof(float) * nb);
  sum_part.ensure(sizeof(double) * (size_t)nb * net_sum_blocks(nb));
  if (training) e1.ensure(h * n1 * 16);
  p1.ensure(h * n2 * 16);
  e2.ensure(h * n2 * 32);
  p2.ensure(h * n3 * 32);
  e3.ensure(h * n3 * 64);
  d3.ensure(h * n3 * 32);
  if (training) a2.ensure(h * n2 * 32);
  d2.ensure(h * n2 * 16);
  if (training) {
    dout.ensure(sizeof(float) * n1);
    g5.ensure(h * n2 * 16);
    gA2.ensure(h * n2 * 32);
    g4.ensure(h * n3 * 32);
    gE3.ensure(h * n3 * 64);
    g3.ensure(h * n3 * 64);
    gP2.ensure(h * n3 * 32);
    g2.ensure(h * n2 * 32);
    gP1.ensure(h * n2 * 16);
    g1.ensure(h * n1 * 16);
    hpart.ensure(sizeof(float) * 18 * (size_t)net_head_blocks(L1));
    cpart.ensure(sizeof(float) * 64 * CS_BLOCKS);
    wpart.ensure(sizeof(float) * std::max(27L * 64 * 32 * WG_CHUNKS, 27L * 16 * WG_CHUNKS_IN1));
    loss.ensure(sizeof(float) * 4);
  }
}

void WarmNet::forward(hipStream_t st, const float* raw, bool keep_e1) {
  const float* p = P.as<float>();
  net_inv_scale(st, raw, L1, sum_part.as<double>(), inv_s.as<float>());
  if (keep_e1) {
    e1.ensure(sizeof(bf16_t) * 16 * (size_t)vox(L1));
    a2.ensure(sizeof(bf16_t) * 32 * (size_t)vox(L2));
    net_first(st, 0, raw, inv_s.as<float>(), wf[0].as<bf16_t>(), p + b_off[0], e1.as<bf16_t>(), nullptr, nullptr,
              nullptr, L1);
    net_pool(st, 16, e1.as<bf16_t>(), p1.as<bf16_t>(), L1);
  } else {
    net_first(st, 1, raw, inv_s.as<float>(), wf[0].as<bf16_t>(), p + b_off[0], p1.as<bf16_t>(), nullptr, nullptr,
              nullptr, L1);
  }
  e1_kept = keep_e1;
  net_conv_pool_16_32(st, p1.as<bf16_t>(), wf[1].as<bf16_t>(), p + b_off[1], e2.as<bf16_t>(), p2.as<bf16_t>(), L2);
  net_conv(st, 32, 64, true, p2.as<bf16_t>(), wf[2].as<bf16_t>(), p + b_off[2], e3.as<bf16_t>(), L3);
  net_conv(st, 64, 32, true, e3.as<bf16_t>(), wf[3].as<bf16_t>(), p + b_off[3], d3.as<bf16_t>(), L3);
  if (keep_e1) {  // training keeps a2 (the input of the decoder's weight gradient)
    net_upadd(st, 32, d3.as<bf16_t>(), e2.as<bf16_t>(), a2.as<bf16_t>(), L2);
    net_conv(st, 32, 16, true, a2.as<bf16_t>(), wf[4].as<bf16_t>(), p + b_off[4], d2.as<bf16_t>(), L2);
  } else {
    net_conv_upadd_32_16(st, d3.as<bf16_t>(), e2.as<bf16_t>(), wf[4].as<bf16_t>(), p + b_off[4], d2.as<bf16_t>(), L2);
  }
}

void WarmNet::head(hipStream_t st, const float* raw, int mode, float* out) {
  if (e1_kept)
    net_head(st, mode, raw, inv_s.as<float>(), e1.as<bf16_t>(), d2.as<bf16_t>(), P.as<float>() + head_off, nullptr,
             out, nullptr, L1);
  else
    net_first(st, 2 + mode, raw, inv_s.as<float>(), wf[0].as<bf16_t>(), P.as<float>() + b_off[0], nullptr,
              d2.as<bf16_t>(), P.as<float>() + head_off, out, L1);
}

void WarmNet::backward(hipStream_t st, const float* raw, const float* tgt, float* loss_slot) {
  const float* p = P.as<float>();
  float* g = Gr.as<float>();
  const float* wh = p + head_off;
  const float* dO = dout.as<float>();
  const int hb = net_head_blocks(L1);
  net_head(st, 2, raw, inv_s.as<float>(), e1.as<bf16_t>(), d2.as<bf16_t>(), wh, tgt, dout.as<float>(),
           hpart.as<float>(), L1);
  net_head_reduce(st, hpart.as<float>(), hb, 1.0 / (double)vox(L1), g, head_off, loss_slot);

  auto wgrad = [&](int l, const bf16_t* in, const bf16_t* gr, const Vol& v) {
    net_wgrad(st, NET_CIN[l], NET_COUT[l], in, gr, v, WG_CHUNKS, wpart.as<float>());
    net_reduce_rows(st, wpart.as<float>(), WG_CHUNKS, 27L * NET_CIN[l] * NET_COUT[l], g + w_off[l]);
    net_chan_sum(st, NET_COUT[l], gr, vox(v), CS_BLOCKS, cpart.as<float>());
    net_reduce_rows(st, cpart.as<float>(), CS_BLOCKS, NET_COUT[l], g + b_off[l]);
  };
  // conv5 (a2 -> d2)
  net_grad_dec(st, 16, d2.as<bf16_t>(), nullptr, dO, wh, g5.as<bf16_t>(), L2);
  wgrad(4, a2.as<bf16_t>(), g5.as<bf16_t>(), L2);
  net_conv(st, 16, 32, false, g5.as<bf16_t>(), wb[4].as<bf16_t>(), nullptr, gA2.as<bf16_t>(), L2);
  // conv4 (e3 -> d3)
  net_grad_dec(st, 32, d3.as<bf16_t>(), gA2.as<bf16_t>(), nullptr, nullptr, g4.as<bf16_t>(), L3);
  wgrad(3, e3.as<bf16_t>(), g4.as<bf16_t>(), L3);
  net_conv(st, 32, 64, false, g4.as<bf16_t>(), wb[3].as<bf16_t>(), nullptr, gE3.as<bf16_t>(), L3);
  // conv3 (p2 -> e3)
  net_grad_enc(st, 64, e3.as<bf16_t>(), gE3.as<bf16_t>(), nullptr, nullptr, nullptr, g3.as<bf16_t>(), L3);
  wgrad(2, p2.as<bf16_t>(), g3.as<bf16_t>(), L3);
  net_conv(st, 64, 32, false, g3.as<bf16_t>(), wb[2].as<bf16_t>(), nullptr, gP2.as<bf16_t>(), L3);
  // conv2 (p1 -> e2); e2 also feeds a2 directly
  net_grad_enc(st, 32, e2.as<bf16_t>(), gA2.as<bf16_t>(), nullptr, nullptr, gP2.as<bf16_t>(), g2.as<bf16_t>(), L2);
  wgrad(1, p1.as<bf16_t>(), g2.as<bf16_t>(), L2);
  net_conv(st, 32, 16, false, g2.as<bf16_t>(), wb[1].as<bf16_t>(), nullptr, gP1.as<bf16_t>(), L2);
  // conv1 (u -> e1); e1 also feeds the head directly
  net_grad_enc(st, 16, e1.as<bf16_t>(), nullptr, dO, wh, gP1.as<bf16_t>(), g1.as<bf16_t>(), L1);
  net_wgrad_in1(st, raw, inv_s.as<float>(), g1.as<bf16_t>(), L1, WG_CHUNKS_IN1, wpart.as<float>());
  net_reduce_rows(st, wpart.as<float>(), WG_CHUNKS_IN1, 27L * 16, g + w_off[0]);
  net_chan_sum(st, 16, g1.as<bf16_t>(), vox(L1), CS_BLOCKS, cpart.as<float>());
  net_reduce_rows(st, cpart.as<float>(), CS_BLOCKS, 16, g + b_off[0]);
}

void WarmNet::adam(hipStream_t st, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  ++adam_t;
  const float lr_t = lr / (float)(1.0 - std::pow((double)b1, adam_t));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, adam_t)));
  net_adam(st, P.as<float>(), Gr.as<float>(), M.as<float>(), V.as<float>(), nparam, lr_t, b1, b2, eps, c2);
  pack(st);
}

long WarmNet::fetch(int which, float* host, hipStream_t st) {
  struct T {
    const DevBuf* b;
    const Vol* v;
    int c;
  };
  const T tab[] = {{&e1, &L1, 16}, {&p1, &L2, 16}, {&e2, &L2, 32}, {&p2, &L3, 32}, {&e3, &L3, 64},
                   {&d3, &L3, 32}, {&a2, &L2, 32}, {&d2, &L2, 16}, {&dout, &L1, 1}, {&g5, &L2, 16},
                   {&gA2, &L2, 32}, {&g4, &L3, 32}, {&gE3, &L3, 64}, {&g3, &L3, 64}, {&gP2, &L3, 32},
                   {&g2, &L2, 32}, {&gP1, &L2, 16}, {&g1, &L1, 16}};
  if (which < 0 || which >= (int)(sizeof(tab) / sizeof(tab[0]))) fail("cudaDecon_net_fetch: unknown tensor");
  if ((which == 0 || which == 6) && !e1_kept)
    fail("cudaDecon_net_fetch: e1 and a2 are not stored by the inference path");
  const T& t = tab[which];
  const long n = vox(*t.v) * t.c;
  if (!t.b->p || t.b->bytes < (size_t)n * (which == 8 ? 4 : 2)) fail("cudaDecon_net_fetch: tensor not computed");
  if (!host) return n;
  if (which == 8) {
    CDN_HIP(hipMemcpyAsync(host, t.b->p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, st));
    CDN_HIP(hipStreamSynchronize(st));
    return n;
  }
  std::vector<bf16_t> tmp((size_t)n);
  CDN_HIP(hipMemcpyAsync(tmp.data(), t.b->p, sizeof(bf16_t) * (size_t)n, hipMemcpyDeviceToHost, st));
  CDN_HIP(hipStreamSynchronize(st));
  for (long i = 0; i < n; ++i) {
    const uint32_t u = (uint32_t)tmp[(size_t)i] << 16;
    float f;
    std::memcpy(&f, &u, 4);
    host[i] = f;
  }
  return n;
}

void WarmNet::add_samples(const float* raw, const float* tgt, long count, int nz, int ny, int nx, hipStream_t st) {
  if (count <= 0) return;
  if (ds_count > 0 && (nz != ds_nz || ny != ds_ny || nx != ds_nx))
    fail("cudaDecon_net_add_samples: every sample must have the shape of the first");
  const size_t per = (size_t)nz * ny * nx;
  const size_t old = (size_t)ds_count * per, add = (size_t)count * per;
  // grow both stores (keeping what is there)
  for (DevBuf* b : {&ds_raw, &ds_tgt}) {
    if (b->bytes < sizeof(float) * (old + add)) {
      void* np = nullptr;
      const size_t cap = std::max(sizeof(float) * (old + add), b->bytes * 2);
      CDN_HIP(hipMalloc(&np, cap));
      try {
        if (old) CDN_HIP(hipMemcpyAsync(np, b->p, sizeof(float) * old, hipMemcpyDeviceToDevice, st));
        CDN_HIP(hipStreamSynchronize(st));
      } catch (...) {
        (void)hipFree(np);
        throw;
      }
      b->release();
      b->p = np;
      b->base = np;
      b->bytes = cap;
    }
  }
  CDN_HIP(hipMemcpyAsync(ds_raw.as<float>() + old, raw, sizeof(float) * add, hipMemcpyHostToDevice, st));
  CDN_HIP(hipMemcpyAsync(ds_tgt.as<float>() + old, tgt, sizeof(float) * add, hipMemcpyHostToDevice, st));
  CDN_HIP(hipStreamSynchronize(st));
  ds_count += count;
  ds_nz = nz;
  ds_ny = ny;
  ds_nx = nx;
}

void WarmNet::train(int steps, int batch, float lr, uint64_t seed, hipStream_t st, float* losses) {
  if (ds_count <= 0) fail("cudaDecon_net_train: empty training set");
  if (batch < 1 || batch > ds_count) fail("cudaDecon_net_train: batch must be in [1, training-set size]");
  geometry(batch, ds_nz, ds_ny, ds_nx, true);
  const size_t per = (size_t)ds_nz * ds_ny * ds_nx;
  b_raw.ensure(sizeof(float) * per * batch);
  b_tgt.ensure(sizeof(float) * per * batch);
  DevBuf lossv;
  lossv.ensure(sizeof(float) * (size_t)std::max(steps, 1));
  SplitMix rng{seed};
  std::vector<long> perm((size_t)ds_count);
  for (long i = 0; i < ds_count; ++i) perm[(size_t)i] = i;
  long pos = ds_count;  // forces a shuffle at the first step
  for (int s = 0; s < steps; ++s) {
    for (int b = 0; b < batch; ++b) {
      if (pos >= ds_count) {
        for (long i = ds_count - 1; i > 0; --i) std::swap(perm[(size_t)i], perm[(size_t)(rng.next() % (uint64_t)(i + 1))]);
        pos = 0;
      }
      const long idx = perm[(size_t)pos++];
      CDN_HIP(hipMemcpyAsync(b_raw.as<float>() + per * b, ds_raw.as<float>() + per * idx, sizeof(float) * per,
                             hipMemcpyDeviceToDevice, st));
      CDN_HIP(hipMemcpyAsync(b_tgt.as<float>() + per * b, ds_tgt.as<float>() + per * idx, sizeof(float) * per,
                             hipMemcpyDeviceToDevice, st));
    }
    forward(st, b_raw.as<float>());
    backward(st, b_raw.as<float>(), b_tgt.as<float>(), lossv.as<float>() + s);
    adam(st, lr);
  }
  if (losses && steps > 0)
    CDN_HIP(hipMemcpyAsync(losses, lossv.p, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost, st));
  CDN_HIP(hipStreamSynchronize(st));
}

}  // namespace cdn
