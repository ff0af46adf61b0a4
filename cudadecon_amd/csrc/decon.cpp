// Host orchestration of the RL-Biggs-Andrews iteration (restates
// src/RL-Biggs-Andrews.cpp:211-367 as 9 fused passes + a device-side
// acceleration factor; no per-iteration host synchronisation).
#include "decon.hpp"
#include "hostcopy.hpp"

#include "rl_kernels.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace cdn {

// Large buffers start a multiple of CDN_STAGGER bytes into their
// allocation, the multiple cycling 0..7 over successive allocations, so that
// the volumes a pass reads and writes at the same index (X, X_prev, Y, G,
// raw, spectrum) do not all share the same offset from a 2 MB boundary.
// Measured at 512 x 2048 x 2048 over three processes each: iteration
// 44.37-44.43 ms with 256 bytes against 44.41-45.32 ms without (the x passes
// varied by up to 7% between processes; the column passes, which touch one
// volume, did not) and 44.30-45.95 ms with 4 KB; at 256 x 1024 x 1024
// 4.47-4.55 against 4.54-4.57 ms.
#ifndef CDN_STAGGER
#define CDN_STAGGER 256
#endif
void DevBuf::ensure(size_t b) {
  if (b <= bytes && p) return;
  if (!owned && p)
    fail("a caller-provided device buffer of " + std::to_string(bytes) + " bytes is too small for " +
         std::to_string(b));
  release();
  if (b == 0) return;
  size_t off = 0;
  if (CDN_STAGGER > 0 && b >= (64u << 20)) {
    static std::atomic<unsigned> k{0};
    off = (size_t)(k++ % 8) * CDN_STAGGER;
  }
  CDN_HIP(hipMalloc(&base, b + off));
  p = static_cast<char*>(base) + off;
  bytes = b;
}

void DevBuf::bind(void* q, size_t b) {
  release();
  p = q;
  base = nullptr;
  bytes = b;
  owned = false;
}

void DevBuf::release() {
  if (p && owned) (void)hipFree(base);
  p = nullptr;
  base = nullptr;
  bytes = 0;
  owned = true;
}

// ---------------------------------------------------------------------------
// Fft3
// ---------------------------------------------------------------------------
void* DevRegion::take(size_t b) {
  const size_t off = (used + 255) / 256 * 256;
  if (off > size || b > size - off)
    fail("device region of " + std::to_string(size) + " bytes is too small (" + std::to_string(off + b) +
         " needed)");
  used = off + b;
  return base + off;
}

// a buffer of b bytes: allocated, or the region's next piece
static void obtain(DevBuf& buf, size_t b, DevRegion* region) {
  if (region)
    region->give(buf, b);
  else
    buf.ensure(b);
}

static void upload_twiddles(DevBuf& b, int n, hipStream_t st, DevRegion* region) {
  const int cnt = fft_twiddle_count(n);
  std::vector<float2> h(cnt);
  fft_twiddles(n, h.data());
  obtain(b, sizeof(float2) * cnt, region);
  CDN_HIP(hipMemcpyAsync(b.p, h.data(), sizeof(float2) * cnt, hipMemcpyHostToDevice, st));
  CDN_HIP(hipStreamSynchronize(st));
}

void Fft3::setup(int x, int y, int z, hipStream_t st, DevRegion* region) {
  for (int n : {x, y, z}) {
    if (!fft_length_supported(n))
      fail("FFT length " + std::to_string(n) + " is not supported (2,3,5,7-smooth, at most 16384)");
    fft_prepare_length(n);
  }
  nx = x;
  ny = y;
  nz = z;
  nxh = nx / 2 + 1;
  pitch = pitch_of(nx);
  nvox = (long)nx * ny * nz;
  nspec = (long)nz * ny * pitch;
  upload_twiddles(twx, nx, st, region);
  upload_twiddles(twy, ny, st, region);
  upload_twiddles(twz, nz, st, region);
  fft_prepare_cols(ny, nxh);
  fft_prepare_cols(nz, nxh);
}

void Fft3::col(hipStream_t st, float2* S, int axis, bool inverse) const {
  if (axis == 1)
    launch_col_fft(ny, inverse, st, S, nxh, pitch, (long)ny * pitch, nz, twy.as<float2>());
  else
    launch_col_fft(nz, inverse, st, S, nxh, (long)ny * pitch, pitch, ny, twz.as<float2>());
}

void Fft3::r2c(hipStream_t st, const float* in, float2* S) const {
  launch_x_r2c(nx, st, in, nullptr, nullptr, nullptr, XIN_PLAIN, S, (long)ny * nz, pitch, twx.as<float2>());
  col(st, S, 1, false);
  col(st, S, 2, false);
}

void Fft3::c2r(hipStream_t st, float2* S, float* out) const {
  col(st, S, 2, true);
  col(st, S, 1, true);
  launch_x_c2r(nx, st, S, out, (long)ny * nz, pitch, twx.as<float2>());
}

// ---------------------------------------------------------------------------
// OTF
// ---------------------------------------------------------------------------
void OtfDev::setup(const OtfTable& t, const float kscale[3], int nz, hipStream_t st, DevRegion* region) {
  ra = (t.nyotf == 1);
  ks[0] = kscale[0];
  ks[1] = kscale[1];
  ks[2] = kscale[2];
  nz_for = nz;
  if (ra) {
    // z-interpolated radial table: tz[r][kz] = T[r][iz](1-az) + T[r][iz1] az
    // for the grid's signed kz, as in dev_otfinterpolate's RA branch
    // (src/RLgpuImpl.cu:234-272); zero where kzindex is outside the table.
    const int nr = t.nxotf, nzo = t.nzotf;
    std::vector<float2> h((size_t)nr * nz);
    for (int kz = 0; kz < nz; ++kz) {
      const int kzi = kz > (nz >> 1) ? kz - nz : kz;
      const float kzs = (float)kzi * ks[2];
      const float kzindex = kzs < 0.f ? kzs + (float)nzo : kzs;
      const bool ok = kzindex < (float)nzo && kzindex >= 0.f;
      const int iz = ok ? (int)std::floor(kzindex) : 0;
      const float az = kzindex - (float)iz;
      const int iz1 = (iz == nzo - 1) ? 0 : iz + 1;
      for (int r = 0; r < nr; ++r) {
        float2 v = make_float2(0.f, 0.f);
        if (ok) {
          const float* a = &t.data[2 * ((size_t)r * nzo + iz)];
          const float* b = &t.data[2 * ((size_t)r * nzo + iz1)];
          v.x = a[0] * (1 - az) + b[0] * az;
          v.y = a[1] * (1 - az) + b[1] * az;
        }
        h[(size_t)r * nz + kz] = v;
      }
    }
    obtain(tz, sizeof(float2) * h.size(), region);
    CDN_HIP(hipMemcpyAsync(tz.p, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice, st));
    // the two rows a radius interpolates between, side by side
    if (nr >= 2) {
      std::vector<float4> h2((size_t)(nr - 1) * nz);
      for (int r = 0; r + 1 < nr; ++r)
        for (int kz = 0; kz < nz; ++kz) {
          const float2 a = h[(size_t)r * nz + kz], b = h[(size_t)(r + 1) * nz + kz];
          h2[(size_t)r * nz + kz] = make_float4(a.x, a.y, b.x, b.y);
        }
      obtain(tz2, sizeof(float4) * h2.size(), region);
      CDN_HIP(hipMemcpyAsync(tz2.p, h2.data(), sizeof(float4) * h2.size(), hipMemcpyHostToDevice, st));
    }
  } else {
    obtain(raw, sizeof(float) * t.data.size(), region);
    CDN_HIP(hipMemcpyAsync(raw.p, t.data.data(), sizeof(float) * t.data.size(), hipMemcpyHostToDevice, st));
  }
  CDN_HIP(hipStreamSynchronize(st));
}

OtfParams OtfDev::params(const OtfTable& t, double inv_n) const {
  OtfParams p{};
  p.ra = ra;
  p.ra_tab.tz = tz.as<float2>();
  p.ra_tab.tz2 = tz2.as<float4>();
  p.ra_tab.nr = t.nxotf;
  p.tab3.raw = raw.as<float2>();
  p.tab3.nxotf = t.nxotf;
  p.tab3.nyotf = t.nyotf;
  p.tab3.nzotf = t.nzotf;
  p.kxscale = ks[0];
  p.kyscale = ks[1];
  p.kzscale = ks[2];
  p.inv_n = (float)inv_n;
  return p;
}

// ---------------------------------------------------------------------------
// RL engine
// ---------------------------------------------------------------------------
// Stacks can share every pass only if no x-update workgroup spans two of
// them (each stack's dot products then come from its own workgroups, in the
// order a one-stack engine sums them).
static void check_stack_rows(int nx, long stack_rows, int nstacks) {
  if (nstacks > 1 && stack_rows % x_update_rows_per_block(nx) != 0)
    fail("rows per stack (ny nz = " + std::to_string(stack_rows) + ") must be a multiple of " +
         std::to_string(x_update_rows_per_block(nx)) + " for several stacks in one engine");
}

void RLEngine::setup(int nx, int ny, int nz, const OtfTable& t, const float kscale[3], hipStream_t st,
                     int nst, const EngineViews* views, DevRegion* region) {
  if (nst < 1) fail("an engine needs at least one stack");
  check_stack_rows(nx, (long)ny * nz, nst);
  if (region || !fft.matches(nx, ny, nz)) fft.setup(nx, ny, nz, st, region);
  nstacks = nst;
  table = t;
  otf.setup(table, kscale, nz, st, region);
  const size_t vb = sizeof(float) * (size_t)fft.nvox * nstacks;
  const size_t sb = sizeof(float2) * (size_t)fft.nspec * nstacks;
  if (views) {
    raw.bind(views->raw, vb);
    xa.bind(views->xa, vb);
    xb.bind(views->xb, vb);
    y.bind(views->y, vb);
    g.bind(views->g, vb);
    spec.bind(views->spec, sb);
  } else {
    for (DevBuf* b : {&raw, &xa, &xb, &y, &g})
      if (!b->owned) b->release();  // an earlier setup's views
    if (!spec.owned) spec.release();
    raw.ensure(vb);
    xa.ensure(vb);
    xb.ensure(vb);
    y.ensure(vb);
    g.ensure(vb);
    spec.ensure(sb);
  }
  nb_part = x_update_blocks(nx, (long)ny * nz);
  // room for any x-pass launch shape the autotuner may install later
  obtain(part, sizeof(double) * 2 * (size_t)x_update_blocks_max((long)ny * nz) * nstacks, region);
  cur = xa.as<float>();
  prev = xb.as<float>();
}

void RLEngine::load_raw(const float* in, bool on_device, hipStream_t st) {
  const size_t vb = sizeof(float) * (size_t)total_vox();
  if (on_device)
    copy_f32(st, in, raw.as<float>(), (size_t)total_vox());
  else
    copy_to_device(st, raw.p, in, vb);
  copy_f32(st, raw.as<float>(), xa.as<float>(), (size_t)total_vox());
  cur = xa.as<float>();
  prev = xb.as<float>();
}

void RLEngine::load_stack(int s, const float* in, bool on_device, hipStream_t st) {
  if (s < 0 || s >= nstacks) fail("stack index out of range");
  const size_t vb = sizeof(float) * (size_t)fft.nvox, off = (size_t)fft.nvox * s;
  if (on_device)
    copy_f32(st, in, raw.as<float>() + off, (size_t)fft.nvox);
  else
    copy_to_device(st, raw.as<float>() + off, in, vb);
  copy_f32(st, raw.as<float>() + off, cur + off, (size_t)fft.nvox);  // its X_0
}

void RLEngine::reset(hipStream_t st) {
  copy_f32(st, raw.as<float>(), xa.as<float>(), (size_t)total_vox());
  cur = xa.as<float>();
  prev = xb.as<float>();
}

void RLEngine::bind_lambdas(DevRegion& region, int n) {
  region.give(lam, sizeof(double) * 4 * (size_t)nstacks * (size_t)std::max(n, 1));
}

void RLEngine::prepare(int n) { lam.ensure(sizeof(double) * 4 * (size_t)nstacks * (size_t)std::max(n, 1)); }

std::vector<const void*> RLEngine::device_addresses() const {
  return {raw.p, xa.p, xb.p, y.p, g.p, spec.p, part.p, lam.p, otf.raw.p, otf.tz.p, otf.tz2.p,
          fft.twx.p, fft.twy.p, fft.twz.p};
}

void RLEngine::load_seed(const float* in, bool on_device, hipStream_t st) {
  const size_t vb = sizeof(float) * (size_t)total_vox();
  CDN_HIP(hipMemcpyAsync(y.p, in, vb, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
}

void RLEngine::iterate(int n, int k0, bool seeded, hipStream_t st, double* lambdas) {
  if (n <= 0) return;
  const long stack_rows = (long)fft.ny * fft.nz, rows = stack_rows * nstacks;
  check_stack_rows(fft.nx, stack_rows, nstacks);  // an autotuned x shape may have changed the blocks
  const float eps = FLT_EPSILON;  // src/RL-Biggs-Andrews.cpp:274
  const OtfParams op = otf.params(table, 1.0 / (double)fft.nvox);
  lam.ensure(sizeof(double) * 4 * (size_t)nstacks * (size_t)n);
  float2* S = spec.as<float2>();
  const float2* twx = fft.twx.as<float2>();
  const float2* twz = fft.twz.as<float2>();
  double* lamd = lam.as<double>();

  std::vector<hipEvent_t> ev;
  std::vector<int> evpass;
  auto mark = [&](int pass) {
    if (!timing) return;
    hipEvent_t e;
    CDN_HIP(hipEventCreate(&e));
    CDN_HIP(hipEventRecord(e, st));
    ev.push_back(e);
    evpass.push_back(pass);
  };
  mark(-1);
  for (int j = 0; j < n; ++j) {
    const int k = k0 + j;
    double* lamk = lamd + 4 * (size_t)nstacks * j;
    int mode;
    const float* yv;
    if (k >= 2) {
      accel_factor(st, part.as<double>(), nb_part, (double)eps, lamk, nstacks);
      mark(PASS_LAMBDA);
      mode = XIN_PRED;
      yv = y.as<float>();
    } else if (k == 0 && seeded) {
      mode = XIN_SEED;
      yv = y.as<float>();
    } else {
      mode = XIN_PLAIN;
      yv = cur;
    }
    launch_x_r2c(fft.nx, st, cur, prev, y.as<float>(), lamk, mode, S, rows, fft.pitch, twx, stack_rows);
    mark(PASS_XR2C);
    for (int f = 0; f < 2; ++f) {
      // the y transforms see the stacks as nz nstacks slabs
      launch_col_fft(fft.ny, 0, st, S, fft.nxh, fft.pitch, (long)fft.ny * fft.pitch, (long)fft.nz * nstacks,
                     fft.twy.as<float2>());
      mark(PASS_YFWD);
      launch_z_filter(fft.nz, st, S, fft.nxh, (long)fft.ny * fft.pitch, fft.pitch, fft.ny, f, op, twz, nstacks,
                      fft.nspec);
      mark(PASS_ZFILT);
      launch_col_fft(fft.ny, 1, st, S, fft.nxh, fft.pitch, (long)fft.ny * fft.pitch, (long)fft.nz * nstacks,
                     fft.twy.as<float2>());
      mark(PASS_YINV);
      if (f == 0) {
        launch_x_lrcore(fft.nx, st, S, raw.as<float>(), rows, fft.pitch, twx, eps);
        mark(PASS_XLR);
      }
    }
    nb_part = launch_x_update(fft.nx, st, S, yv, prev, g.as<float>(), part.as<double>(), k >= 1, rows, fft.pitch,
                              twx) /
              nstacks;
    mark(PASS_XUPD);
    std::swap(cur, prev);
  }
  if (timing || lambdas) CDN_HIP(hipStreamSynchronize(st));
  if (timing) {
    for (size_t i = 1; i < ev.size(); ++i) {
      float ms = 0.f;
      CDN_HIP(hipEventElapsedTime(&ms, ev[i - 1], ev[i]));
      pass_ms[evpass[i]] += ms;
      pass_calls[evpass[i]] += 1;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  if (lambdas) {
    std::vector<double> h(4 * (size_t)nstacks * n);
    CDN_HIP(hipMemcpy(h.data(), lam.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int j = 0; j < n; ++j)
      for (int s = 0; s < nstacks; ++s)
        lambdas[(size_t)j * nstacks + s] = (k0 + j >= 2) ? h[4 * ((size_t)j * nstacks + s)] : 0.0;
  }
}

}  // namespace cdn
