// Host orchestration of the RL-Biggs-Andrews path on one GPU.
#pragma once

#include "common.hpp"
#include "fft_kernels.hpp"
#include "geometry.hpp"

#include <string>
#include <vector>

namespace cdn {

// Device buffer: an owned allocation that only grows (kept across calls;
// freed by the owner's destructor or release()), or a view of memory the
// caller owns (bind), which is never freed here and never grows.
struct DevBuf {
  void* p = nullptr;
  void* base = nullptr;  // the allocation (p may start a stagger offset into it)
  size_t bytes = 0;
  bool owned = true;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), base(o.base), bytes(o.bytes), owned(o.owned) {
    o.p = nullptr;
    o.base = nullptr;
    o.bytes = 0;
    o.owned = true;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      base = o.base;
      bytes = o.bytes;
      owned = o.owned;
      o.p = nullptr;
      o.base = nullptr;
      o.bytes = 0;
      o.owned = true;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void ensure(size_t b);         // a view must already hold b bytes
  void bind(void* q, size_t b);  // view q[0, b)
  void release();
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Pieces of one caller-provided device region, 256-byte aligned, handed
// out in order (a bump allocator; nothing is ever freed back).
struct DevRegion {
  char* base = nullptr;
  size_t size = 0, used = 0;
  void* take(size_t b);
  // bind buf to the next b bytes
  void give(DevBuf& buf, size_t b) { buf.bind(take(b), b); }
};

// Where an engine's volumes and spectrum live when the caller provides them
// (each nstacks stacks long): raw, the X ping-pong, Y, G, the spectrum.
struct EngineViews {
  float *raw, *xa, *xb, *y, *g;
  float2* spec;
};

// 3-D R2C / C2R plan for one (nx, ny, nz) grid.
struct Fft3 {
  int nx = 0, ny = 0, nz = 0, nxh = 0;
  long pitch = 0;          // half-spectrum row pitch (complex elements)
  long nvox = 0;           // nx ny nz
  long nspec = 0;          // nz ny pitch (allocated complex elements)
  DevBuf twx, twy, twz;
  bool matches(int x, int y, int z) const { return nx == x && ny == y && nz == z; }
  // row pitch (complex elements) of the half spectrum of an x length
  static long pitch_of(int x) { return (x / 2 + 1 + 15) / 16 * 16; }
  // twiddle tables allocated, or taken from `region` when given
  void setup(int x, int y, int z, hipStream_t st, DevRegion* region = nullptr);
  void r2c(hipStream_t st, const float* in, float2* S) const;   // unnormalised
  void c2r(hipStream_t st, float2* S, float* out) const;        // unnormalised, consumes S
  void col(hipStream_t st, float2* S, int axis, bool inverse) const;  // axis 1 = y, 2 = z
};

struct OtfDev {
  int ra = 1, nz_for = -1;
  float ks[3] = {0, 0, 0};
  DevBuf raw, tz, tz2;
  void setup(const OtfTable& t, const float kscale[3], int nz, hipStream_t st, DevRegion* region = nullptr);
  OtfParams params(const OtfTable& t, double inv_n) const;
};

enum Pass { PASS_XR2C = 0, PASS_YFWD, PASS_ZFILT, PASS_YINV, PASS_XLR, PASS_XUPD, PASS_LAMBDA, PASS_COUNT };

// RL-BA engine: all state resident in HBM; the acceleration factor stays on
// the device so an iteration issues no host synchronisation.  It holds
// `nstacks` stacks of one shape back to back (stack s at s nvox in the real
// volumes, s nspec in the spectrum) and runs every pass once per iteration
// over all of them; each stack keeps its own acceleration factor, and its
// results are those of a one-stack engine.
struct RLEngine {
  Fft3 fft;
  OtfDev otf;
  OtfTable table;
  DevBuf raw, xa, xb, y, g, spec, part, lam;
  float* cur = nullptr;   // X_k after the last iteration
  float* prev = nullptr;  // X_{k-1}
  int nstacks = 1;
  int nb_part = 0;        // dot-product partials per stack
  // per-pass timing (ms, summed) when enabled
  bool timing = false;
  double pass_ms[PASS_COUNT] = {};
  long pass_calls[PASS_COUNT] = {};

  // views: the volumes and spectrum live there; region: every other device
  // buffer (twiddles, OTF tables, partials) is taken from it, nothing is
  // allocated (then call bind_lambdas before iterating)
  void setup(int nx, int ny, int nz, const OtfTable& t, const float kscale[3], hipStream_t st, int nstacks = 1,
             const EngineViews* views = nullptr, DevRegion* region = nullptr);
  // the factor slots of up to n iterations, from the region
  void bind_lambdas(DevRegion& region, int n);
  long stack_vox() const { return fft.nvox; }
  long total_vox() const { return fft.nvox * nstacks; }
  // Stage the (preprocessed) stacks: R := in, X_0 := in (all stacks, stack
  // after stack).  `in` is a device pointer when on_device, else host.
  void load_raw(const float* in, bool on_device, hipStream_t st);
  // stage stack s only
  void load_stack(int s, const float* in, bool on_device, hipStream_t st);
  void load_seed(const float* in, bool on_device, hipStream_t st);
  // X_0 := R again (restart from the staged stack)
  void reset(hipStream_t st);
  // size every buffer for calls of up to n iterations (no allocation inside
  // iterate afterwards, e.g. during stream capture)
  void prepare(int n);
  // Run iterations k = k0 .. k0+n-1 continuing the resident state (k0 > 0
  // requires the previous call to have ended at k0-1); seeded => Y_0 from the
  // seed buffer.  lambdas (optional, host, n x nstacks entries, stack
  // fastest) receives the factor used at each k (0 for k < 2).
  void iterate(int n, int k0, bool seeded, hipStream_t st, double* lambdas);
  float* result() const { return cur; }
  // every device address an iteration reads or writes (buffers, tables,
  // twiddles): a captured graph of the iteration stays valid while these
  // are unchanged
  std::vector<const void*> device_addresses() const;
};

}  // namespace cdn
