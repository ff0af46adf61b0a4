// extern "C" boundary of libcudaDecon.so (declarations in include/cudaDecon.h).
//
// Session state mirrors the reference library's process globals
// (src/RL-Biggs-Andrews.cpp:45-47, 459-466).  Every entry point catches all
// exceptions and returns 0 (or a neutral value) with the message kept for
// cudaDecon_last_error().
#include "../../include/cudaDecon.h"

#include "batch.hpp"
#include "decon.hpp"
#include "geometry.hpp"
#include "hostcopy.hpp"
#include "net.hpp"
#include "noalloc.hpp"
#include "rl_kernels.hpp"
#include "tiff.hpp"
#include "tune.hpp"

#include <cfloat>
#include <cmath>
#include <algorithm>
#include <array>
#include <cstring>
#include <memory>
#include <string>

namespace cdn {
namespace {

thread_local std::string g_err;

// the session's sizes and matrices (SessionGeometry) plus its device state
struct Session : SessionGeometry {
  hipStream_t st = nullptr;
  OtfTable otf;
  RLEngine eng;
  DevBuf u16, bufA, bufB, red_part, red_acc, red_mean;
  ~Session() {
    if (st) (void)hipStreamDestroy(st);
  }
};

std::unique_ptr<Session> g_s;
// bleach-correction statics (src/RL-Biggs-Andrews.cpp:46-47)
double g_intensity0 = 0.0;
bool g_first = true;
// output_nz survives RL_cleanup in the reference and feeds Deskew_interface
// (src/RL-Biggs-Andrews.cpp:459,741)
unsigned g_output_nz = 0;
// flat start of RL_interface: Y_0 filled with a value (the CLI's FlatStart,
// src/linearDecon.cpp:1143-1145 and src/RL-Biggs-Andrews.cpp:305-316) until
// RL_cleanup
bool g_flat_on = false;
float g_flat_value = 0.f;

hipStream_t make_stream() {
  hipStream_t s;
  CDN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

void kscales(const Session& S, unsigned fft_nx, float ks[3]) { session_kscales(S, S.otf, fft_nx, ks); }

// camera-correction parameters of camcor_interface_init (process global, as
// the reference's intended context, src/llspy.cpp:70-77)
struct Camcor {
  hipStream_t st = nullptr;
  int nx = 0, ny = 0, nz = 0;
  DevBuf cam, in, out;
};
Camcor& camcor_state() {
  static Camcor c;
  return c;
}

template <class F> int guard(F&& f) {
  try {
    g_err.clear();
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
  } catch (...) {
    g_err = "unknown error";
  }
  return 0;
}

// A NaN or Inf anywhere in a stack spreads through the FFT to every voxel of
// the result, so float inputs of the RL path are checked on the device
// (one read of the staged stack) and rejected with a message.
void check_finite(hipStream_t st, const float* dev, long n, DevBuf& red, const std::string& what) {
  red.ensure(sizeof(double) * 3 * 1024);
  const long bad = count_nonfinite(st, dev, n, red.as<double>());
  if (bad > 0) fail(what + ": " + std::to_string(bad) + " NaN or Inf value(s) in the stack");
}

}  // namespace
}  // namespace cdn

using namespace cdn;

struct cudaDecon_batch {
  Batch b;
  DevBuf red;  // partials of the input checks
};

struct cudaDecon_net {
  hipStream_t st = nullptr;
  WarmNet net;
  ~cudaDecon_net() {
    if (st) (void)hipStreamDestroy(st);
  }
};

struct cudaDecon_rl {
  hipStream_t st = nullptr;
  RLEngine eng;
  int k_next = 0;
  DevBuf red, ref;  // reduction partials of the diagnostics (not the engine's), snapshot of X
  DevBuf probe;     // destination of the copy-rate probe (never an engine buffer)
  DevBuf u16;       // staging of uint16 stacks
  DevBuf post[2];   // device-resident deskew / affine results
  int post_cur = -1, pnx = 0, pny = 0, pnz = 0;
  ~cudaDecon_rl() {
    if (st) (void)hipStreamDestroy(st);
  }
};

// entry points defined on one stack (the engine's other stacks would be ignored)
static void one_stack(const cudaDecon_rl* h, const char* what) {
  if (h && h->eng.nstacks != 1) fail(std::string(what) + ": defined for one-stack engines only");
}

extern "C" {

const char* cudaDecon_last_error(void) { return g_err.c_str(); }

const char* get_version(void) { return "0.7.0"; }

int cudaDecon_fft_supported(int n) { return fft_length_supported(n) ? 1 : 0; }

unsigned cudaDecon_find_optimal_dimension(unsigned n, int step) { return find_optimal_dimension(n, step); }

unsigned cudaDecon_deskew_width(int nx, float dz, float dr, float deskewAngle, int outputWidth) {
  return deskew_interface_width(nx, dz, dr, deskewAngle, outputWidth, g_output_nz);
}

int RL_interface_init(int nx, int ny, int nz, float dr, float dz, float dr_psf, float dz_psf, float deskewAngle,
                      float rotationAngle, int outputWidth, bool bSkewedDecon, char* OTF_file_name) {
  return guard([&]() -> int {
    if (!OTF_file_name) fail("RL_interface_init: OTF file name is null");
    auto S = std::make_unique<Session>();
    try {
      static_cast<SessionGeometry&>(*S) =
          session_geometry(nx, ny, nz, dr, dz, deskewAngle, rotationAngle, outputWidth, bSkewedDecon);
    } catch (const std::exception& e) {
      fail(std::string("RL_interface_init: ") + e.what());
    }
    if (bSkewedDecon) dz_psf *= std::fabs(std::sin(deskewAngle * M_PI / 180.));
    S->otf = load_otf_file(OTF_file_name, dr_psf, dz_psf);
    // precompiled lengths need nothing; others are compiled now (like a
    // cuFFT plan, src/RL-Biggs-Andrews.cpp:622-631), without touching the GPU
    for (unsigned n : {S->fft_nx, S->output_ny, S->output_nz}) {
      if (!fft_length_supported((int)n))
        fail("RL_interface_init: FFT length " + std::to_string(n) + " is not supported (2,3,5,7-smooth, at most 16384)");
      fft_prepare_length((int)n);
    }
    g_output_nz = S->output_nz;  // the stream is created on first use
    g_s = std::move(S);
    return 1;
  });
}

unsigned get_output_nx(void) {
  if (!g_s) return 0;
  if (g_s->rot_on) {
    int nxo, nzo;
    rotated_dims(g_s->rot, g_s->deskewedXdim, g_s->output_nz, &nxo, &nzo);
    return (unsigned)nxo;
  }
  return g_s->deskewedXdim;
}
unsigned get_output_ny(void) { return g_s ? g_s->output_ny : 0; }

int cudaDecon_get_session_kscale(float* ks3) {
  return guard([&]() -> int {
    if (!g_s) fail("cudaDecon_get_session_kscale: no session");
    if (!ks3) fail("cudaDecon_get_session_kscale: null argument");
    kscales(*g_s, g_s->fft_nx, ks3);
    return 1;
  });
}

unsigned cudaDecon_get_session_dims(int which) {
  if (!g_s) return 0;
  switch (which) {
    case 0: return g_s->output_nz;
    case 1: return g_s->deskewedXdim;
    case 2: return g_s->output_nx;
    case 3: return g_s->fft_nx;
    default: return 0;
  }
}
unsigned get_output_nz(void) {
  if (!g_s) return 0;
  if (g_s->rot_on) {
    int nxo, nzo;
    rotated_dims(g_s->rot, g_s->deskewedXdim, g_s->output_nz, &nxo, &nzo);
    return (unsigned)nzo;
  }
  return g_s->output_nz;
}

int RL_interface(const unsigned short* const raw_data, int nx, int ny, int nz, float* result,
                 float* raw_deskewed_result, float background, bool bDoRescale, bool bSaveDeskewedRaw, int nIters,
                 int extraShift, int napodize, int nZblend, float padVal, bool bDupRevStack, bool bSkewedDecon) {
  // bDupRevStack has no effect on the library path's result: its FFT plans and
  // kernel sizes cover only the first nz slices (src/RL-Biggs-Andrews.cpp:193-205
  // against :622-642), so the mirrored half is never filtered.
  (void)bDupRevStack;
  return guard([&]() -> int {
    if (!g_s) fail("RL_interface called before RL_interface_init");
    Session& S = *g_s;
    if (!raw_data || !result) fail("RL_interface: null buffer");
    if (!std::isfinite(background) || !std::isfinite(padVal)) fail("RL_interface: non-finite background or pad value");
    if (nIters < 0) fail("RL_interface: negative iteration count");
    if (!S.st) S.st = make_stream();
    if ((unsigned)nx < S.output_nx || (unsigned)ny < S.output_ny || (unsigned)nz < S.output_nz)
      fail("RL_interface: stack smaller than the initialised size");
    // the reference indexes outside the stack for these (apodize and z-blend
    // loops run napodize / nZblend planes from each edge)
    if (napodize < 0 || (unsigned)napodize > std::min(S.output_ny, S.output_nx))
      fail("RL_interface: napodize exceeds the stack's x or y size");
    if (nZblend < 0 || (unsigned)nZblend > S.output_nz)
      fail("RL_interface: nZblend exceeds the stack's z size");
    const hipStream_t st = S.st;
    const size_t nin = (size_t)nx * ny * nz;
    S.u16.ensure(nin * sizeof(uint16_t));
    copy_to_device(st, S.u16.p, raw_data, nin * sizeof(uint16_t));
    unsigned cx = S.output_nx, cy = S.output_ny, cz = S.output_nz;
    auto vox = [&]() { return (size_t)cx * cy * cz; };
    S.bufA.ensure(vox() * sizeof(float));
    float* X = S.bufA.as<float>();
    u16_to_f32_crop(st, S.u16.as<uint16_t>(), nx, ny, nz, X, (int)cx, (int)cy, (int)cz);
    // scratch buffer that is not X
    auto other = [&](size_t bytes) -> float* {
      DevBuf& b = (X == S.bufA.as<float>()) ? S.bufB : S.bufA;
      b.ensure(bytes);
      return b.as<float>();
    };
    const bool deskew_on = std::fabs(S.deskewFactor) > 0.0;
    if (nIters > 0 || (bSaveDeskewedRaw && deskew_on) || S.rot_on) {
      apodize(st, X, (int)cx, (int)cy, (int)cz, napodize);
      bgsub(st, X, (long)vox(), background);
      S.red_part.ensure(sizeof(double) * 2048);
      S.red_acc.ensure(sizeof(double) * 4);
      S.red_mean.ensure(sizeof(float));
      const double inten = mean_above_background(st, X, (long)vox(), S.red_part.as<double>(), S.red_acc.as<double>(),
                                                 S.red_mean.as<float>());
      if (bDoRescale) {
        if (g_first) {
          g_intensity0 = inten;
          g_first = false;
        } else {
          rescale(st, X, (long)vox(), (float)(g_intensity0 / inten));
        }
      }
      if ((!bSkewedDecon || (bSaveDeskewedRaw && nIters == 0)) && deskew_on) {
        float* D = other(sizeof(float) * (size_t)S.deskewedXdim * cy * cz);
        deskew(st, X, (int)cx, (int)cy, (int)cz, S.deskewFactor, D, (int)S.deskewedXdim, extraShift, padVal);
        X = D;
        cx = S.deskewedXdim;
        if (bSaveDeskewedRaw) {
          if (!raw_deskewed_result) fail("RL_interface: raw_deskewed_result is null");
          copy_to_host(st, raw_deskewed_result, X, vox() * sizeof(float));
          if (nIters == 0) {
            // the reference returns uninitialised memory in `result` here
            copy_to_host(st, result, X, vox() * sizeof(float));
            return 1;
          }
        }
      }
      zblend(st, X, (int)cx, (int)cy, (int)cz, nZblend);
      if (nIters > 0) {
        float ks[3];
        kscales(S, cx, ks);
        S.eng.setup((int)cx, (int)cy, (int)cz, S.otf, ks, st);
        S.eng.load_raw(X, true, st);
        if (g_flat_on) fill(st, S.eng.y.as<float>(), S.eng.total_vox(), g_flat_value);
        S.eng.iterate(nIters, 0, g_flat_on, st, nullptr);
        X = S.eng.result();
      }
    }
    if (bSkewedDecon && deskew_on) {
      float* D = other(sizeof(float) * (size_t)S.deskewedXdim * cy * cz);
      deskew(st, X, (int)cx, (int)cy, (int)cz, S.deskewFactor, D, (int)S.deskewedXdim, extraShift, 0.f);
      X = D;
      cx = S.deskewedXdim;
    }
    if (S.rot_on) {
      int nxo, nzo;
      rotated_dims(S.rot, cx, cz, &nxo, &nzo);
      if (nxo <= 0 || nzo <= 0) fail("RL_interface: rotation gives an empty result");
      float* R = other(sizeof(float) * (size_t)nxo * cy * nzo);
      rotate(st, X, (int)cx, (int)cy, (int)cz, S.rot, R, nxo, nzo);
      X = R;
      cx = (unsigned)nxo;
      cz = (unsigned)nzo;
    }
    // uint16 input and a finite OTF cannot give NaN or Inf; report it if the
    // arithmetic ever does, rather than return a poisoned stack as a success
    check_finite(st, X, (long)vox(), S.red_part, "RL_interface result");
    copy_to_host(st, result, X, vox() * sizeof(float));
    return 1;
  });
}

void RL_cleanup(void) {
  // src/RL-Biggs-Andrews.cpp:707-715
  g_intensity0 = 0.;
  g_first = true;
  g_flat_on = false;
  g_s.reset();
}

int cudaDecon_set_flat_start(int enable, float value) {
  return guard([&]() -> int {
    if (enable && !std::isfinite(value)) fail("cudaDecon_set_flat_start: non-finite value");
    g_flat_on = enable != 0;
    g_flat_value = value;
    return 1;
  });
}

void cuda_reset(void) {
  g_s.reset();
  (void)hipDeviceReset();
}

int Deskew_interface(const float* const raw_data, int nx, int ny, int nz, float dz, float dr, float deskewAngle,
                     float* const result, int outputWidth, int extraShift, float padVal) {
  return guard([&]() -> int {
    if (!raw_data || !result) fail("Deskew_interface: null buffer");
    if (nx <= 0 || ny <= 0 || nz <= 0) fail("Deskew_interface: non-positive dimensions");
    const unsigned xdim = deskew_interface_width(nx, dz, dr, deskewAngle, outputWidth, g_output_nz);
    if (xdim == 0) fail("Deskew_interface: zero output width");
    double factor = 0.;
    if (std::fabs(deskewAngle) > 0.0) {
      if (deskewAngle < 0) deskewAngle += 180.;
      factor = std::cos(deskewAngle * M_PI / 180.) * dz / dr;
    }
    hipStream_t st = make_stream();
    DevBuf din, dout;
    try {
      din.ensure(sizeof(float) * (size_t)nx * ny * nz);
      dout.ensure(sizeof(float) * (size_t)xdim * ny * nz);
      copy_to_device(st, din.p, raw_data, sizeof(float) * (size_t)nx * ny * nz);
      deskew(st, din.as<float>(), nx, ny, nz, factor, dout.as<float>(), (int)xdim, extraShift, padVal);
      copy_to_host(st, result, dout.p, sizeof(float) * (size_t)xdim * ny * nz);
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    (void)hipStreamDestroy(st);
    return 1;
  });
}

static int affine_common(const float* raw_data, int nx, int ny, int nz, float* result, const float* affMat, int ra,
                         float dx, float dy, float dz) {
  return guard([&]() -> int {
    if (!raw_data || !result || !affMat) fail("Affine_interface: null buffer");
    const size_t n = (size_t)nx * ny * nz;
    hipStream_t st = make_stream();
    DevBuf din, dout;
    try {
      din.ensure(sizeof(float) * n);
      dout.ensure(sizeof(float) * n);
      copy_to_device(st, din.p, raw_data, sizeof(float) * n);
      affine(st, din.as<float>(), nx, ny, nz, dout.as<float>(), affMat, ra, dx, dy, dz);
      copy_to_host(st, result, dout.p, sizeof(float) * n);
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    (void)hipStreamDestroy(st);
    return 1;
  });
}

int Affine_interface(const float* const raw_data, int nx, int ny, int nz, float* const result, const float* affMat) {
  return affine_common(raw_data, nx, ny, nz, result, affMat, 0, 1.f, 1.f, 1.f);
}

int Affine_interface_RA(const float* const raw_data, int nx, int ny, int nz, float dx, float dy, float dz,
                        float* const result, const float* affMat) {
  return affine_common(raw_data, nx, ny, nz, result, affMat, 1, dx, dy, dz);
}

int camcor_interface_init(int nx, int ny, int nz, const float* const camparam) {
  return guard([&]() -> int {
    if (!camparam || nx <= 0 || ny <= 0 || nz <= 0) fail("camcor_interface_init: bad argument");
    Camcor& C = camcor_state();
    if (!C.st) C.st = make_stream();
    const size_t b = sizeof(float) * 3 * (size_t)nx * ny;
    C.cam.ensure(b);
    CDN_HIP(hipMemcpyAsync(C.cam.p, camparam, b, hipMemcpyHostToDevice, C.st));
    CDN_HIP(hipStreamSynchronize(C.st));
    C.nx = nx;
    C.ny = ny;
    C.nz = nz;
    return 1;
  });
}

int camcor_interface(const unsigned short* const raw_data, int nx, int ny, int nz, unsigned short* const result) {
  return guard([&]() -> int {
    Camcor& C = camcor_state();
    if (!C.st || C.nx == 0) fail("camcor_interface called before camcor_interface_init");
    if (!raw_data || !result || nz <= 0) fail("camcor_interface: bad argument");
    if (nx != C.nx || ny != C.ny) fail("camcor_interface: nx, ny differ from camcor_interface_init");
    const size_t n = (size_t)nx * ny * nz;
    C.in.ensure(sizeof(uint16_t) * n);
    C.out.ensure(sizeof(uint16_t) * n);
    copy_to_device(C.st, C.in.p, raw_data, sizeof(uint16_t) * n);
    camcor(C.st, C.in.as<uint16_t>(), C.cam.as<float>(), nx, ny, nz, C.out.as<uint16_t>());
    copy_to_host(C.st, result, C.out.p, sizeof(uint16_t) * n);
    return 1;
  });
}

void* cudaDecon_device_malloc(size_t bytes) {
  void* p = nullptr;
  guard([&]() -> int {
    if (bytes == 0) fail("cudaDecon_device_malloc: zero bytes");
    CDN_HIP(hipMalloc(&p, bytes));
    return 1;
  });
  return p;
}

int cudaDecon_device_free(void* p) {
  return guard([&]() -> int {
    if (p) CDN_HIP(hipFree(p));
    return 1;
  });
}

int cudaDecon_device_copy(void* dst, const void* src, size_t bytes, int kind) {
  return guard([&]() -> int {
    if (!dst || !src) fail("cudaDecon_device_copy: null pointer");
    const hipMemcpyKind k = kind == 1   ? hipMemcpyHostToDevice
                            : kind == 2 ? hipMemcpyDeviceToHost
                            : kind == 3 ? hipMemcpyDeviceToDevice
                                        : hipMemcpyDefault;
    if (kind < 1 || kind > 3) fail("cudaDecon_device_copy: kind must be 1, 2 or 3");
    CDN_HIP(hipMemcpy(dst, src, bytes, k));
    return 1;
  });
}

int cudaDecon_device_sync(void) {
  return guard([]() -> int {
    CDN_HIP(hipDeviceSynchronize());
    return 1;
  });
}

// ---- standalone FFT / filter -------------------------------------------------
int cudaDecon_fft_r2c(const float* in, int nx, int ny, int nz, float* out) {
  return guard([&]() -> int {
    hipStream_t st = make_stream();
    Fft3 F;
    DevBuf dv, ds;
    try {
      F.setup(nx, ny, nz, st);
      dv.ensure(sizeof(float) * F.nvox);
      ds.ensure(sizeof(float2) * F.nspec);
      copy_to_device(st, dv.p, in, sizeof(float) * F.nvox);
      F.r2c(st, dv.as<float>(), ds.as<float2>());
      CDN_HIP(hipMemcpy2DAsync(out, sizeof(float2) * F.nxh, ds.p, sizeof(float2) * F.pitch, sizeof(float2) * F.nxh,
                               (size_t)ny * nz, hipMemcpyDeviceToHost, st));
      CDN_HIP(hipStreamSynchronize(st));
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    (void)hipStreamDestroy(st);
    return 1;
  });
}

int cudaDecon_fft_c2r(const float* in, int nx, int ny, int nz, float* out) {
  return guard([&]() -> int {
    hipStream_t st = make_stream();
    Fft3 F;
    DevBuf dv, ds;
    try {
      F.setup(nx, ny, nz, st);
      dv.ensure(sizeof(float) * F.nvox);
      ds.ensure(sizeof(float2) * F.nspec);
      CDN_HIP(hipMemcpy2DAsync(ds.p, sizeof(float2) * F.pitch, in, sizeof(float2) * F.nxh, sizeof(float2) * F.nxh,
                               (size_t)ny * nz, hipMemcpyHostToDevice, st));
      F.c2r(st, ds.as<float2>(), dv.as<float>());
      copy_to_host(st, out, dv.p, sizeof(float) * F.nvox);
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    (void)hipStreamDestroy(st);
    return 1;
  });
}

int cudaDecon_filter(float* vol, int nx, int ny, int nz, const float* otf, int otf_width, int otf_height,
                     int otf_depth, float kxscale, float kyscale, float kzscale, int conj) {
  return guard([&]() -> int {
    hipStream_t st = make_stream();
    try {
      Fft3 F;
      F.setup(nx, ny, nz, st);
      OtfTable T;
      T.set_from_image(otf, otf_width, otf_height, otf_depth, 1.f, 1.f);
      OtfDev D;
      const float ks[3] = {kxscale, kyscale, kzscale};
      D.setup(T, ks, nz, st);
      DevBuf dv, ds;
      dv.ensure(sizeof(float) * F.nvox);
      ds.ensure(sizeof(float2) * F.nspec);
      copy_to_device(st, dv.p, vol, sizeof(float) * F.nvox);
      float2* S = ds.as<float2>();
      launch_x_r2c(nx, st, dv.as<float>(), nullptr, nullptr, nullptr, XIN_PLAIN, S, (long)ny * nz, F.pitch,
                   F.twx.as<float2>());
      F.col(st, S, 1, false);
      launch_z_filter(nz, st, S, F.nxh, (long)ny * F.pitch, F.pitch, ny, conj, D.params(T, 1.0 / (double)F.nvox),
                      F.twz.as<float2>());
      F.col(st, S, 1, true);
      launch_x_c2r(nx, st, S, dv.as<float>(), (long)ny * nz, F.pitch, F.twx.as<float2>());
      copy_to_host(st, vol, dv.p, sizeof(float) * F.nvox);
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    (void)hipStreamDestroy(st);
    return 1;
  });
}

int cudaDecon_wiener(float* vol, int nx, int ny, int nz, const float* otf, int otf_width, int otf_height,
                     float kxscale, float kyscale, float kzscale, float dkx, float dky, float rcutoff,
                     float wiener, float background) {
  return guard([&]() -> int {
    if (!vol || !otf) fail("cudaDecon_wiener: null argument");
    if (!(rcutoff > 0.f)) fail("cudaDecon_wiener: rcutoff must be positive");
    hipStream_t st = make_stream();
    try {
      Fft3 F;
      F.setup(nx, ny, nz, st);
      OtfTable T;
      T.set_from_image(otf, otf_width, otf_height, 1, 1.f, 1.f);
      if (T.nyotf != 1) fail("cudaDecon_wiener: needs a rotationally averaged OTF (depth 1)");
      OtfDev D;
      const float ks[3] = {kxscale, kyscale, kzscale};
      D.setup(T, ks, nz, st);
      DevBuf dv, ds;
      dv.ensure(sizeof(float) * F.nvox);
      ds.ensure(sizeof(float2) * F.nspec);
      copy_to_device(st, dv.p, vol, sizeof(float) * F.nvox);
      if (background != 0.f) add_scalar(st, dv.as<float>(), F.nvox, -background);
      float2* S = ds.as<float2>();
      launch_x_r2c(nx, st, dv.as<float>(), nullptr, nullptr, nullptr, XIN_PLAIN, S, (long)ny * nz, F.pitch,
                   F.twx.as<float2>());
      F.col(st, S, 1, false);
      F.col(st, S, 2, false);
      wiener_multiply(st, S, F.nxh, ny, nz, F.pitch, D.tz.as<float2>(), T.nxotf, kxscale, kyscale, dkx, dky,
                      rcutoff, wiener * wiener, (float)(1.0 / (double)F.nvox));
      F.col(st, S, 2, true);
      F.col(st, S, 1, true);
      launch_x_c2r(nx, st, S, dv.as<float>(), (long)ny * nz, F.pitch, F.twx.as<float2>());
      copy_to_host(st, vol, dv.p, sizeof(float) * F.nvox);
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    (void)hipStreamDestroy(st);
    return 1;
  });
}

int cudaDecon_autotune(const int* dims_zyx, int count, int reps, int use_cache, char* report, int report_len) {
  return guard([&]() -> int {
    if (!dims_zyx || count <= 0) fail("cudaDecon_autotune: no grids");
    std::vector<std::array<int, 3>> g((size_t)count);
    for (int i = 0; i < count; ++i) g[(size_t)i] = {dims_zyx[3 * i], dims_zyx[3 * i + 1], dims_zyx[3 * i + 2]};
    const std::string r = autotune(g, reps, use_cache != 0);
    if (report && report_len > 0) {
      const size_t n = std::min(r.size(), (size_t)report_len - 1);
      std::memcpy(report, r.data(), n);
      report[n] = 0;
    }
    return 1;
  });
}

int cudaDecon_autotune_clear(void) {
  return guard([]() -> int {
    fft_clear_variants();
    return 1;
  });
}

int cudaDecon_fft_variant(int n, int* fields, int nfields) {
  FftVariant v;
  const bool tuned = fft_get_variant(n, &v);
  const int f[7] = {v.cols, v.xthreads, v.col_nt_load, v.xcd_remap, v.col_tpt, v.zf_waves, v.zf_rows};
  for (int i = 0; fields && i < nfields && i < 7; ++i) fields[i] = f[i];
  return tuned ? 1 : 0;
}

// ---- RL engine handle -----------------------------------------------------
cudaDecon_rl* cudaDecon_rl_create(int nx, int ny, int nz, const float* otf, int otf_width, int otf_height,
                                  int otf_depth, float kxscale, float kyscale, float kzscale) {
  cudaDecon_rl* out = nullptr;
  guard([&]() -> int {
    auto h = std::make_unique<cudaDecon_rl>();
    h->st = make_stream();
    OtfTable T;
    T.set_from_image(otf, otf_width, otf_height, otf_depth, 1.f, 1.f);
    const float ks[3] = {kxscale, kyscale, kzscale};
    h->eng.setup(nx, ny, nz, T, ks, h->st);
    out = h.release();
    return 1;
  });
  return out;
}

cudaDecon_rl* cudaDecon_rl_create_stacks(int nx, int ny, int nz, int nstacks, const float* otf, int otf_width,
                                         int otf_height, int otf_depth, float kxscale, float kyscale,
                                         float kzscale) {
  cudaDecon_rl* out = nullptr;
  guard([&]() -> int {
    if (nstacks < 1) fail("cudaDecon_rl_create_stacks: nstacks < 1");
    auto h = std::make_unique<cudaDecon_rl>();
    h->st = make_stream();
    OtfTable T;
    T.set_from_image(otf, otf_width, otf_height, otf_depth, 1.f, 1.f);
    const float ks[3] = {kxscale, kyscale, kzscale};
    h->eng.setup(nx, ny, nz, T, ks, h->st, nstacks);
    out = h.release();
    return 1;
  });
  return out;
}

int cudaDecon_rl_stacks(cudaDecon_rl* h) { return h ? h->eng.nstacks : 0; }

int cudaDecon_rl_set_raw_stack(cudaDecon_rl* h, int s, const float* raw) {
  return guard([&]() -> int {
    if (!h || !raw) fail("cudaDecon_rl_set_raw_stack: null argument");
    h->eng.load_stack(s, raw, false, h->st);
    if (s >= 0 && s < h->eng.nstacks)
      check_finite(h->st, h->eng.raw.as<float>() + (size_t)h->eng.fft.nvox * s, h->eng.fft.nvox, h->red,
                   "cudaDecon_rl_set_raw_stack");
    CDN_HIP(hipStreamSynchronize(h->st));
    h->k_next = 0;
    return 1;
  });
}

int cudaDecon_rl_get_stack(cudaDecon_rl* h, int s, float* X) {
  return guard([&]() -> int {
    if (!h || !X) fail("cudaDecon_rl_get_stack: null argument");
    if (s < 0 || s >= h->eng.nstacks) fail("cudaDecon_rl_get_stack: stack index out of range");
    copy_to_host(h->st, X, h->eng.result() + (size_t)h->eng.fft.nvox * s, sizeof(float) * h->eng.fft.nvox);
    return 1;
  });
}

int cudaDecon_rl_set_raw(cudaDecon_rl* h, const float* raw) {
  return guard([&]() -> int {
    if (!h || !raw) fail("cudaDecon_rl_set_raw: null argument");
    h->eng.load_raw(raw, false, h->st);
    check_finite(h->st, h->eng.raw.as<float>(), h->eng.total_vox(), h->red, "cudaDecon_rl_set_raw");
    CDN_HIP(hipStreamSynchronize(h->st));
    h->k_next = 0;
    return 1;
  });
}

int cudaDecon_rl_set_seed(cudaDecon_rl* h, const float* seed) {
  return guard([&]() -> int {
    if (!h || !seed) fail("cudaDecon_rl_set_seed: null argument");
    h->eng.load_seed(seed, false, h->st);
    check_finite(h->st, h->eng.y.as<float>(), h->eng.total_vox(), h->red, "cudaDecon_rl_set_seed");
    CDN_HIP(hipStreamSynchronize(h->st));
    return 1;
  });
}

int cudaDecon_rl_iterate(cudaDecon_rl* h, int n, int k0, int seeded, double* lambdas, double* elapsed_ms) {
  return guard([&]() -> int {
    if (!h) fail("cudaDecon_rl_iterate: null handle");
    if (n < 0) fail("cudaDecon_rl_iterate: negative iteration count");
    if (k0 != 0 && k0 != h->k_next) fail("cudaDecon_rl_iterate: k0 must be 0 or continue the previous call");
    if (n == 0) {
      if (elapsed_ms) *elapsed_ms = 0.0;
      return 1;
    }
    if (k0 == 0) h->eng.reset(h->st);
    hipEvent_t e0, e1;
    CDN_HIP(hipEventCreate(&e0));
    CDN_HIP(hipEventCreate(&e1));
    CDN_HIP(hipStreamSynchronize(h->st));
    CDN_HIP(hipEventRecord(e0, h->st));
    h->eng.iterate(n, k0, seeded != 0, h->st, nullptr);
    CDN_HIP(hipEventRecord(e1, h->st));
    CDN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CDN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) *elapsed_ms = ms;
    if (lambdas) {
      // the engine's lambda slots of this call (4 doubles per stack and
      // iteration): k < 2 carries no factor; n x nstacks values, stack fastest
      const size_t ns = (size_t)h->eng.nstacks;
      std::vector<double> hl(4 * ns * (size_t)std::max(n, 1));
      CDN_HIP(hipMemcpy(hl.data(), h->eng.lam.p, sizeof(double) * 4 * ns * (size_t)n, hipMemcpyDeviceToHost));
      for (int j = 0; j < n; ++j)
        for (size_t s = 0; s < ns; ++s) lambdas[j * ns + s] = (k0 + j >= 2) ? hl[4 * (j * ns + s)] : 0.0;
    }
    h->k_next = k0 + n;
    return 1;
  });
}

int cudaDecon_rl_get(cudaDecon_rl* h, float* X) {
  return guard([&]() -> int {
    if (!h || !X) fail("cudaDecon_rl_get: null argument");
    copy_to_host(h->st, X, h->eng.result(), sizeof(float) * h->eng.total_vox());
    return 1;
  });
}

int cudaDecon_rl_timing(cudaDecon_rl* h, int enable, double* pass_ms, long* pass_calls) {
  return guard([&]() -> int {
    if (!h) fail("cudaDecon_rl_timing: null handle");
    if (pass_ms)
      for (int i = 0; i < PASS_COUNT; ++i) pass_ms[i] = h->eng.pass_ms[i];
    if (pass_calls)
      for (int i = 0; i < PASS_COUNT; ++i) pass_calls[i] = h->eng.pass_calls[i];
    if (enable >= 0) {
      const bool on = enable != 0;
      if (on && !h->eng.timing)
        for (int i = 0; i < PASS_COUNT; ++i) {
          h->eng.pass_ms[i] = 0;
          h->eng.pass_calls[i] = 0;
        }
      h->eng.timing = on;
    }
    return 1;
  });
}

int cudaDecon_rl_geometry(cudaDecon_rl* h, long* nvox, long* nspec_alg, long* pitch) {
  return guard([&]() -> int {
    if (!h) fail("cudaDecon_rl_geometry: null handle");
    if (nvox) *nvox = h->eng.fft.nvox;
    if (nspec_alg) *nspec_alg = (long)h->eng.fft.nz * h->eng.fft.ny * h->eng.fft.nxh;
    if (pitch) *pitch = h->eng.fft.pitch;
    return 1;
  });
}

static void block_copy(cudaDecon_rl* h, float* dev, const float* host_in, float* host_out, int bx, int by, int bz,
                       int ox, int oy, int oz) {
  const Fft3& F = h->eng.fft;
  if (bx <= 0 || by <= 0 || bz <= 0 || ox < 0 || oy < 0 || oz < 0 || ox + bx > F.nx || oy + by > F.ny ||
      oz + bz > F.nz)
    fail("block outside the stack");
  for (int z = 0; z < bz; ++z) {
    float* d = dev + ((size_t)(oz + z) * F.ny + oy) * F.nx + ox;
    if (host_in)
      CDN_HIP(hipMemcpy2DAsync(d, sizeof(float) * F.nx, host_in + (size_t)z * bx * by, sizeof(float) * bx,
                               sizeof(float) * bx, by, hipMemcpyHostToDevice, h->st));
    else
      CDN_HIP(hipMemcpy2DAsync(host_out + (size_t)z * bx * by, sizeof(float) * bx, d, sizeof(float) * F.nx,
                               sizeof(float) * bx, by, hipMemcpyDeviceToHost, h->st));
  }
  CDN_HIP(hipStreamSynchronize(h->st));
}

int cudaDecon_rl_set_raw_block(cudaDecon_rl* h, const float* block, int bx, int by, int bz, int ox, int oy, int oz,
                               float background) {
  return guard([&]() -> int {
    one_stack(h, "cudaDecon_rl_set_raw_block");
    if (!h || !block) fail("cudaDecon_rl_set_raw_block: null argument");
    fill(h->st, h->eng.raw.as<float>(), h->eng.fft.nvox, background);
    if (!std::isfinite(background)) fail("cudaDecon_rl_set_raw_block: non-finite background");
    block_copy(h, h->eng.raw.as<float>(), block, nullptr, bx, by, bz, ox, oy, oz);
    check_finite(h->st, h->eng.raw.as<float>(), h->eng.fft.nvox, h->red, "cudaDecon_rl_set_raw_block");
    h->eng.reset(h->st);
    CDN_HIP(hipStreamSynchronize(h->st));
    h->k_next = 0;
    return 1;
  });
}

int cudaDecon_rl_get_block(cudaDecon_rl* h, float* out, int bx, int by, int bz, int ox, int oy, int oz) {
  return guard([&]() -> int {
    one_stack(h, "cudaDecon_rl_get_block");
    if (!h || !out) fail("cudaDecon_rl_get_block: null argument");
    block_copy(h, h->eng.result(), nullptr, out, bx, by, bz, ox, oy, oz);
    return 1;
  });
}

int cudaDecon_rl_stats(cudaDecon_rl* h, double* out6) {
  return guard([&]() -> int {
    one_stack(h, "cudaDecon_rl_stats");
    if (!h || !out6) fail("cudaDecon_rl_stats: null argument");
    h->red.ensure(sizeof(double) * 3 * 1024);
    stats(h->st, h->eng.raw.as<float>(), h->eng.fft.nvox, h->red.as<double>(), out6);
    stats(h->st, h->eng.result(), h->eng.fft.nvox, h->red.as<double>(), out6 + 3);
    return 1;
  });
}

int cudaDecon_rl_set_raw_u16(cudaDecon_rl* h, const unsigned short* raw, float offset) {
  return guard([&]() -> int {
    one_stack(h, "cudaDecon_rl_set_raw_u16");
    if (!h || !raw) fail("cudaDecon_rl_set_raw_u16: null argument");
    const Fft3& F = h->eng.fft;
    h->u16.ensure(sizeof(uint16_t) * (size_t)F.nvox);
    copy_to_device(h->st, h->u16.p, raw, sizeof(uint16_t) * (size_t)F.nvox);
    // src/RLgpuImpl.cu:124-132 with bg = offset
    u16_to_f32_bgsub(h->st, h->u16.as<uint16_t>(), F.nx, F.ny, F.nz, h->eng.raw.as<float>(), offset);
    h->eng.reset(h->st);
    CDN_HIP(hipStreamSynchronize(h->st));
    h->k_next = 0;
    h->post_cur = -1;
    return 1;
  });
}

int cudaDecon_rl_deskew(cudaDecon_rl* h, float dz, float dr, float deskewAngle, int outputWidth, int extraShift,
                        float padVal, int* nx_out) {
  return guard([&]() -> int {
    one_stack(h, "cudaDecon_rl_deskew");
    if (!h) fail("cudaDecon_rl_deskew: null handle");
    const Fft3& F = h->eng.fft;
    double factor = 0.;
    int xdim = F.nx;
    if (std::fabs(deskewAngle) > 0.0) {
      double a = deskewAngle;
      if (a < 0) a += 180.;
      factor = std::cos(a * M_PI / 180.) * dz / dr;
      xdim = outputWidth > 0 ? outputWidth
                             : (int)find_optimal_dimension(
                                   (unsigned)(F.nx + std::floor(F.nz * dz * std::fabs(std::cos(a * M_PI / 180.)) / dr)));
    }
    DevBuf& out = h->post[0];
    out.ensure(sizeof(float) * (size_t)xdim * F.ny * F.nz);
    deskew(h->st, h->eng.result(), F.nx, F.ny, F.nz, factor, out.as<float>(), xdim, extraShift, padVal);
    CDN_HIP(hipStreamSynchronize(h->st));
    h->post_cur = 0;
    h->pnx = xdim;
    h->pny = F.ny;
    h->pnz = F.nz;
    if (nx_out) *nx_out = xdim;
    return 1;
  });
}

int cudaDecon_rl_affine(cudaDecon_rl* h, const float* affMat) {
  return guard([&]() -> int {
    one_stack(h, "cudaDecon_rl_affine");
    if (!h || !affMat) fail("cudaDecon_rl_affine: null argument");
    const Fft3& F = h->eng.fft;
    const float* in = h->post_cur >= 0 ? h->post[h->post_cur].as<float>() : h->eng.result();
    if (h->post_cur < 0) {
      h->pnx = F.nx;
      h->pny = F.ny;
      h->pnz = F.nz;
    }
    const int dst = h->post_cur == 0 ? 1 : 0;
    h->post[dst].ensure(sizeof(float) * (size_t)h->pnx * h->pny * h->pnz);
    affine(h->st, in, h->pnx, h->pny, h->pnz, h->post[dst].as<float>(), affMat, 0, 0.f, 0.f, 0.f);
    CDN_HIP(hipStreamSynchronize(h->st));
    h->post_cur = dst;
    return 1;
  });
}

int cudaDecon_rl_post_dims(cudaDecon_rl* h, int* nx, int* ny, int* nz) {
  return guard([&]() -> int {
    if (!h || h->post_cur < 0) fail("cudaDecon_rl_post_dims: no deskew or affine result");
    if (nx) *nx = h->pnx;
    if (ny) *ny = h->pny;
    if (nz) *nz = h->pnz;
    return 1;
  });
}

int cudaDecon_rl_post_get(cudaDecon_rl* h, float* out) {
  return guard([&]() -> int {
    if (!h || !out || h->post_cur < 0) fail("cudaDecon_rl_post_get: no deskew or affine result");
    copy_to_host(h->st, out, h->post[h->post_cur].p, sizeof(float) * (size_t)h->pnx * h->pny * h->pnz);
    return 1;
  });
}

int cudaDecon_rl_post_get_block(cudaDecon_rl* h, int buf, float* out, int bx, int by, int bz, int ox, int oy,
                                int oz) {
  return guard([&]() -> int {
    if (!h || !out || buf < 0 || buf > 1 || !h->post[buf].p || h->post_cur < 0)
      fail("cudaDecon_rl_post_get_block: no such deskew or affine result");
    if (bx <= 0 || by <= 0 || bz <= 0 || ox < 0 || oy < 0 || oz < 0 || ox + bx > h->pnx || oy + by > h->pny ||
        oz + bz > h->pnz)
      fail("cudaDecon_rl_post_get_block: block outside the result");
    const float* src = h->post[buf].as<float>();
    for (int z = 0; z < bz; ++z)
      CDN_HIP(hipMemcpy2DAsync(out + (size_t)z * bx * by, sizeof(float) * bx,
                               src + ((size_t)(oz + z) * h->pny + oy) * h->pnx + ox, sizeof(float) * h->pnx,
                               sizeof(float) * bx, by, hipMemcpyDeviceToHost, h->st));
    CDN_HIP(hipStreamSynchronize(h->st));
    return 1;
  });
}

int cudaDecon_rl_buffer_stats(cudaDecon_rl* h, int which, double* out3) {
  return guard([&]() -> int {
    one_stack(h, "cudaDecon_rl_buffer_stats");
    if (!h || !out3 || which < 0 || which > 3) fail("cudaDecon_rl_buffer_stats: bad argument");
    const float* b = which == 0 ? h->eng.raw.as<float>()
                     : which == 1 ? h->eng.result()
                     : which == 2 ? h->eng.y.as<float>()
                                  : h->eng.g.as<float>();
    h->red.ensure(sizeof(double) * 3 * 1024);
    stats(h->st, b, h->eng.fft.nvox, h->red.as<double>(), out3);
    return 1;
  });
}

int cudaDecon_rl_post_stats(cudaDecon_rl* h, double* out3) {
  return guard([&]() -> int {
    if (!h || !out3 || h->post_cur < 0) fail("cudaDecon_rl_post_stats: no deskew or affine result");
    h->red.ensure(sizeof(double) * 3 * 1024);
    stats(h->st, h->post[h->post_cur].as<float>(), (long)h->pnx * h->pny * h->pnz, h->red.as<double>(), out3);
    return 1;
  });
}

int cudaDecon_rl_snapshot(cudaDecon_rl* h) {
  return guard([&]() -> int {
    one_stack(h, "cudaDecon_rl_snapshot");
    if (!h) fail("cudaDecon_rl_snapshot: null handle");
    const size_t b = sizeof(float) * (size_t)h->eng.fft.nvox;
    h->ref.ensure(b);
    CDN_HIP(hipMemcpyAsync(h->ref.p, h->eng.result(), b, hipMemcpyDeviceToDevice, h->st));
    CDN_HIP(hipStreamSynchronize(h->st));
    return 1;
  });
}

int cudaDecon_rl_distance(cudaDecon_rl* h, int which, double* rel) {
  return guard([&]() -> int {
    one_stack(h, "cudaDecon_rl_distance");
    if (!h || !rel || which < 0 || which > 1) fail("cudaDecon_rl_distance: bad argument");
    if (!h->ref.p) fail("cudaDecon_rl_distance: no snapshot taken");
    h->red.ensure(sizeof(double) * 3 * 1024);
    double o[2];
    const float* v = which == 0 ? h->eng.result() : h->eng.y.as<float>();
    diff_norm(h->st, v, h->ref.as<float>(), h->eng.fft.nvox, h->red.as<double>(), o);
    *rel = o[1] > 0 ? std::sqrt(o[0] / o[1]) : std::sqrt(o[0]);
    return 1;
  });
}

int cudaDecon_rl_copy_probe(cudaDecon_rl* h, int n, double* ms) {
  return guard([&]() -> int {
    if (!h || n <= 0) fail("cudaDecon_rl_copy_probe: bad argument");
    const size_t bytes = sizeof(float) * (size_t)h->eng.fft.nvox / 16 * 16;
    h->probe.ensure(bytes);
    hipEvent_t e0, e1;
    CDN_HIP(hipEventCreate(&e0));
    CDN_HIP(hipEventCreate(&e1));
    CDN_HIP(hipEventRecord(e0, h->st));
    for (int i = 0; i < n; ++i) stream_copy(h->st, h->eng.raw.p, h->probe.p, bytes);
    CDN_HIP(hipEventRecord(e1, h->st));
    CDN_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    CDN_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t / n;
    return 1;
  });
}

void cudaDecon_rl_destroy(cudaDecon_rl* h) { delete h; }

int cudaDecon_rl_run_noalloc(const float* d_raw, float* d_out, int nx, int ny, int nz, const float* otf,
                             int otf_width, int otf_height, int otf_depth, float kxscale, float kyscale,
                             float kzscale, int n_iters, void* stream) {
  return guard([&]() -> int {
    const float ks[3] = {kxscale, kyscale, kzscale};
    try {
      noalloc_run(d_raw, d_out, nx, ny, nz, otf, otf_width, otf_height, otf_depth, ks, n_iters,
                  static_cast<hipStream_t>(stream));
    } catch (const std::exception& e) {
      fail(std::string("cudaDecon_rl_run_noalloc: ") + e.what());
    }
    return 1;
  });
}

// ---- batch of stacks ------------------------------------------------------
cudaDecon_batch* cudaDecon_batch_create(int n, const int* dims_zyx, const float* otf, int otf_width, int otf_height,
                                        int otf_depth, const float* kscales, int nstreams) {
  cudaDecon_batch* out = nullptr;
  guard([&]() -> int {
    if (!dims_zyx || !otf || !kscales || n <= 0) fail("cudaDecon_batch_create: bad argument");
    auto h = std::make_unique<cudaDecon_batch>();
    std::vector<OtfTable> T(1);
    T[0].set_from_image(otf, otf_width, otf_height, otf_depth, 1.f, 1.f);
    const std::vector<int> table_of((size_t)n, 0);
    h->b.setup(n, dims_zyx, std::move(T), table_of.data(), kscales, nstreams);
    out = h.release();
    return 1;
  });
  return out;
}

cudaDecon_batch* cudaDecon_batch_create_tables(int n, const int* dims_zyx, int ntables, const float* otf_data,
                                               const int* otf_whd, const int* table_of, const float* kscales,
                                               int nstreams) {
  cudaDecon_batch* out = nullptr;
  guard([&]() -> int {
    if (!dims_zyx || !otf_data || !otf_whd || !table_of || !kscales || n <= 0 || ntables <= 0)
      fail("cudaDecon_batch_create_tables: bad argument");
    auto h = std::make_unique<cudaDecon_batch>();
    std::vector<OtfTable> T((size_t)ntables);
    size_t off = 0;
    for (int t = 0; t < ntables; ++t) {
      const int w = otf_whd[3 * t], ht = otf_whd[3 * t + 1], d = otf_whd[3 * t + 2];
      if (w <= 0 || ht <= 0 || d <= 0) fail("cudaDecon_batch_create_tables: OTF " + std::to_string(t) + " is empty");
      T[t].set_from_image(otf_data + off, w, ht, d, 1.f, 1.f);
      off += (size_t)w * ht * d;
    }
    h->b.setup(n, dims_zyx, std::move(T), table_of, kscales, nstreams);
    out = h.release();
    return 1;
  });
  return out;
}

int cudaDecon_batch_lambdas(cudaDecon_batch* h, int i, int iters, double* out) {
  return guard([&]() -> int {
    if (!h || !out || iters <= 0) fail("cudaDecon_batch_lambdas: bad argument");
    const std::vector<double> l = h->b.lambdas(i, iters);
    std::copy(l.begin(), l.end(), out);
    return 1;
  });
}

int cudaDecon_batch_set_grid_cap(cudaDecon_batch* h, long max_workgroups) {
  return guard([&]() -> int {
    if (!h || max_workgroups < 0) fail("cudaDecon_batch_set_grid_cap: bad argument");
    h->b.grid_cap = max_workgroups;
    return 1;
  });
}

int cudaDecon_batch_set_partitions(cudaDecon_batch* h, int n) {
  return guard([&]() -> int {
    if (!h) fail("cudaDecon_batch_set_partitions: bad argument");
    h->b.set_partitions(n);
    return 1;
  });
}

int cudaDecon_batch_launches(cudaDecon_batch* h) { return h ? h->b.grouped_launches() : 0; }

int cudaDecon_batch_set_raw(cudaDecon_batch* h, int i, const float* raw) {
  return guard([&]() -> int {
    if (!h || !raw || i < 0 || i >= (int)h->b.engines.size()) fail("cudaDecon_batch_set_raw: bad argument");
    RLEngine& e = *h->b.engines[i];
    e.load_raw(raw, false, h->b.streams[0]);
    check_finite(h->b.streams[0], e.raw.as<float>(), e.total_vox(), h->red, "cudaDecon_batch_set_raw");
    CDN_HIP(hipStreamSynchronize(h->b.streams[0]));
    return 1;
  });
}

int cudaDecon_batch_set_raw_block(cudaDecon_batch* h, int i, const float* block, int bx, int by, int bz, int ox,
                                  int oy, int oz, float background) {
  return guard([&]() -> int {
    if (!h || !block || i < 0 || i >= (int)h->b.engines.size()) fail("cudaDecon_batch_set_raw_block: bad argument");
    RLEngine& e = *h->b.engines[i];
    hipStream_t st = h->b.streams[0];
    const Fft3& F = e.fft;
    if (bx <= 0 || by <= 0 || bz <= 0 || ox < 0 || oy < 0 || oz < 0 || ox + bx > F.nx || oy + by > F.ny ||
        oz + bz > F.nz)
      fail("block outside the stack");
    fill(st, e.raw.as<float>(), F.nvox, background);
    for (int z = 0; z < bz; ++z)
      CDN_HIP(hipMemcpy2DAsync(e.raw.as<float>() + ((size_t)(oz + z) * F.ny + oy) * F.nx + ox, sizeof(float) * F.nx,
                               block + (size_t)z * bx * by, sizeof(float) * bx, sizeof(float) * bx, by,
                               hipMemcpyHostToDevice, st));
    check_finite(st, e.raw.as<float>(), F.nvox, h->red, "cudaDecon_batch_set_raw_block");
    e.reset(st);
    CDN_HIP(hipStreamSynchronize(st));
    return 1;
  });
}

int cudaDecon_batch_run(cudaDecon_batch* h, int iters, int mode, double* ms) {
  return guard([&]() -> int {
    if (!h || iters <= 0 || mode < 0 || mode > 4) fail("cudaDecon_batch_run: bad argument");
    const double t = h->b.run(iters, mode);
    if (ms) *ms = t;
    return 1;
  });
}

int cudaDecon_batch_get(cudaDecon_batch* h, int i, float* X) {
  return guard([&]() -> int {
    if (!h || !X || i < 0 || i >= (int)h->b.engines.size()) fail("cudaDecon_batch_get: bad argument");
    RLEngine& e = *h->b.engines[i];
    CDN_HIP(hipMemcpy(X, e.result(), sizeof(float) * e.fft.nvox, hipMemcpyDeviceToHost));
    return 1;
  });
}

void cudaDecon_batch_destroy(cudaDecon_batch* h) { delete h; }

// ---- warm-start network ---------------------------------------------------
cudaDecon_net* cudaDecon_net_create(unsigned long long seed) {
  cudaDecon_net* out = nullptr;
  guard([&]() -> int {
    auto h = std::make_unique<cudaDecon_net>();
    h->st = make_stream();
    h->net.init(seed, h->st);
    out = h.release();
    return 1;
  });
  return out;
}

long cudaDecon_net_num_params(void) { return WarmNet::num_params(); }

int cudaDecon_net_get_params(cudaDecon_net* h, float* params) {
  return guard([&]() -> int {
    if (!h || !params) fail("cudaDecon_net_get_params: null argument");
    h->net.get_params(params, h->st);
    return 1;
  });
}

int cudaDecon_net_set_params(cudaDecon_net* h, const float* params) {
  return guard([&]() -> int {
    if (!h || !params) fail("cudaDecon_net_set_params: null argument");
    h->net.set_params(params, h->st);
    return 1;
  });
}

namespace {
// stage a host batch (and optional target) in the network's batch buffers
void stage_batch(cudaDecon_net* h, const float* raw, const float* tgt, int nb, int nz, int ny, int nx, bool training) {
  if (!raw) fail("warm-start network: null input");
  h->net.geometry(nb, nz, ny, nx, training);
  const size_t n = (size_t)nb * nz * ny * nx;
  h->net.b_raw.ensure(sizeof(float) * n);
  copy_to_device(h->st, h->net.b_raw.p, raw, sizeof(float) * n);
  if (tgt) {
    h->net.b_tgt.ensure(sizeof(float) * n);
    copy_to_device(h->st, h->net.b_tgt.p, tgt, sizeof(float) * n);
  }
}
}  // namespace

int cudaDecon_net_forward(cudaDecon_net* h, const float* raw, int nb, int nz, int ny, int nx, int mode, float* out) {
  return guard([&]() -> int {
    if (!h || !out || mode < 0 || mode > 3) fail("cudaDecon_net_forward: bad argument");
    stage_batch(h, raw, nullptr, nb, nz, ny, nx, false);
    const size_t n = (size_t)nb * nz * ny * nx;
    DevBuf o;
    o.ensure(sizeof(float) * n);
    h->net.forward(h->st, h->net.b_raw.as<float>(), mode < 2);
    h->net.head(h->st, h->net.b_raw.as<float>(), mode & 1, o.as<float>());
    copy_to_host(h->st, out, o.p, sizeof(float) * n);
    return 1;
  });
}

long cudaDecon_net_fetch(cudaDecon_net* h, int which, float* out) {
  long n = -1;
  guard([&]() -> int {
    if (!h) fail("cudaDecon_net_fetch: null handle");
    n = h->net.fetch(which, out, h->st);
    return 1;
  });
  return n;
}

int cudaDecon_net_loss_grad(cudaDecon_net* h, const float* raw, const float* target, int nb, int nz, int ny, int nx,
                            double* loss, float* grad) {
  return guard([&]() -> int {
    if (!h || !target) fail("cudaDecon_net_loss_grad: null argument");
    stage_batch(h, raw, target, nb, nz, ny, nx, true);
    WarmNet& N = h->net;
    N.forward(h->st, N.b_raw.as<float>());
    N.backward(h->st, N.b_raw.as<float>(), N.b_tgt.as<float>(), N.loss.as<float>());
    float l = 0.f;
    CDN_HIP(hipMemcpyAsync(&l, N.loss.p, sizeof(float), hipMemcpyDeviceToHost, h->st));
    if (grad) CDN_HIP(hipMemcpyAsync(grad, N.Gr.p, sizeof(float) * N.nparam, hipMemcpyDeviceToHost, h->st));
    CDN_HIP(hipStreamSynchronize(h->st));
    if (loss) *loss = l;
    return 1;
  });
}

int cudaDecon_net_step(cudaDecon_net* h, const float* raw, const float* target, int nb, int nz, int ny, int nx,
                       float lr, double* loss) {
  return guard([&]() -> int {
    if (!h || !target) fail("cudaDecon_net_step: null argument");
    stage_batch(h, raw, target, nb, nz, ny, nx, true);
    WarmNet& N = h->net;
    N.forward(h->st, N.b_raw.as<float>());
    N.backward(h->st, N.b_raw.as<float>(), N.b_tgt.as<float>(), N.loss.as<float>());
    N.adam(h->st, lr);
    float l = 0.f;
    CDN_HIP(hipMemcpyAsync(&l, N.loss.p, sizeof(float), hipMemcpyDeviceToHost, h->st));
    CDN_HIP(hipStreamSynchronize(h->st));
    if (loss) *loss = l;
    return 1;
  });
}

int cudaDecon_net_add_samples(cudaDecon_net* h, const float* raw, const float* target, int count, int nz, int ny,
                              int nx) {
  return guard([&]() -> int {
    if (!h || !raw || !target || count <= 0) fail("cudaDecon_net_add_samples: bad argument");
    if (nz % 4 || ny % 4 || nx % 4 || nz < 4 || ny < 4 || nx < 4)
      fail("cudaDecon_net_add_samples: every dimension must be a positive multiple of 4");
    h->net.add_samples(raw, target, count, nz, ny, nx, h->st);
    return 1;
  });
}

int cudaDecon_net_train(cudaDecon_net* h, int steps, int batch, float lr, unsigned long long seed, float* losses,
                        double* elapsed_ms) {
  return guard([&]() -> int {
    if (!h || steps < 0) fail("cudaDecon_net_train: bad argument");
    hipEvent_t e0, e1;
    CDN_HIP(hipEventCreate(&e0));
    CDN_HIP(hipEventCreate(&e1));
    CDN_HIP(hipEventRecord(e0, h->st));
    h->net.train(steps, batch, lr, seed, h->st, losses);
    CDN_HIP(hipEventRecord(e1, h->st));
    CDN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CDN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) *elapsed_ms = ms;
    return 1;
  });
}

void cudaDecon_net_destroy(cudaDecon_net* h) { delete h; }

int cudaDecon_rl_warm_start(cudaDecon_rl* h, cudaDecon_net* net, double* elapsed_ms) {
  return guard([&]() -> int {
    if (!h || !net) fail("cudaDecon_rl_warm_start: null argument");
    const Fft3& F = h->eng.fft;
    WarmNet& N = net->net;
    N.geometry(h->eng.nstacks, F.nz, F.ny, F.nx, false);  // one sample per stack, each normalised by its mean
    CDN_HIP(hipStreamSynchronize(net->st));  // parameter uploads of the network's stream
    hipEvent_t e0, e1;
    CDN_HIP(hipEventCreate(&e0));
    CDN_HIP(hipEventCreate(&e1));
    CDN_HIP(hipStreamSynchronize(h->st));
    CDN_HIP(hipEventRecord(e0, h->st));
    N.forward(h->st, h->eng.raw.as<float>(), false);
    N.head(h->st, h->eng.raw.as<float>(), 1, h->eng.y.as<float>());
    CDN_HIP(hipEventRecord(e1, h->st));
    CDN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CDN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) *elapsed_ms = ms;
    return 1;
  });
}

}  // extern "C"
