// RL-BA run that allocates nothing: every device buffer the iteration needs
// beyond the caller's stack and result (X ping-pong, prediction, update
// difference, spectrum, twiddle and OTF tables, dot-product partials,
// factors) is carved from `scratch`, a static device array of
// kNoAllocScratchBytes that each GPU holds from the library's load until the
// process ends.  For callers that cannot allocate device memory (code under
// stream capture, tight serving loops, harnesses that forbid it).
#include "noalloc.hpp"

#include "decon.hpp"
#include "fft_kernels.hpp"

#include <map>
#include <mutex>

namespace cdn {

namespace {

__device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_mu;                 // one run at a time per process: they share the array
std::map<int, char*> g_scratch;  // device -> address of its copy of scratch

// The device address of this device's scratch, looked up once per device
// with that device current.  In host code `scratch` names a host-side
// stand-in, never the device memory, so only this address is used.
char* device_scratch() {
  int dev = -1;
  CDN_HIP(hipGetDevice(&dev));
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  void* p = nullptr;
  CDN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) fail("no device address for the static scratch array");
  g_scratch[dev] = static_cast<char*>(p);
  return static_cast<char*>(p);
}

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

void noalloc_run(const float* d_raw, float* d_out, int nx, int ny, int nz, const float* otf, int otf_width,
                 int otf_height, int otf_depth, const float kscale[3], int n_iters, hipStream_t st) {
  if (!d_raw || !d_out || !otf) fail("null argument");
  if (nx <= 0 || ny <= 0 || nz <= 0) fail("non-positive dimensions");
  if (n_iters < 0) fail("negative iteration count");
  for (int n : {nx, ny, nz})
    if (!fft_length_precompiled(n))
      fail("FFT length " + std::to_string(n) +
           " is not built into the library (run-time compiled kernels load code, which allocates)");
  // the volumes and spectrum first: a stack too large is refused before any
  // GPU work
  const size_t vb = sizeof(float) * (size_t)nx * ny * nz;
  const size_t sb = sizeof(float2) * (size_t)nz * ny * Fft3::pitch_of(nx);
  const size_t big = 4 * up256(vb) + up256(sb);
  if (big > kNoAllocScratchBytes)
    fail("the stack needs " + std::to_string(big) + " bytes of scratch; the static scratch holds " +
         std::to_string(kNoAllocScratchBytes) + " (about 20 bytes per voxel: at most " +
         std::to_string(kNoAllocScratchBytes / 20) + " voxels)");
  OtfTable T;
  T.set_from_image(otf, otf_width, otf_height, otf_depth, 1.f, 1.f);

  std::lock_guard<std::mutex> lk(g_mu);
  DevRegion R{device_scratch(), kNoAllocScratchBytes, 0};
  const BuiltinShapesOnly builtin;
  // the engine reads the caller's stack as its (read-only) raw stack
  EngineViews v{};
  v.raw = const_cast<float*>(d_raw);
  v.xa = static_cast<float*>(R.take(vb));
  v.xb = static_cast<float*>(R.take(vb));
  v.y = static_cast<float*>(R.take(vb));
  v.g = static_cast<float*>(R.take(vb));
  v.spec = static_cast<float2*>(R.take(sb));
  RLEngine eng;
  eng.setup(nx, ny, nz, T, kscale, st, 1, &v, &R);
  eng.bind_lambdas(R, n_iters);
  // every scratch byte the iteration reads is written first in this call
  // (X_0 copied from the stack; Y, G, partials and factors written before
  // the iterations that read them), so the last call's contents need no
  // clearing
  eng.reset(st);
  eng.iterate(n_iters, 0, false, st, nullptr);
  CDN_HIP(hipMemcpyAsync(d_out, eng.result(), vb, hipMemcpyDeviceToDevice, st));
  CDN_HIP(hipStreamSynchronize(st));
}

}  // namespace cdn
