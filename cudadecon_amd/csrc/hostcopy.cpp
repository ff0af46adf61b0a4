// Staged host <-> device copies (see hostcopy.hpp).
#include "hostcopy.hpp"

#include "common.hpp"
#include "copypool.hpp"

#include <algorithm>
#include <mutex>

namespace cdn {
namespace {

constexpr size_t kChunk = 32UL << 20;   // bytes per staging buffer
constexpr size_t kDirect = 16UL << 20;  // smaller copies go straight through the runtime

struct Stage {
  std::mutex mu;
  CopyPool* pool = nullptr;
  void* buf[2] = {nullptr, nullptr};  // pinned, portable across devices
  void ensure() {
    for (int i = 0; i < 2; ++i)
      if (!buf[i]) CDN_HIP(hipHostMalloc(&buf[i], kChunk, hipHostMallocPortable));
    if (!pool) pool = new CopyPool;
  }
};

// The two events of one call (the last DMA through buf[0] and buf[1]),
// created on the device of the call's stream and destroyed when the call
// returns, after every DMA through the staging buffers has finished.  No
// event outlives its call, so none is ever waited on after the stream it
// was recorded on has been destroyed (the caller may destroy it right after
// the call).
struct CallEvents {
  hipEvent_t e[2] = {nullptr, nullptr};
  explicit CallEvents(hipStream_t st) {
    int dev = 0, cur = 0;
    CDN_HIP(hipStreamGetDevice(st, &dev));
    CDN_HIP(hipGetDevice(&cur));
    if (cur != dev) CDN_HIP(hipSetDevice(dev));
    hipError_t rc = hipSuccess;
    for (int i = 0; i < 2 && rc == hipSuccess; ++i) rc = hipEventCreateWithFlags(&e[i], hipEventDisableTiming);
    if (cur != dev) (void)hipSetDevice(cur);
    if (rc != hipSuccess) {
      release();
      CDN_HIP(rc);
    }
  }
  void release() {
    for (hipEvent_t& x : e)
      if (x) {
        (void)hipEventDestroy(x);
        x = nullptr;
      }
  }
  ~CallEvents() { release(); }
};

// process lifetime: the pinned buffers are released when the process ends
Stage& stage() {
  static Stage* s = new Stage;
  return *s;
}

// on a failed call: no DMA through the staging buffers may still be in
// flight when the next call takes them (the error itself is reported)
void drain(hipStream_t st) { (void)hipStreamSynchronize(st); }

}  // namespace

void copy_to_device(hipStream_t st, void* dev, const void* host, size_t bytes) {
  if (bytes < kDirect) {
    CDN_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
    return;
  }
  Stage& S = stage();
  std::lock_guard<std::mutex> lock(S.mu);
  S.ensure();
  CallEvents E(st);
  const size_t n = (bytes + kChunk - 1) / kChunk;
  try {
    for (size_t i = 0; i < n; ++i) {
      const int b = (int)(i & 1);
      const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
      if (i >= 2) CDN_HIP(hipEventSynchronize(E.e[b]));  // the DMA of chunk i - 2 left buf[b]
      S.pool->copy(S.buf[b], static_cast<const char*>(host) + off, len);
      CDN_HIP(hipMemcpyAsync(static_cast<char*>(dev) + off, S.buf[b], len, hipMemcpyHostToDevice, st));
      CDN_HIP(hipEventRecord(E.e[b], st));
    }
    // the staging buffers are free for the next call
    for (size_t b = 0; b < 2 && b < n; ++b) CDN_HIP(hipEventSynchronize(E.e[b]));
  } catch (...) {
    drain(st);
    throw;
  }
}

void copy_to_host(hipStream_t st, void* host, const void* dev, size_t bytes) {
  if (bytes < kDirect) {
    CDN_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
    CDN_HIP(hipStreamSynchronize(st));
    return;
  }
  Stage& S = stage();
  std::lock_guard<std::mutex> lock(S.mu);
  S.ensure();
  CallEvents E(st);
  const size_t n = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t i) {
    const int b = (int)(i & 1);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    CDN_HIP(hipMemcpyAsync(S.buf[b], static_cast<const char*>(dev) + off, len, hipMemcpyDeviceToHost, st));
    CDN_HIP(hipEventRecord(E.e[b], st));
  };
  try {
    issue(0);
    for (size_t i = 0; i < n; ++i) {
      // buf[(i + 1) & 1] was emptied by the host copy of chunk i - 1
      if (i + 1 < n) issue(i + 1);
      const int b = (int)(i & 1);
      const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
      CDN_HIP(hipEventSynchronize(E.e[b]));
      S.pool->copy(static_cast<char*>(host) + off, S.buf[b], len);
    }
  } catch (...) {
    drain(st);
    throw;
  }
}

}  // namespace cdn
