// Staged host <-> device copies (see hostcopy.hpp).
#include "hostcopy.hpp"

#include "common.hpp"
#include "copypool.hpp"

#include <algorithm>
#include <map>
#include <mutex>

namespace cdn {
namespace {

constexpr size_t kChunk = 32UL << 20;   // bytes per staging buffer
constexpr size_t kDirect = 16UL << 20;  // smaller copies go straight through the runtime

struct Stage {
  std::mutex mu;
  CopyPool* pool = nullptr;
  void* buf[2] = {nullptr, nullptr};      // pinned, portable across devices
  std::map<int, hipEvent_t> ev[2];        // last DMA through buf[i], per device
  // the event for buf[i] on the device of stream st (created there)
  hipEvent_t event(int i, hipStream_t st) {
    int dev = 0;
    CDN_HIP(hipStreamGetDevice(st, &dev));
    auto it = ev[i].find(dev);
    if (it != ev[i].end()) return it->second;
    int cur = 0;
    CDN_HIP(hipGetDevice(&cur));
    if (cur != dev) CDN_HIP(hipSetDevice(dev));
    hipEvent_t e = nullptr;
    const hipError_t rc = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (cur != dev) CDN_HIP(hipSetDevice(cur));
    CDN_HIP(rc);
    ev[i][dev] = e;
    return e;
  }
  void ensure() {
    for (int i = 0; i < 2; ++i)
      if (!buf[i]) CDN_HIP(hipHostMalloc(&buf[i], kChunk, hipHostMallocPortable));
    if (!pool) pool = new CopyPool;
  }
  // every DMA through the staging buffers, on any device, has finished
  void drain() {
    for (int i = 0; i < 2; ++i)
      for (auto& kv : ev[i]) CDN_HIP(hipEventSynchronize(kv.second));
  }
};

// process lifetime: the pinned buffers are released when the process ends
Stage& stage() {
  static Stage* s = new Stage;
  return *s;
}

}  // namespace

void copy_to_device(hipStream_t st, void* dev, const void* host, size_t bytes) {
  if (bytes < kDirect) {
    CDN_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
    return;
  }
  Stage& S = stage();
  std::lock_guard<std::mutex> lock(S.mu);
  S.ensure();
  S.drain();
  const size_t n = (bytes + kChunk - 1) / kChunk;
  for (size_t i = 0; i < n; ++i) {
    const int b = (int)(i & 1);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    hipEvent_t e = S.event(b, st);
    if (i >= 2) CDN_HIP(hipEventSynchronize(e));  // the DMA of chunk i - 2 left buf[b]
    S.pool->copy(S.buf[b], static_cast<const char*>(host) + off, len);
    CDN_HIP(hipMemcpyAsync(static_cast<char*>(dev) + off, S.buf[b], len, hipMemcpyHostToDevice, st));
    CDN_HIP(hipEventRecord(e, st));
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
  S.drain();
  const size_t n = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t i) {
    const int b = (int)(i & 1);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    CDN_HIP(hipMemcpyAsync(S.buf[b], static_cast<const char*>(dev) + off, len, hipMemcpyDeviceToHost, st));
    CDN_HIP(hipEventRecord(S.event(b, st), st));
  };
  issue(0);
  for (size_t i = 0; i < n; ++i) {
    // buf[(i + 1) & 1] was emptied by the host copy of chunk i - 1
    if (i + 1 < n) issue(i + 1);
    const int b = (int)(i & 1);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    CDN_HIP(hipEventSynchronize(S.event(b, st)));
    S.pool->copy(static_cast<char*>(host) + off, S.buf[b], len);
  }
}

}  // namespace cdn
