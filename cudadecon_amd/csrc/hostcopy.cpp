// Staged host <-> device copies (see hostcopy.hpp).
#include "hostcopy.hpp"

#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace cdn {
namespace {

constexpr size_t kChunk = 32UL << 20;   // bytes per staging buffer
constexpr size_t kDirect = 16UL << 20;  // smaller copies go straight through the runtime

// memcpy split over host threads (slices of at least 2 MB, at most 8 threads)
void par_memcpy(void* dst, const void* src, size_t n) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>({8, hw, std::max<size_t>(1, n / (2UL << 20))});
  if (nt <= 1) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t per = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  for (size_t t = 1; t < nt; ++t) {
    const size_t a = t * per;
    if (a >= n) break;
    const size_t len = std::min(per, n - a);
    th.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, len); });
  }
  std::memcpy(dst, src, std::min(per, n));
  for (auto& t : th) t.join();
}

struct Stage {
  std::mutex mu;
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
    par_memcpy(S.buf[b], static_cast<const char*>(host) + off, len);
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
    par_memcpy(static_cast<char*>(host) + off, S.buf[b], len);
  }
}

}  // namespace cdn
