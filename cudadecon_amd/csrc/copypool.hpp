// A small pool of host threads for large memcpy's (no HIP dependency).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace cdn {

// memcpy split over host threads (slices of at least 2 MB, at most 8
// threads, the caller's included).  The workers start with the pool and wait
// for the next job until it is destroyed.  Starting threads for every 32 MB
// chunk instead halved the rate of a chunked memcpy on an 8-core host (22
// against 40 GB/s); on the MI355X host the staged copies are bound by the DMA
// either way (RL_interface at 256x1024x1024: 0.099 s end to end with both).
// One job at a time: callers serialise copy() themselves.
class CopyPool {
 public:
  CopyPool() {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned k = std::min(8u, hw);
    th_.reserve(k);
    for (unsigned w = 1; w < k; ++w) {
      try {
        th_.emplace_back([this, w] { run(w); });
      } catch (...) {
        break;  // fewer workers; copy() splits over the ones that started
      }
    }
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> l(mu_);
      stop_ = true;
    }
    go_.notify_all();
    for (auto& t : th_) t.join();
  }
  CopyPool(const CopyPool&) = delete;
  CopyPool& operator=(const CopyPool&) = delete;
  size_t threads() const { return th_.size() + 1; }
  void copy(void* dst, const void* src, size_t n) {
    const size_t nt = std::min<size_t>(th_.size() + 1, std::max<size_t>(1, n / (2UL << 20)));
    if (nt <= 1) {
      std::memcpy(dst, src, n);
      return;
    }
    const size_t per = (n + nt - 1) / nt;
    {
      std::lock_guard<std::mutex> l(mu_);
      dst_ = static_cast<char*>(dst);
      src_ = static_cast<const char*>(src);
      n_ = n;
      per_ = per;
      nt_ = nt;
      pending_ = nt - 1;
      ++gen_;
    }
    go_.notify_all();
    std::memcpy(dst, src, std::min(per, n));
    std::unique_lock<std::mutex> l(mu_);
    done_.wait(l, [&] { return pending_ == 0; });
  }

 private:
  void run(size_t w) {
    unsigned long seen = 0;
    std::unique_lock<std::mutex> l(mu_);
    for (;;) {
      go_.wait(l, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      if (w >= nt_) continue;  // this job has fewer slices
      char* d = dst_;
      const char* s = src_;
      const size_t a = w * per_, n = n_, per = per_;
      l.unlock();
      if (a < n) std::memcpy(d + a, s + a, std::min(per, n - a));
      l.lock();
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::mutex mu_;
  std::condition_variable go_, done_;
  std::vector<std::thread> th_;
  unsigned long gen_ = 0;
  bool stop_ = false;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t n_ = 0, per_ = 0, nt_ = 0, pending_ = 0;
};

}  // namespace cdn
