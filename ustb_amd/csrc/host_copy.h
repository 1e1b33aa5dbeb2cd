// host_copy.h — host-only helper of the host-buffer entry: a memcpy split
// over a small persistent thread pool (no HIP; also compiled on its own with
// the host sanitizers by the CPU tests).
#pragma once
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace ustbx {

// threads - 1 workers; copy() gives the caller the first part and the workers
// the others, and returns when all parts are done. Copies below 1 MiB run on
// the caller alone. One pooled copy at a time (concurrent callers queue).
class CopyPool {
 public:
  explicit CopyPool(int threads) : threads_(std::max(1, threads)) {
    for (int i = 1; i < threads_; ++i) workers_.emplace_back([this, i] { run(i); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> l(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  CopyPool(const CopyPool&) = delete;
  CopyPool& operator=(const CopyPool&) = delete;

  void copy(void* dst, const void* src, size_t n) {
    if (n < ((size_t)1 << 20) || threads_ == 1) {
      memcpy(dst, src, n);
      return;
    }
    std::lock_guard<std::mutex> one(use_);
    {
      std::lock_guard<std::mutex> l(m_);
      dst_ = (char*)dst;
      src_ = (const char*)src;
      n_ = n;
      pending_ = threads_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    part(0, (char*)dst, (const char*)src, n);
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [this] { return pending_ == 0; });
  }

 private:
  // part i of n bytes: ceil(n / threads) rounded up to 64 bytes per part, so
  // that the parts cover all n bytes
  void part(int i, char* dst, const char* src, size_t n) const {
    const size_t per = ((n + threads_ - 1) / threads_ + 63) & ~size_t(63);
    const size_t a = std::min(n, per * i), b = std::min(n, per * (i + 1));
    if (b > a) memcpy(dst + a, src + a, b - a);
  }
  void run(int i) {
    unsigned seen = 0;
    for (;;) {
      char* dst;
      const char* src;
      size_t n;
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        dst = dst_;
        src = src_;
        n = n_;
      }
      part(i, dst, src, n);
      {
        std::lock_guard<std::mutex> l(m_);
        if (--pending_ == 0) done_.notify_one();
      }
    }
  }

  const int threads_;
  std::vector<std::thread> workers_;
  std::mutex use_, m_;
  std::condition_variable cv_, done_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t n_ = 0;
  int pending_ = 0;
  unsigned gen_ = 0;
  bool stop_ = false;
};

}  // namespace ustbx
