// fm_copy.h -- host-side copy helpers shared by the plain host-buffer ops
// and the serving loop.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace fm {

// True when [p, p + bytes) lies inside one allocation made by fm_host_alloc
// (pinned: the DMA engines can move it without a host copy).
bool host_pinned(const void* p, uint64_t bytes);

// Host copy workers.  Buffers in ordinary (pageable) memory go through a
// pinned staging buffer; one thread moves them at a few GB/s, so large
// copies are cut into pieces that the workers and the caller share.
// Buffers from fm_host_alloc skip this: the DMA engines move them directly.
class CopyPool {
 public:
  struct Span {
    uint8_t* dst;
    const uint8_t* src;
    uint64_t bytes;
  };
  explicit CopyPool(unsigned workers) {
    for (unsigned i = 0; i < workers; ++i) threads_.emplace_back([this] { loop(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> l(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (std::thread& t : threads_) t.join();
  }
  // Copies every span; returns when all bytes are in place.
  void run(const std::vector<Span>& spans) {
    if (spans.empty()) return;  // nothing to copy: wake no worker
    constexpr uint64_t kPiece = 256 << 10;
    Call call;
    {
      std::lock_guard<std::mutex> l(mu_);
      for (const Span& sp : spans)
        for (uint64_t o = 0; o < sp.bytes; o += kPiece) {
          jobs_.push_back(Job{sp.dst + o, sp.src + o, std::min(kPiece, sp.bytes - o), &call});
          ++call.left;
        }
    }
    cv_.notify_all();
    for (;;) {  // the caller copies too
      Job j;
      {
        std::lock_guard<std::mutex> l(mu_);
        if (jobs_.empty()) break;
        j = jobs_.front();
        jobs_.pop_front();
      }
      do_job(j);
    }
    std::unique_lock<std::mutex> l(mu_);
    done_cv_.wait(l, [&] { return call.left == 0; });
  }

 private:
  struct Call {
    uint64_t left = 0;  // pieces not yet copied (under mu_)
  };
  struct Job {
    uint8_t* dst = nullptr;
    const uint8_t* src = nullptr;
    uint64_t bytes = 0;
    Call* call = nullptr;
  };
  void do_job(const Job& j) {
    std::memcpy(j.dst, j.src, j.bytes);
    std::lock_guard<std::mutex> l(mu_);
    if (--j.call->left == 0) done_cv_.notify_all();
  }
  void loop() {
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [&] { return stop_ || !jobs_.empty(); });
        if (jobs_.empty()) return;
        j = jobs_.front();
        jobs_.pop_front();
      }
      do_job(j);
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<Job> jobs_;
  bool stop_ = false;
  std::vector<std::thread> threads_;
};

}  // namespace fm
