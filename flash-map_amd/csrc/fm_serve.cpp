// fm_serve.cpp -- continuous-batching serving loop over one GPU map.
//
// Threads:
//   arrival    releases request i at start + arrival_s[i] (or all at once);
//   dispatcher takes released requests strictly in arrival order, coalesces
//              adjacent requests of the same kind into one device batch,
//              admits the batch when its staging fits in the ring (waiting
//              for completions otherwise), copies it host->device on the copy
//              stream, runs the op on the map's stream after that copy, and
//              queues the device->host copy of the outputs on a third stream;
//   completion waits for each batch's outputs in order, scatters them into
//              the requests' buffers, attributes per-key outcome flags to
//              requests (counts), frees the staging and wakes the dispatcher.
// Batch k+1's copy overlaps batch k's kernels, whose results copy back while
// batch k+2 is staged.
//
// Semantics: arrival-order application.  Coalescing adjacent gets is exact
// (read-only); adjacent inserts (resp. removes) applied as one batch in index
// order equal applying them one after another.  TableFull is decided per
// insert request with the table length before it: len is an upper bound
// while inserts or removes are in flight (an insert raises it by its batch
// size, a remove lowers it only once its outcome flags are back), and a
// request that might not fit waits for them to settle so the exact length
// decides (rayon_cpu.rs:134-142).  No op waits for the device at dispatch:
// a remove's conflict replay is gated on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fm_copy.h"
#include "fm_internal.h"

namespace fm {

// Allocations made by fm_host_alloc (base -> bytes).  A request buffer that
// lies inside one of them is pinned: the DMA engines read or write it
// directly, with no copy through the pinned staging ring.
static std::mutex g_pinned_mu;
static std::map<uintptr_t, uint64_t> g_pinned;

bool host_pinned(const void* p, uint64_t bytes) {
  if (bytes == 0) return true;  // nothing to move
  if (!p) return false;
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  std::lock_guard<std::mutex> l(g_pinned_mu);
  auto it = g_pinned.upper_bound(a);
  if (it == g_pinned.begin()) return false;
  --it;
  return a - it->first <= it->second && bytes <= it->second - (a - it->first);
}

namespace {

using Clock = std::chrono::steady_clock;


inline uint64_t align256(uint64_t b) { return (b + 255) & ~uint64_t(255); }

// FIFO ring allocator: regions are freed in allocation order.
class Ring {
 public:
  explicit Ring(uint64_t size) : size_(size) {}
  bool alloc(uint64_t bytes, uint64_t* off) {
    bytes = align256(std::max<uint64_t>(bytes, 1));
    if (bytes > size_) return false;
    if (live_.empty()) head_ = tail_ = 0;
    const bool wrapped = !live_.empty() && head_ <= tail_;
    if (!wrapped) {
      if (size_ - head_ >= bytes) {
        *off = head_;
      } else if (tail_ >= bytes) {  // wrap; the end of the buffer stays unused
        *off = 0;
      } else {
        return false;
      }
    } else {
      if (tail_ - head_ < bytes) return false;
      *off = head_;
    }
    head_ = *off + bytes;
    live_.push_back(*off);
    return true;
  }
  void free_front() {
    live_.pop_front();
    if (!live_.empty()) tail_ = live_.front();
  }
  uint64_t size() const { return size_; }

 private:
  uint64_t size_, head_ = 0, tail_ = 0;
  std::deque<uint64_t> live_;
};

// Issues hipMemcpyAsync calls, merging a copy into the previous one when
// both its source and its destination continue where the previous ended
// (adjacent requests laid out back to back need one call, not one each:
// a call costs ~20 us of host time).
class CopyMerger {
 public:
  CopyMerger(hipMemcpyKind kind, hipStream_t stream) : kind_(kind), stream_(stream) {}
  void add(void* dst, const void* src, uint64_t bytes) {
    if (bytes == 0 || err_ != hipSuccess) return;
    uint8_t* d = static_cast<uint8_t*>(dst);
    const uint8_t* s = static_cast<const uint8_t*>(src);
    if (n_ && d == dst_ + n_ && s == src_ + n_) {
      n_ += bytes;
      return;
    }
    (void)flush();
    dst_ = d;
    src_ = s;
    n_ = bytes;
  }
  hipError_t flush() {
    if (n_ && err_ == hipSuccess) err_ = hipMemcpyAsync(dst_, src_, n_, kind_, stream_);
    n_ = 0;
    return err_;
  }

 private:
  hipMemcpyKind kind_;
  hipStream_t stream_;
  uint8_t* dst_ = nullptr;
  const uint8_t* src_ = nullptr;
  uint64_t n_ = 0;
  hipError_t err_ = hipSuccess;
};

struct Batch {
  uint32_t op = 0;
  std::vector<uint64_t> reqs;  // request indices, arrival order
  uint64_t n = 0;              // keys
  uint64_t off = 0;            // staging offset (host ring == device ring)
  uint64_t keys_at = 0, vals_at = 0, out_at = 0, found_at = 0, flags_at = 0;
  hipEvent_t h2d = nullptr, done = nullptr, d2h = nullptr;
  int32_t id = 0;
  bool direct_out = false;  // every request's outputs are pinned: DMA straight into them
};

class Server {
 public:
  Server(fm_map* m, fm_serve_request* reqs, uint64_t n, const fm_serve_config& cfg)
      : m_(m),
        reqs_(reqs),
        n_(n),
        cfg_(cfg),
        ring_(cfg.staging_bytes),
        copies_(kCopyWorkers) {}
  ~Server();
  fm_status run(fm_serve_stats* stats);

 private:
  double now() const { return std::chrono::duration<double>(Clock::now() - t0_).count(); }
  fm_status setup();
  void arrival_loop();
  void completion_loop();
  fm_status dispatch_one(uint64_t first);
  uint64_t batch_bytes(uint32_t op, uint64_t n) const;
  fm_status fail(fm_status c, const std::string& msg) {
    set_error(c, msg);
    return c;
  }

  fm_map* m_;
  fm_serve_request* reqs_;
  uint64_t n_;
  fm_serve_config cfg_;
  Ring ring_;
  uint32_t K_ = 0, V_ = 0;
  uint64_t cap_ = 0;
  int dev_ = 0;
  hipStream_t compute_ = nullptr, h2d_ = nullptr, d2h_ = nullptr;
  uint8_t* d_stage_ = nullptr;
  uint8_t* h_stage_ = nullptr;
  Clock::time_point t0_;

  std::mutex mu_;
  std::condition_variable cv_;
  uint64_t released_ = 0;   // requests released by the arrival thread
  uint64_t next_ = 0;       // next request to dispatch
  std::deque<Batch> inflight_;
  uint64_t inflight_inserts_ = 0;
  uint64_t inflight_removes_ = 0;
  bool stop_ = false;
  fm_status async_err_ = FM_OK;
  uint64_t batches_ = 0, max_inflight_ = 0, staging_waits_ = 0;
  std::vector<uint8_t> pin_in_, pin_out_;  // per request: inputs / outputs pinned
  std::vector<hipEvent_t> events_;          // free events (under mu_), reused across batches
  hipError_t take_event(hipEvent_t* ev) {
    {
      std::lock_guard<std::mutex> l(mu_);
      if (!events_.empty()) {
        *ev = events_.back();
        events_.pop_back();
        return hipSuccess;
      }
    }
    return hipEventCreateWithFlags(ev, hipEventDisableTiming);
  }
  static constexpr unsigned kCopyWorkers = 10;
  CopyPool copies_;
};

Server::~Server() {
  if (compute_ || h2d_ || d2h_) (void)hipDeviceSynchronize();
  for (hipEvent_t ev : events_) (void)hipEventDestroy(ev);
  if (h2d_) (void)hipStreamDestroy(h2d_);
  if (d2h_) (void)hipStreamDestroy(d2h_);
  if (d_stage_) (void)hipFree(d_stage_);
  if (h_stage_) (void)hipHostFree(h_stage_);
}

fm_status Server::setup() {
  Backend& be = *m_->be;
  if (be.kind() != FM_BACKEND_GPU) return gpu_required();
  K_ = be.key_size();
  V_ = be.value_size();
  cap_ = be.capacity();
  dev_ = be.device_id();
  compute_ = static_cast<hipStream_t>(be.device_stream());
  hipError_t e;
  if ((e = hipSetDevice(dev_)) != hipSuccess || (e = hipStreamCreateWithFlags(&h2d_, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&d2h_, hipStreamNonBlocking)) != hipSuccess)
    return fail(FM_ERR_DEVICE_INIT, std::string("serve streams: ") + hipGetErrorString(e));
  if ((e = hipMalloc(&d_stage_, ring_.size())) != hipSuccess ||
      (e = hipHostMalloc(&h_stage_, ring_.size(), hipHostMallocDefault)) != hipSuccess)
    return fail(FM_ERR_GPU_ALLOC, std::string("serve staging: ") + hipGetErrorString(e));
  for (uint64_t i = 0; i < n_; ++i) {
    fm_serve_request& r = reqs_[i];
    if (r.op > FM_OP_REMOVE || (r.n && !r.keys) || (r.op == FM_OP_INSERT && r.n && !r.values) ||
        (r.op == FM_OP_GET && r.n && !r.out_values))
      return fail(FM_ERR_INVALID_ARGUMENT, "fm_serve_log: request " + std::to_string(i) + " is malformed");
    if (batch_bytes(r.op, r.n) > ring_.size())
      return fail(FM_ERR_INVALID_ARGUMENT, "fm_serve_log: request " + std::to_string(i) + " exceeds the staging ring");
    r.count = 0;
    r.status = FM_OK;
    r.batch = -1;
    r.admit_s = r.done_s = 0;
  }
  pin_in_.assign(n_, 0);
  pin_out_.assign(n_, 0);
  for (uint64_t i = 0; i < n_; ++i) {
    const fm_serve_request& r = reqs_[i];
    pin_in_[i] = host_pinned(r.keys, r.n * K_) && (r.op != FM_OP_INSERT || host_pinned(r.values, r.n * V_));
    pin_out_[i] = r.op == FM_OP_GET ? (r.out_found && host_pinned(r.out_values, r.n * V_) && host_pinned(r.out_found, r.n))
                                    : (r.out_flags && host_pinned(r.out_flags, r.n));
  }
  // Size the map's per-batch scratch for the largest batch now, not mid-run.
  uint64_t largest = 0;
  for (uint64_t i = 0; i < n_; ++i) largest = std::max<uint64_t>(largest, reqs_[i].n);
  const uint64_t most = std::max(largest, std::min<uint64_t>(cfg_.max_batch_keys, ring_.size() / (K_ + 1)));
  std::lock_guard<std::mutex> ml(m_->mu);
  return be.serve_reserve(most);
}

uint64_t Server::batch_bytes(uint32_t op, uint64_t n) const {
  uint64_t b = align256(n * K_) + align256(n);  // keys + flags/found
  if (op == FM_OP_INSERT || op == FM_OP_GET) b += align256(n * V_);
  return b;
}

void Server::arrival_loop() {
  for (uint64_t i = 0; i < n_; ++i) {
    if (cfg_.realtime) {
      const auto when = t0_ + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(reqs_[i].arrival_s));
      std::this_thread::sleep_until(when);
    }
    std::lock_guard<std::mutex> l(mu_);
    if (stop_) return;
    released_ = i + 1;
    cv_.notify_all();
  }
}

// Forms, admits and launches the batch that starts at request `first`.
// Called with mu_ unlocked; returns after the batch is queued (or the
// request completed as TableFull).
fm_status Server::dispatch_one(uint64_t first) {
  Backend& be = *m_->be;
  const uint32_t op = reqs_[first].op;
  Batch b;
  b.op = op;
  for (;;) {
    std::unique_lock<std::mutex> l(mu_);
    // Extend over already-released adjacent requests of the same kind.
    b.reqs.clear();
    b.n = 0;
    uint64_t len_upper = 0;
    {
      std::lock_guard<std::mutex> ml(m_->mu);
      len_upper = be.len();
    }
    for (uint64_t i = first; i < released_ && reqs_[i].op == op; ++i) {
      const uint64_t n = reqs_[i].n;
      if (!b.reqs.empty() && b.n + n > cfg_.max_batch_keys) break;
      if (batch_bytes(op, b.n + n) > ring_.size()) break;
      if (op == FM_OP_INSERT && len_upper + b.n + n > cap_) {
        if (!b.reqs.empty()) break;  // decide this one after the batch
        // The table length before this request is needed exactly: wait for
        // every in-flight insert and remove to settle.
        cv_.wait(l, [&] { return (inflight_inserts_ == 0 && inflight_removes_ == 0) || stop_; });
        {
          std::lock_guard<std::mutex> ml(m_->mu);
          len_upper = be.len();
        }
        if (len_upper + n > cap_) {  // TableFull for this request alone
          fm_serve_request& r = reqs_[i];
          set_table_full(len_upper, cap_);
          r.status = FM_ERR_TABLE_FULL;
          r.count = 0;
          r.admit_s = r.done_s = now();
          next_ = i + 1;
          cv_.notify_all();
          return FM_OK;
        }
      }
      b.reqs.push_back(i);
      b.n += n;
    }
    uint64_t off = 0;
    if (ring_.alloc(batch_bytes(op, b.n), &off)) {
      b.off = off;
      next_ = b.reqs.back() + 1;
      if (op == FM_OP_INSERT) ++inflight_inserts_;
      if (op == FM_OP_REMOVE) ++inflight_removes_;
      break;
    }
    ++staging_waits_;
    cv_.wait(l, [&] { return inflight_.empty() || stop_ || released_ > first + b.reqs.size(); });
    if (stop_) return FM_OK;
  }
  // Lay out [keys | values | out | flags-or-found] inside the region.
  b.keys_at = b.off;
  uint64_t at = b.off + align256(b.n * K_);
  if (op != FM_OP_REMOVE) {
    (op == FM_OP_INSERT ? b.vals_at : b.out_at) = at;
    at += align256(b.n * V_);
  }
  (op == FM_OP_GET ? b.found_at : b.flags_at) = at;
  const double admit = now();
  uint64_t pos = 0;
  bool any_pinned_in = false;
  b.direct_out = true;
  std::vector<CopyPool::Span> spans;
  for (uint64_t i : b.reqs) {
    fm_serve_request& r = reqs_[i];
    r.admit_s = admit;
    any_pinned_in = any_pinned_in || pin_in_[i];
    b.direct_out = b.direct_out && pin_out_[i];
    if (!pin_in_[i]) {
      spans.push_back({h_stage_ + b.keys_at + pos * K_, static_cast<const uint8_t*>(r.keys), r.n * K_});
      if (op == FM_OP_INSERT)
        spans.push_back({h_stage_ + b.vals_at + pos * V_, static_cast<const uint8_t*>(r.values), r.n * V_});
    }
    pos += r.n;
  }
  copies_.run(spans);
  b.id = static_cast<int32_t>(batches_++);
  hipError_t e;
  if ((e = take_event(&b.h2d)) != hipSuccess || (e = take_event(&b.done)) != hipSuccess ||
      (e = take_event(&b.d2h)) != hipSuccess)
    return fail(FM_ERR_KERNEL_LAUNCH, std::string("serve events: ") + hipGetErrorString(e));
  if (!any_pinned_in) {  // the whole input region from the staging ring in one copy
    const uint64_t in_bytes = op == FM_OP_INSERT ? (b.vals_at + b.n * V_ - b.keys_at) : b.n * K_;
    e = hipMemcpyAsync(d_stage_ + b.keys_at, h_stage_ + b.keys_at, in_bytes, hipMemcpyHostToDevice, h2d_);
  } else {  // per request: straight from a pinned buffer, else from the staging ring
    CopyMerger keys(hipMemcpyHostToDevice, h2d_), vals(hipMemcpyHostToDevice, h2d_);
    pos = 0;
    for (uint64_t i : b.reqs) {
      const fm_serve_request& r = reqs_[i];
      keys.add(d_stage_ + b.keys_at + pos * K_,
               pin_in_[i] ? static_cast<const uint8_t*>(r.keys) : h_stage_ + b.keys_at + pos * K_, r.n * K_);
      if (op == FM_OP_INSERT)
        vals.add(d_stage_ + b.vals_at + pos * V_,
                 pin_in_[i] ? static_cast<const uint8_t*>(r.values) : h_stage_ + b.vals_at + pos * V_, r.n * V_);
      pos += r.n;
    }
    if ((e = keys.flush()) == hipSuccess) e = vals.flush();
  }
  if (e != hipSuccess || (e = hipEventRecord(b.h2d, h2d_)) != hipSuccess ||
      (e = hipStreamWaitEvent(compute_, b.h2d, 0)) != hipSuccess)
    return fail(FM_ERR_TRANSFER, std::string("serve H2D: ") + hipGetErrorString(e));
  fm_status st;
  {
    std::lock_guard<std::mutex> ml(m_->mu);
    if (op == FM_OP_GET)
      st = be.serve_get(d_stage_ + b.keys_at, b.n, d_stage_ + b.out_at, d_stage_ + b.found_at);
    else if (op == FM_OP_INSERT)
      st = be.serve_insert(d_stage_ + b.keys_at, d_stage_ + b.vals_at, b.n, d_stage_ + b.flags_at);
    else
      st = be.serve_remove(d_stage_ + b.keys_at, b.n, d_stage_ + b.flags_at);
  }
  if (st != FM_OK) return st;
  const uint64_t out_from = op == FM_OP_GET ? b.out_at : b.flags_at;
  const uint64_t out_bytes = op == FM_OP_GET ? (b.found_at + b.n - b.out_at) : b.n;
  if ((e = hipEventRecord(b.done, compute_)) == hipSuccess) e = hipStreamWaitEvent(d2h_, b.done, 0);
  if (!b.direct_out) {  // into the staging ring; the completion thread scatters
    if (e == hipSuccess)
      e = hipMemcpyAsync(h_stage_ + out_from, d_stage_ + out_from, out_bytes, hipMemcpyDeviceToHost, d2h_);
  } else if (e == hipSuccess) {  // straight into each request's pinned outputs
    CopyMerger vals(hipMemcpyDeviceToHost, d2h_), flags(hipMemcpyDeviceToHost, d2h_);
    pos = 0;
    for (uint64_t i : b.reqs) {
      const fm_serve_request& r = reqs_[i];
      if (op == FM_OP_GET) {
        vals.add(r.out_values, d_stage_ + b.out_at + pos * V_, r.n * V_);
        flags.add(r.out_found, d_stage_ + b.found_at + pos, r.n);
      } else {
        flags.add(r.out_flags, d_stage_ + b.flags_at + pos, r.n);
      }
      pos += r.n;
    }
    if ((e = vals.flush()) == hipSuccess) e = flags.flush();
  }
  if (e != hipSuccess || (e = hipEventRecord(b.d2h, d2h_)) != hipSuccess)
    return fail(FM_ERR_TRANSFER, std::string("serve D2H: ") + hipGetErrorString(e));
  std::lock_guard<std::mutex> l(mu_);
  inflight_.push_back(b);
  max_inflight_ = std::max<uint64_t>(max_inflight_, inflight_.size());
  cv_.notify_all();
  return FM_OK;
}

void Server::completion_loop() {
  for (;;) {
    Batch b;
    {
      std::unique_lock<std::mutex> l(mu_);
      cv_.wait(l, [&] { return !inflight_.empty() || stop_; });
      if (inflight_.empty()) return;
      b = inflight_.front();
    }
    hipError_t e = hipEventSynchronize(b.d2h);
    uint64_t pos = 0, settled = 0;
    std::vector<CopyPool::Span> spans;
    for (uint64_t i : b.reqs) {
      if (b.direct_out) break;  // the DMA wrote every request's outputs
      fm_serve_request& r = reqs_[i];
      if (b.op == FM_OP_GET) {
        spans.push_back({static_cast<uint8_t*>(r.out_values), h_stage_ + b.out_at + pos * V_, r.n * V_});
        if (r.out_found) spans.push_back({r.out_found, h_stage_ + b.found_at + pos, r.n});
      } else if (r.out_flags) {
        spans.push_back({r.out_flags, h_stage_ + b.flags_at + pos, r.n});
      }
      pos += r.n;
    }
    copies_.run(spans);
    const double done = now();
    pos = 0;
    for (uint64_t i : b.reqs) {
      fm_serve_request& r = reqs_[i];
      uint64_t count = 0;
      const uint8_t* f = b.direct_out ? (b.op == FM_OP_GET ? r.out_found : r.out_flags)
                                      : h_stage_ + (b.op == FM_OP_GET ? b.found_at : b.flags_at) + pos;
      for (uint64_t k = 0; k < r.n; ++k) count += f[k];
      settled += count;
      r.count = count;
      r.batch = b.id;
      r.done_s = done;
      pos += r.n;
    }
    if (b.op == FM_OP_INSERT) {
      std::lock_guard<std::mutex> ml(m_->mu);
      m_->be->serve_settle_insert(b.n, settled);
    } else if (b.op == FM_OP_REMOVE) {
      std::lock_guard<std::mutex> ml(m_->mu);
      m_->be->serve_settle_remove(settled);
    }
    std::lock_guard<std::mutex> l(mu_);
    events_.push_back(b.h2d);
    events_.push_back(b.done);
    events_.push_back(b.d2h);
    if (e != hipSuccess && async_err_ == FM_OK) {
      async_err_ = FM_ERR_TRANSFER;
      set_error(FM_ERR_TRANSFER, std::string("serve completion: ") + hipGetErrorString(e));
    }
    inflight_.pop_front();
    ring_.free_front();
    if (b.op == FM_OP_INSERT) --inflight_inserts_;
    if (b.op == FM_OP_REMOVE) --inflight_removes_;
    cv_.notify_all();
  }
}

fm_status Server::run(fm_serve_stats* stats) {
  fm_status st = setup();
  if (st != FM_OK) return st;
  struct ServeMode {  // hot keys are expected: inserts deduplicate while serving
    Backend& be;
    explicit ServeMode(Backend& b) : be(b) { be.serve_mode(true); }
    ~ServeMode() { be.serve_mode(false); }
  } mode(*m_->be);
  t0_ = Clock::now();
  std::thread arrivals(&Server::arrival_loop, this);
  std::thread completions(&Server::completion_loop, this);
  fm_status err = FM_OK;
  for (;;) {
    uint64_t first;
    {
      std::unique_lock<std::mutex> l(mu_);
      cv_.wait(l, [&] { return next_ < released_ || next_ >= n_; });
      if (next_ >= n_) break;
      first = next_;
    }
    if ((err = dispatch_one(first)) != FM_OK) break;
  }
  {
    std::unique_lock<std::mutex> l(mu_);
    if (err != FM_OK) stop_ = true;
    cv_.wait(l, [&] { return inflight_.empty() || stop_; });
    stop_ = true;
    cv_.notify_all();
  }
  arrivals.join();
  completions.join();
  if (err == FM_OK) err = async_err_;
  if (err != FM_OK) return err;
  if (stats) {
    std::vector<double> lat;
    lat.reserve(n_);
    double first_arrival = 1e300, last_done = 0, sum = 0;
    uint64_t keys = 0;
    for (uint64_t i = 0; i < n_; ++i) {
      const double l = (reqs_[i].done_s - (cfg_.realtime ? reqs_[i].arrival_s : 0.0)) * 1e3;
      lat.push_back(l);
      sum += l;
      keys += reqs_[i].n;
      first_arrival = std::min(first_arrival, cfg_.realtime ? reqs_[i].arrival_s : 0.0);
      last_done = std::max(last_done, reqs_[i].done_s);
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      if (lat.empty()) return 0.0;
      const size_t k = static_cast<size_t>(std::min<double>(lat.size() - 1, std::ceil(p * lat.size()) - 1));
      return lat[k];
    };
    stats->requests = n_;
    stats->batches = batches_;
    stats->keys = keys;
    stats->wall_s = last_done - (n_ ? first_arrival : 0);
    stats->p50_ms = pct(0.50);
    stats->p99_ms = pct(0.99);
    stats->mean_ms = n_ ? sum / n_ : 0;
    stats->max_ms = lat.empty() ? 0 : lat.back();
    stats->max_inflight_batches = max_inflight_;
    stats->staging_waits = staging_waits_;
  }
  return FM_OK;
}

}  // namespace
}  // namespace fm

extern "C" fm_status fm_serve_log(fm_map* m, fm_serve_request* reqs, uint64_t n_reqs, const fm_serve_config* cfg,
                                  fm_serve_stats* stats) {
  if (!m || (!reqs && n_reqs) || !cfg || cfg->staging_bytes < 4096 || cfg->max_batch_keys == 0) {
    fm::set_error(FM_ERR_INVALID_ARGUMENT, "fm_serve_log: bad arguments");
    return FM_ERR_INVALID_ARGUMENT;
  }
  fm::Server s(m, reqs, n_reqs, *cfg);
  return s.run(stats);
}

extern "C" fm_status fm_host_alloc(uint64_t bytes, void** out) {
  if (!out || bytes == 0) {
    fm::set_error(FM_ERR_INVALID_ARGUMENT, "fm_host_alloc: bad arguments");
    return FM_ERR_INVALID_ARGUMENT;
  }
  void* p = nullptr;
  const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    fm::set_error(FM_ERR_GPU_ALLOC, std::string("fm_host_alloc: ") + hipGetErrorString(e));
    return FM_ERR_GPU_ALLOC;
  }
  {
    std::lock_guard<std::mutex> l(fm::g_pinned_mu);
    fm::g_pinned[reinterpret_cast<uintptr_t>(p)] = bytes;
  }
  *out = p;
  return FM_OK;
}

extern "C" fm_status fm_host_free(void* p) {
  if (!p) return FM_OK;
  {
    std::lock_guard<std::mutex> l(fm::g_pinned_mu);
    auto it = fm::g_pinned.find(reinterpret_cast<uintptr_t>(p));
    if (it == fm::g_pinned.end()) {
      fm::set_error(FM_ERR_INVALID_ARGUMENT, "fm_host_free: not a pointer from fm_host_alloc");
      return FM_ERR_INVALID_ARGUMENT;
    }
    fm::g_pinned.erase(it);
  }
  const hipError_t e = hipHostFree(p);
  if (e != hipSuccess) {
    fm::set_error(FM_ERR_GPU_ALLOC, std::string("fm_host_free: ") + hipGetErrorString(e));
    return FM_ERR_GPU_ALLOC;
  }
  return FM_OK;
}
