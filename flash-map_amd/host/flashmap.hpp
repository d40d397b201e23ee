// flashmap.hpp -- C++ mirror of the reference's FlashMap API over the C ABI.
//
// FlashMap<K,V>       src/lib.rs:94-351   (host-facing and device-resident API)
// FlashMapBuilder     src/lib.rs:377-449  (GPU first, CPU fallback)
// HashStrategy        src/hash.rs:3-27
// FlashMapError       src/error.rs:4-39   (thrown; code() is the fm_status)
// AsyncFlashMap<K,V>  src/async_map.rs:32-147 (worker thread per op, shared
//                     lock for gets, exclusive lock for mutations)
//
// K and V must be trivially copyable (bytemuck::Pod); they cross the ABI as
// raw bytes, so padding bytes are part of the key.
#pragma once

#include <cstdint>
#include <cstring>
#include <future>
#include <memory>
#include <optional>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/flashmap.h"

namespace flash_map {

enum class HashStrategy : uint32_t { Identity = FM_HASH_IDENTITY, Murmur3 = FM_HASH_MURMUR3 };
enum class Backend : int { None = FM_BACKEND_NONE, Gpu = FM_BACKEND_GPU, Cpu = FM_BACKEND_CPU };

class FlashMapError : public std::runtime_error {
 public:
  FlashMapError(fm_status code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  fm_status code() const { return code_; }

 private:
  fm_status code_;
};

// FlashMapError::TableFull { occupied, capacity, load_factor (percent) }.
class TableFull : public FlashMapError {
 public:
  TableFull(const std::string& msg, uint64_t occ, uint64_t cap, double lf)
      : FlashMapError(FM_ERR_TABLE_FULL, msg), occupied(occ), capacity(cap), load_factor(lf) {}
  uint64_t occupied;
  uint64_t capacity;
  double load_factor;
};

namespace detail {
inline void check(fm_status st) {
  if (st == FM_OK) return;
  std::string msg(fm_last_error_message(nullptr, 0) + 1, '\0');
  fm_last_error_message(&msg[0], msg.size());
  msg.resize(msg.size() - 1);
  if (st == FM_ERR_TABLE_FULL) {
    uint64_t occ = 0, cap = 0;
    double lf = 0;
    fm_last_table_full(&occ, &cap, &lf);
    throw TableFull(msg, occ, cap, lf);
  }
  throw FlashMapError(st, msg);
}
struct MapDeleter {
  void operator()(fm_map* m) const { fm_destroy(m); }
};
}  // namespace detail

template <class K, class V>
class FlashMap;

class FlashMapBuilder {
 public:
  explicit FlashMapBuilder(uint64_t capacity) { fm_builder_init(&b_, capacity, 0, 0); }
  FlashMapBuilder& hash_strategy(HashStrategy s) {
    b_.hash_strategy = static_cast<uint32_t>(s);
    return *this;
  }
  FlashMapBuilder& device_id(int id) {
    b_.device_id = id;
    return *this;
  }
  FlashMapBuilder& force_cpu() {
    b_.force_cpu = 1;
    return *this;
  }
  FlashMapBuilder& require_gpu() {
    b_.require_gpu = 1;
    return *this;
  }
  FlashMapBuilder& quiet() {
    b_.quiet = 1;
    return *this;
  }
  template <class K, class V>
  FlashMap<K, V> build() const;

 private:
  fm_builder b_;
};

// A device allocation owned by a map (CudaSlice<u8> in the reference).
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(fm_map* owner, void* p, uint64_t bytes) : owner_(owner), p_(p), bytes_(bytes) {}
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept { *this = std::move(o); }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    std::swap(owner_, o.owner_);
    std::swap(p_, o.p_);
    std::swap(bytes_, o.bytes_);
    return *this;
  }
  ~DeviceBuffer() {
    if (p_) fm_free_device(owner_, p_);
  }
  void* get() const { return p_; }
  uint64_t size() const { return bytes_; }

 private:
  fm_map* owner_ = nullptr;
  void* p_ = nullptr;
  uint64_t bytes_ = 0;
};

template <class K, class V>
class FlashMap {
  static_assert(std::is_trivially_copyable<K>::value && std::is_trivially_copyable<V>::value,
                "FlashMap keys and values must be plain old data");

 public:
  explicit FlashMap(fm_map* h) : h_(h) {}

  // lib.rs:110-117
  static FlashMap with_capacity(uint64_t capacity) { return FlashMapBuilder(capacity).build<K, V>(); }
  static FlashMapBuilder builder(uint64_t capacity) { return FlashMapBuilder(capacity); }

  // ---- host-facing API (lib.rs:123-207) ----
  std::vector<std::optional<V>> bulk_get(const std::vector<K>& keys) const {
    std::vector<std::optional<V>> out(keys.size());
    if (keys.empty()) return out;
    std::vector<V> vals(keys.size());
    std::vector<uint8_t> found(keys.size());
    detail::check(fm_bulk_get(h_.get(), keys.data(), keys.size(), vals.data(), found.data()));
    for (size_t i = 0; i < keys.size(); ++i)
      if (found[i]) out[i] = vals[i];
    return out;
  }

  // Returns the number of newly occupied slots; duplicates resolve as the
  // sequential CPU path does (last value wins).
  size_t bulk_insert(const std::vector<std::pair<K, V>>& pairs) {
    std::vector<K> ks(pairs.size());
    std::vector<V> vs(pairs.size());
    for (size_t i = 0; i < pairs.size(); ++i) {  // gpu.rs:212-217
      ks[i] = pairs[i].first;
      vs[i] = pairs[i].second;
    }
    uint64_t added = 0;
    detail::check(fm_bulk_insert(h_.get(), ks.data(), vs.data(), pairs.size(), &added));
    return static_cast<size_t>(added);
  }

  size_t bulk_remove(const std::vector<K>& keys) {
    uint64_t removed = 0;
    detail::check(fm_bulk_remove(h_.get(), keys.data(), keys.size(), &removed));
    return static_cast<size_t>(removed);
  }

  size_t len() const { return static_cast<size_t>(fm_len(h_.get())); }
  bool is_empty() const { return len() == 0; }
  size_t capacity() const { return static_cast<size_t>(fm_capacity(h_.get())); }
  double load_factor() const { return fm_load_factor(h_.get()); }
  void clear() { detail::check(fm_clear(h_.get())); }
  Backend backend() const { return static_cast<Backend>(fm_backend(h_.get())); }

  // impl Debug (lib.rs:353-370)
  std::string debug_string() const {
    char buf[256];
    detail::check(fm_debug_string(h_.get(), buf, sizeof buf));
    return buf;
  }

  // ---- device-resident API (lib.rs:213-350) ----
  void* device_stream() const { return fm_device_stream(h_.get()); }
  DeviceBuffer alloc_device(uint64_t bytes) const {
    void* p = nullptr;
    detail::check(fm_alloc_device(h_.get(), bytes, &p));
    return DeviceBuffer(h_.get(), p, bytes);
  }
  DeviceBuffer upload_keys(const std::vector<K>& keys) const { return upload(keys.data(), keys.size() * sizeof(K)); }
  DeviceBuffer upload_values(const std::vector<V>& vals) const {
    return upload(vals.data(), vals.size() * sizeof(V));
  }
  std::vector<uint8_t> download(const DeviceBuffer& b) const {
    std::vector<uint8_t> out(b.size());
    detail::check(fm_download(h_.get(), b.get(), b.size(), out.data()));
    return out;
  }
  // Returns (values, found) device buffers as gpu.rs:422-478 does.
  std::pair<DeviceBuffer, DeviceBuffer> bulk_get_device(const DeviceBuffer& keys, uint64_t count) const {
    DeviceBuffer vals = alloc_device(count * sizeof(V));
    DeviceBuffer found = alloc_device(count);
    detail::check(fm_bulk_get_device(h_.get(), keys.get(), count, vals.get(), static_cast<uint8_t*>(found.get())));
    return {std::move(vals), std::move(found)};
  }
  DeviceBuffer bulk_get_values_device(const DeviceBuffer& keys, uint64_t count) const {
    DeviceBuffer vals = alloc_device(count * sizeof(V));
    detail::check(fm_bulk_get_values_device(h_.get(), keys.get(), count, vals.get()));
    return vals;
  }
  size_t bulk_insert_device(const DeviceBuffer& keys, const DeviceBuffer& vals, uint64_t count) {
    uint64_t added = 0;
    detail::check(fm_bulk_insert_device(h_.get(), keys.get(), vals.get(), count, &added));
    return static_cast<size_t>(added);
  }
  void bulk_insert_device_uncounted(const DeviceBuffer& keys, const DeviceBuffer& vals, uint64_t count) {
    detail::check(fm_bulk_insert_device_uncounted(h_.get(), keys.get(), vals.get(), count));
  }
  size_t recount() {
    uint64_t n = 0;
    detail::check(fm_recount(h_.get(), &n));
    return static_cast<size_t>(n);
  }

  fm_map* raw() const { return h_.get(); }

 private:
  DeviceBuffer upload(const void* p, uint64_t bytes) const {
    void* d = nullptr;
    detail::check(fm_upload(h_.get(), p, bytes, &d));
    return DeviceBuffer(h_.get(), d, bytes);
  }
  std::unique_ptr<fm_map, detail::MapDeleter> h_;
};

template <class K, class V>
FlashMap<K, V> FlashMapBuilder::build() const {
  fm_builder b = b_;
  b.key_size = sizeof(K);
  b.value_size = sizeof(V);
  fm_map* h = nullptr;
  detail::check(fm_build(&b, &h));
  return FlashMap<K, V>(h);
}

// src/async_map.rs:32-147: each op runs on its own thread (spawn_blocking);
// gets hold a shared lock, mutations an exclusive one.  Copies share state.
template <class K, class V>
class AsyncFlashMap {
 public:
  explicit AsyncFlashMap(FlashMap<K, V>&& map) : s_(std::make_shared<State>(std::move(map))) {}

  std::future<std::vector<std::optional<V>>> bulk_get(std::vector<K> keys) const {
    auto s = s_;
    return std::async(std::launch::async, [s, keys = std::move(keys)] {
      std::shared_lock<std::shared_mutex> g(s->mu);
      return s->map.bulk_get(keys);
    });
  }
  std::future<size_t> bulk_insert(std::vector<std::pair<K, V>> pairs) const {
    auto s = s_;
    return std::async(std::launch::async, [s, pairs = std::move(pairs)] {
      std::unique_lock<std::shared_mutex> g(s->mu);
      return s->map.bulk_insert(pairs);
    });
  }
  std::future<size_t> bulk_remove(std::vector<K> keys) const {
    auto s = s_;
    return std::async(std::launch::async, [s, keys = std::move(keys)] {
      std::unique_lock<std::shared_mutex> g(s->mu);
      return s->map.bulk_remove(keys);
    });
  }
  std::future<void> clear() const {
    auto s = s_;
    return std::async(std::launch::async, [s] {
      std::unique_lock<std::shared_mutex> g(s->mu);
      s->map.clear();
    });
  }
  size_t len() const {
    std::shared_lock<std::shared_mutex> g(s_->mu);
    return s_->map.len();
  }
  bool is_empty() const { return len() == 0; }
  size_t capacity() const {
    std::shared_lock<std::shared_mutex> g(s_->mu);
    return s_->map.capacity();
  }
  double load_factor() const {
    std::shared_lock<std::shared_mutex> g(s_->mu);
    return s_->map.load_factor();
  }

 private:
  struct State {
    explicit State(FlashMap<K, V>&& m) : map(std::move(m)) {}
    std::shared_mutex mu;
    FlashMap<K, V> map;
  };
  std::shared_ptr<State> s_;
};

}  // namespace flash_map
