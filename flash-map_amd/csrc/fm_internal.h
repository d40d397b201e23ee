// fm_internal.h -- backend interface behind the C ABI (include/flashmap.h).
//
// Mirrors the reference's backend seam: enum FlashMapBackend { Gpu, Rayon }
// matched by every FlashMap method (src/lib.rs:98-103, :124-350).  Here the
// arms are classes implementing Backend; fm_map owns one.
#pragma once

#include <stdint.h>

#include <memory>
#include <mutex>
#include <string>

#include "../../include/flashmap.h"

namespace fm {

// Thread-local error detail (Display text of FlashMapError, error.rs:4-39).
void set_error(fm_status code, const std::string& msg);
void set_table_full(uint64_t occupied, uint64_t capacity);
fm_status gpu_required();

class Backend {
 public:
  virtual ~Backend() = default;
  virtual fm_backend_kind kind() const = 0;
  virtual uint64_t capacity() const = 0;
  virtual uint64_t len() const = 0;
  virtual uint32_t key_size() const = 0;
  virtual uint32_t value_size() const = 0;

  virtual fm_status bulk_get(const void* keys, uint64_t n, void* out_values, uint8_t* out_found) = 0;
  virtual fm_status bulk_insert(const void* keys, const void* values, uint64_t n, uint64_t* out_new) = 0;
  virtual fm_status bulk_remove(const void* keys, uint64_t n, uint64_t* out_removed) = 0;
  virtual fm_status clear() = 0;
  virtual fm_status export_slots(uint32_t* flags, void* keys, void* values) = 0;
  virtual fm_status import_slots(const uint32_t* flags, const void* keys, const void* values, uint64_t len) = 0;

  // Device-resident family (GPU only; lib.rs:235 returns GpuRequired).
  virtual void* device_stream() { return nullptr; }
  virtual int32_t device_id() const { return -1; }
  virtual fm_status alloc_device(uint64_t, void**) { return gpu_required(); }
  virtual fm_status free_device(void*) { return gpu_required(); }
  virtual fm_status upload(const void*, uint64_t, void**) { return gpu_required(); }
  virtual fm_status download(const void*, uint64_t, void*) { return gpu_required(); }
  virtual fm_status bulk_get_device(const void*, uint64_t, void*, uint8_t*) { return gpu_required(); }
  virtual fm_status bulk_insert_device(const void*, const void*, uint64_t, uint64_t*, bool /*counted*/) {
    return gpu_required();
  }  virtual fm_status bulk_insert_device_noalloc(const void*, const void*, uint64_t, uint64_t*) {
    return gpu_required();
  }

  virtual fm_status bulk_remove_device(const void*, uint64_t, uint64_t*) { return gpu_required(); }
  virtual fm_status recount(uint64_t*) { return gpu_required(); }
  virtual fm_status synchronize() { return FM_OK; }
  virtual fm_status copy_from(Backend&, uint32_t /*parts*/) { return gpu_required(); }

  // Serving-loop hooks (GPU only).  Device buffers; work is ordered on the
  // map's stream and none of them waits for it.  serve_insert raises len by
  // n (an upper bound, so len never undercounts occupied slots) and
  // serve_settle_insert lowers it by the keys that did not claim a slot once
  // their outcome flags are known; serve_remove leaves len alone and
  // serve_settle_remove lowers it by the removed count read from the flags.
  virtual fm_status serve_get(const void*, uint64_t, void*, uint8_t*) { return gpu_required(); }
  virtual fm_status serve_insert(const void*, const void*, uint64_t, uint8_t*) { return gpu_required(); }
  virtual void serve_settle_insert(uint64_t, uint64_t) {}
  virtual void serve_mode(bool) {}  // inside fm_serve_log: inserts deduplicate hot keys
  virtual fm_status serve_reserve(uint64_t /*max_keys*/) { return FM_OK; }  // per-batch scratch up front
  virtual fm_status serve_remove(const void*, uint64_t, uint8_t*) { return gpu_required(); }
  virtual void serve_settle_remove(uint64_t) {}
  virtual fm_status profile_enable(bool) { return gpu_required(); }
  virtual fm_status profile_only(const char*) { return gpu_required(); }
  virtual fm_status profile_read(fm_kernel_stat*, uint32_t, uint32_t*) { return gpu_required(); }
  virtual fm_status get_tuning(fm_tuning*) const { return gpu_required(); }
  virtual fm_status set_tuning(const fm_tuning*) { return gpu_required(); }
};

// Capacity rounding used by every backend: max(c,16).next_power_of_two()
// (rayon_cpu.rs:64).  The reference GPU tier omits the minimum of 16
// (gpu.rs:54); the home mask decides results, so all tiers use one rule.
uint64_t round_capacity(uint64_t requested);

std::unique_ptr<Backend> make_gpu_backend(const fm_builder& b, fm_status* st, std::string* err);
std::unique_ptr<Backend> make_cpu_backend(const fm_builder& b, fm_status* st, std::string* err);

}  // namespace fm

// FlashMap<K,V> (lib.rs:94-96): one backend arm behind a per-handle lock.
struct fm_map {
  std::unique_ptr<fm::Backend> be;
  mutable std::mutex mu;
};
