// fm_capi.cpp -- extern "C" entry points of include/flashmap.h.
//
// fm_map plays the role of FlashMap<K,V> (src/lib.rs:94-103): it owns one
// backend arm and dispatches every method to it.  fm_build mirrors
// FlashMapBuilder::build (lib.rs:414-449): GPU first, CPU on failure with a
// message on stderr.
#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fm_internal.h"

namespace fm {
namespace {
thread_local fm_status t_code = FM_OK;
thread_local std::string t_msg;
thread_local uint64_t t_occ = 0, t_cap = 0;
thread_local double t_lf = 0.0;
}  // namespace

void set_error(fm_status code, const std::string& msg) {
  t_code = code;
  t_msg = msg;
}

// error.rs:17-24; load_factor in percent (gpu.rs:204, rayon_cpu.rs:140).
void set_table_full(uint64_t occupied, uint64_t capacity) {
  t_occ = occupied;
  t_cap = capacity;
  t_lf = capacity ? static_cast<double>(occupied) / static_cast<double>(capacity) * 100.0 : 0.0;
  char buf[160];
  std::snprintf(buf, sizeof buf, "table full: %" PRIu64 " occupied of %" PRIu64 " capacity (load factor %.1f%%)",
                occupied, capacity, t_lf);
  set_error(FM_ERR_TABLE_FULL, buf);
}

// error.rs:32-33
fm_status gpu_required() {
  set_error(FM_ERR_GPU_REQUIRED, "device-resident API requires GPU backend (currently using CPU)");
  return FM_ERR_GPU_REQUIRED;
}

uint64_t round_capacity(uint64_t requested) {
  uint64_t c = requested < 16 ? 16 : requested;
  uint64_t p = 1;
  while (p < c) p <<= 1;
  return p;
}

}  // namespace fm

namespace {

fm_status bad_arg(const char* what) {
  fm::set_error(FM_ERR_INVALID_ARGUMENT, what);
  return FM_ERR_INVALID_ARGUMENT;
}

// error.rs Display texts for the status codes raised at construction.
std::string describe(fm_status st, const std::string& detail) {
  switch (st) {
    case FM_ERR_DEVICE_INIT: return "device initialization failed: " + detail;
    case FM_ERR_GPU_ALLOC: return "GPU memory allocation failed: " + detail;
    case FM_ERR_KERNEL_LAUNCH: return "kernel launch failed: " + detail;
    case FM_ERR_TRANSFER: return "host-device transfer failed: " + detail;
    case FM_ERR_ZERO_CAPACITY: return "capacity must be positive";
    case FM_ERR_NO_BACKEND: return "no backend available";
    default: return detail;
  }
}

}  // namespace

extern "C" {

void fm_builder_init(fm_builder* b, uint64_t capacity, uint32_t key_size, uint32_t value_size) {
  if (!b) return;
  std::memset(b, 0, sizeof *b);
  b->capacity = capacity;
  b->key_size = key_size;
  b->value_size = value_size;
  b->hash_strategy = FM_HASH_IDENTITY;
  b->device_id = 0;
}

fm_status fm_build(const fm_builder* b, fm_map** out) {
  if (!b || !out) return bad_arg("fm_build: null argument");
  *out = nullptr;
  if (b->key_size == 0 || b->value_size == 0) return bad_arg("fm_build: key and value sizes must be positive");
  if (b->hash_strategy > FM_HASH_MURMUR3) return bad_arg("fm_build: unknown hash strategy");
  std::unique_ptr<fm::Backend> be;
  fm_status gst = FM_OK;
  std::string gerr;
  if (!b->force_cpu) {
    be = fm::make_gpu_backend(*b, &gst, &gerr);
    if (!be) {
      const std::string msg = describe(gst, gerr);
      if (b->require_gpu) {
        fm::set_error(gst, msg);
        return gst;
      }
      if (!b->quiet) std::fprintf(stderr, "[flash-map] GPU unavailable (%s), using CPU backend\n", msg.c_str());
    }
  } else if (b->require_gpu) {
    return bad_arg("fm_build: force_cpu and require_gpu are exclusive");
  }
  if (!be) {
    fm_status cst = FM_OK;
    std::string cerr;
    be = fm::make_cpu_backend(*b, &cst, &cerr);
    if (!be) {
      fm::set_error(cst, describe(cst, cerr));
      return cst;
    }
  }
  fm_map* m = new (std::nothrow) fm_map();
  if (!m) {
    fm::set_error(FM_ERR_GPU_ALLOC, "host allocation failed");
    return FM_ERR_GPU_ALLOC;
  }
  m->be = std::move(be);
  *out = m;
  return FM_OK;
}

fm_status fm_with_capacity(uint64_t capacity, uint32_t key_size, uint32_t value_size, fm_map** out) {
  fm_builder b;
  fm_builder_init(&b, capacity, key_size, value_size);
  return fm_build(&b, out);
}

void fm_destroy(fm_map* m) { delete m; }

#define FM_LOCK(m)                                 \
  if (!(m)) return bad_arg("null map handle");     \
  std::lock_guard<std::mutex> fm_guard((m)->mu)

fm_status fm_bulk_get(fm_map* m, const void* keys, uint64_t n, void* out_values, uint8_t* out_found) {
  FM_LOCK(m);
  if (n == 0) return FM_OK;
  if (!keys || !out_values || !out_found) return bad_arg("fm_bulk_get: null buffer");
  return m->be->bulk_get(keys, n, out_values, out_found);
}

fm_status fm_bulk_insert(fm_map* m, const void* keys, const void* values, uint64_t n, uint64_t* out_new) {
  FM_LOCK(m);
  uint64_t dummy = 0;
  if (!out_new) out_new = &dummy;
  *out_new = 0;
  if (n == 0) return FM_OK;
  if (!keys || !values) return bad_arg("fm_bulk_insert: null buffer");
  return m->be->bulk_insert(keys, values, n, out_new);
}

fm_status fm_bulk_remove(fm_map* m, const void* keys, uint64_t n, uint64_t* out_removed) {
  FM_LOCK(m);
  uint64_t dummy = 0;
  if (!out_removed) out_removed = &dummy;
  *out_removed = 0;
  if (n == 0) return FM_OK;
  if (!keys) return bad_arg("fm_bulk_remove: null buffer");
  return m->be->bulk_remove(keys, n, out_removed);
}

uint64_t fm_len(const fm_map* m) {
  if (!m) return 0;
  std::lock_guard<std::mutex> g(m->mu);
  return m->be->len();
}

uint64_t fm_capacity(const fm_map* m) { return m ? m->be->capacity() : 0; }

double fm_load_factor(const fm_map* m) {
  if (!m) return 0.0;
  std::lock_guard<std::mutex> g(m->mu);
  return static_cast<double>(m->be->len()) / static_cast<double>(m->be->capacity());
}

fm_status fm_clear(fm_map* m) {
  FM_LOCK(m);
  return m->be->clear();
}

fm_backend_kind fm_backend(const fm_map* m) { return m ? m->be->kind() : FM_BACKEND_NONE; }
uint32_t fm_key_size(const fm_map* m) { return m ? m->be->key_size() : 0; }
uint32_t fm_value_size(const fm_map* m) { return m ? m->be->value_size() : 0; }

fm_status fm_debug_string(const fm_map* m, char* buf, size_t buf_len) {
  if (!m || !buf) return bad_arg("fm_debug_string: null argument");
  std::lock_guard<std::mutex> g(m->mu);
  const uint64_t len = m->be->len(), cap = m->be->capacity();
  std::snprintf(buf, buf_len, "FlashMap { backend: \"%s\", len: %" PRIu64 ", capacity: %" PRIu64
                ", load_factor: \"%.1f%%\" }",
                m->be->kind() == FM_BACKEND_GPU ? "GPU" : "CPU", len, cap,
                static_cast<double>(len) / static_cast<double>(cap) * 100.0);
  return FM_OK;
}

size_t fm_last_error_message(char* buf, size_t buf_len) {
  const std::string& s = fm::t_msg;
  if (buf && buf_len) {
    const size_t k = std::min(buf_len - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = 0;
  }
  return s.size();
}

fm_status fm_last_table_full(uint64_t* occupied, uint64_t* capacity, double* load_factor_pct) {
  if (fm::t_code != FM_ERR_TABLE_FULL) return bad_arg("no TableFull error recorded on this thread");
  if (occupied) *occupied = fm::t_occ;
  if (capacity) *capacity = fm::t_cap;
  if (load_factor_pct) *load_factor_pct = fm::t_lf;
  return FM_OK;
}

void* fm_device_stream(fm_map* m) { return m ? m->be->device_stream() : nullptr; }
int32_t fm_device_id(const fm_map* m) { return m ? m->be->device_id() : -1; }

fm_status fm_alloc_device(fm_map* m, uint64_t bytes, void** out) {
  FM_LOCK(m);
  if (!out) return bad_arg("fm_alloc_device: null out");
  return m->be->alloc_device(bytes, out);
}

fm_status fm_free_device(fm_map* m, void* p) {
  FM_LOCK(m);
  if (!p) return FM_OK;
  return m->be->free_device(p);
}

fm_status fm_upload(fm_map* m, const void* host, uint64_t bytes, void** out) {
  FM_LOCK(m);
  if (!out || (!host && bytes)) return bad_arg("fm_upload: null argument");
  return m->be->upload(host, bytes, out);
}

fm_status fm_download(fm_map* m, const void* dptr, uint64_t bytes, void* host) {
  FM_LOCK(m);
  if (bytes == 0) return FM_OK;
  if (!dptr || !host) return bad_arg("fm_download: null argument");
  return m->be->download(dptr, bytes, host);
}

fm_status fm_bulk_get_device(fm_map* m, const void* d_keys, uint64_t count, void* d_out_values,
                             uint8_t* d_out_found) {
  FM_LOCK(m);
  if (!d_out_found && m->be->kind() == FM_BACKEND_GPU && count) return bad_arg("fm_bulk_get_device: null found");
  return m->be->bulk_get_device(d_keys, count, d_out_values, d_out_found);
}

fm_status fm_bulk_get_values_device(fm_map* m, const void* d_keys, uint64_t count, void* d_out_values) {
  FM_LOCK(m);
  return m->be->bulk_get_device(d_keys, count, d_out_values, nullptr);
}

fm_status fm_bulk_insert_device(fm_map* m, const void* d_keys, const void* d_values, uint64_t count,
                                uint64_t* out_new) {
  FM_LOCK(m);
  return m->be->bulk_insert_device(d_keys, d_values, count, out_new, true);
}

fm_status fm_bulk_insert_device_noalloc(fm_map* m, const void* d_keys, const void* d_values, uint64_t count,
                                        uint64_t* out_new) {
  FM_LOCK(m);
  return m->be->bulk_insert_device_noalloc(d_keys, d_values, count, out_new);
}

fm_status fm_bulk_insert_device_uncounted(fm_map* m, const void* d_keys, const void* d_values, uint64_t count) {
  FM_LOCK(m);
  return m->be->bulk_insert_device(d_keys, d_values, count, nullptr, false);
}

fm_status fm_bulk_remove_device(fm_map* m, const void* d_keys, uint64_t count, uint64_t* out_removed) {
  FM_LOCK(m);
  return m->be->bulk_remove_device(d_keys, count, out_removed);
}

fm_status fm_recount(fm_map* m, uint64_t* out_len) {
  FM_LOCK(m);
  uint64_t dummy = 0;
  return m->be->recount(out_len ? out_len : &dummy);
}

fm_status fm_synchronize(fm_map* m) {
  FM_LOCK(m);
  return m->be->synchronize();
}

fm_status fm_export_slots(fm_map* m, uint32_t* flags, void* keys, void* values) {
  FM_LOCK(m);
  if (!flags) return bad_arg("fm_export_slots: null flags");
  return m->be->export_slots(flags, keys, values);
}

fm_status fm_import_slots(fm_map* m, const uint32_t* flags, const void* keys, const void* values, uint64_t len) {
  FM_LOCK(m);
  if (!flags || !keys || !values) return bad_arg("fm_import_slots: null argument");
  return m->be->import_slots(flags, keys, values, len);
}

fm_status fm_copy_table(fm_map* dst, fm_map* src, uint32_t parts) {
  if (!dst || !src) return bad_arg("fm_copy_table: null map");
  if (parts > FM_COPY_STRUCTURE) return bad_arg("fm_copy_table: unknown parts");
  if (dst == src) return FM_OK;
  std::scoped_lock guard(dst->mu, src->mu);
  return dst->be->copy_from(*src->be, parts);
}

fm_status fm_profile_enable(fm_map* m, int enable) {
  FM_LOCK(m);
  return m->be->profile_enable(enable != 0);
}

fm_status fm_profile_only(fm_map* m, const char* kernel_name) {
  FM_LOCK(m);
  return m->be->profile_only(kernel_name ? kernel_name : "");
}

fm_status fm_profile_read(fm_map* m, fm_kernel_stat* out, uint32_t max, uint32_t* count) {
  FM_LOCK(m);
  if (!out || !count) return bad_arg("fm_profile_read: null argument");
  return m->be->profile_read(out, max, count);
}

fm_status fm_get_tuning(const fm_map* m, fm_tuning* out) {
  if (!m || !out) return bad_arg("fm_get_tuning: null argument");
  std::lock_guard<std::mutex> g(m->mu);
  return m->be->get_tuning(out);
}

fm_status fm_set_tuning(fm_map* m, const fm_tuning* t) {
  FM_LOCK(m);
  return m->be->set_tuning(t);
}

// Exhaustive over get_group x get_block (35 shapes): gets change nothing,
// so every shape runs on the live table.  Time per shape: median wall time
// of `reps` bulk_get_device calls, each synchronised, after one untimed call.
fm_status fm_autotune_get(fm_map* m, const void* d_keys, uint64_t count, uint32_t reps, fm_tuning* out_best,
                          double* out_get_us) {
  FM_LOCK(m);
  if (m->be->kind() != FM_BACKEND_GPU) return bad_arg("fm_autotune_get: the map is not on a GPU");
  if (!d_keys || count == 0 || reps == 0 || reps > 64) return bad_arg("fm_autotune_get: need keys, count > 0, reps 1..64");
  fm_tuning cur{};
  fm_status st = m->be->get_tuning(&cur);
  if (st != FM_OK) return st;
  const uint64_t V = m->be->value_size();
  void* d_out = nullptr;
  if ((st = m->be->alloc_device(count * V + count, &d_out)) != FM_OK) return st;
  uint8_t* d_found = static_cast<uint8_t*>(d_out) + count * V;
  static const uint32_t kGroups[] = {1, 2, 4, 8, 16};
  static const uint32_t kBlocks[] = {64, 128, 256, 512, 1024};
  fm_tuning top = cur;
  double top_us = -1;
  for (uint32_t g : kGroups) {
    for (uint32_t b : kBlocks) {
      fm_tuning t = cur;
      t.get_group = g;
      t.get_block = b;
      if ((st = m->be->set_tuning(&t)) != FM_OK) break;
      std::vector<double> us;
      for (uint32_t r = 0; r <= reps && st == FM_OK; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        st = m->be->bulk_get_device(d_keys, count, d_out, d_found);
        if (st == FM_OK) st = m->be->synchronize();
        const auto t1 = std::chrono::steady_clock::now();
        if (r) us.push_back(std::chrono::duration<double, std::micro>(t1 - t0).count());
      }
      if (st != FM_OK) break;
      std::sort(us.begin(), us.end());
      if (top_us < 0 || us[us.size() / 2] < top_us) top = t, top_us = us[us.size() / 2];
    }
    if (st != FM_OK) break;
  }
  const fm_status fst = m->be->free_device(d_out);
  if (st != FM_OK) {
    (void)m->be->set_tuning(&cur);
    return st;
  }
  if (fst != FM_OK) return fst;
  if ((st = m->be->set_tuning(&top)) != FM_OK) return st;
  if (out_best) *out_best = top;
  if (out_get_us) *out_get_us = top_us;
  return FM_OK;
}

}  // extern "C"
