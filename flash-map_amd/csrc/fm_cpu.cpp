// fm_cpu.cpp -- CPU backend of the FlashMap C ABI (the builder's fallback arm;
// replaces RayonFlashMap, src/rayon_cpu.rs:45-353).
//
// Gets run on all host threads; inserts and removes apply the batch in
// index order, which is the deterministic refinement of the reference's
// parallel CAS loop (its outcome for duplicate keys is otherwise
// schedule-dependent, lib.rs:138-142).
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "fm_hash.h"
#include "fm_internal.h"

namespace fm {
namespace {

constexpr uint32_t kEmpty = 0, kOcc = 1, kTomb = 2;

class CpuBackend final : public Backend {
 public:
  fm_status init(const fm_builder& b, std::string* err) {
    cap_ = round_capacity(b.capacity == 0 ? 16 : b.capacity);  // rayon_cpu.rs:64 maps 0 to 16
    mask_ = cap_ - 1;
    K_ = b.key_size;
    V_ = b.value_size;
    mode_ = b.hash_strategy == FM_HASH_MURMUR3 ? 1 : 0;
    if (K_ == 0 || V_ == 0) {
      *err = "key and value sizes must be positive";
      return FM_ERR_INVALID_ARGUMENT;
    }
    try {
      keys_.assign(cap_ * K_, 0);
      vals_.assign(cap_ * V_, 0);
      flags_.assign(cap_, kEmpty);
    } catch (...) {
      *err = "host allocation failed";
      return FM_ERR_GPU_ALLOC;
    }
    threads_ = std::max(1u, std::thread::hardware_concurrency());
    return FM_OK;
  }

  fm_backend_kind kind() const override { return FM_BACKEND_CPU; }
  uint64_t capacity() const override { return cap_; }
  uint64_t len() const override { return len_; }
  uint32_t key_size() const override { return K_; }
  uint32_t value_size() const override { return V_; }

  fm_status bulk_get(const void* keys, uint64_t n, void* out_values, uint8_t* out_found) override {
    const uint8_t* q = static_cast<const uint8_t*>(keys);
    uint8_t* ov = static_cast<uint8_t*>(out_values);
    auto work = [&](uint64_t b, uint64_t e) {
      for (uint64_t i = b; i < e; ++i) {
        const int64_t s = find(q + i * K_);
        if (s >= 0) {
          std::memcpy(ov + i * V_, &vals_[s * V_], V_);
          out_found[i] = 1;
        } else {
          std::memset(ov + i * V_, 0, V_);
          out_found[i] = 0;
        }
      }
    };
    const uint64_t nt = std::min<uint64_t>(threads_, (n + 4095) / 4096);
    if (nt <= 1) {
      work(0, n);
      return FM_OK;
    }
    std::vector<std::thread> pool;
    const uint64_t chunk = (n + nt - 1) / nt;
    for (uint64_t t = 0; t < nt; ++t) {
      const uint64_t b = t * chunk, e = std::min(n, b + chunk);
      if (b < e) pool.emplace_back(work, b, e);
    }
    for (auto& th : pool) th.join();
    return FM_OK;
  }

  fm_status bulk_insert(const void* keys, const void* values, uint64_t n, uint64_t* out_new) override {
    if (n == 0) {
      *out_new = 0;
      return FM_OK;
    }
    if (len_ + n > cap_) {  // rayon_cpu.rs:134-142
      set_table_full(len_, cap_);
      return FM_ERR_TABLE_FULL;
    }
    const uint8_t* kin = static_cast<const uint8_t*>(keys);
    const uint8_t* vin = static_cast<const uint8_t*>(values);
    std::vector<uint8_t> ck(K_), cv(V_), tk(K_), tv(V_);
    uint64_t added = 0;
    for (uint64_t i = 0; i < n; ++i) {
      std::memcpy(ck.data(), kin + i * K_, K_);
      std::memcpy(cv.data(), vin + i * V_, V_);
      uint64_t home = hash_key(ck.data(), K_, mode_) & mask_;
      uint64_t p = 0;
      for (;;) {
        if (p > mask_) break;  // carried pair dropped (rayon_cpu.rs:167-169)
        const uint64_t s = (home + p) & mask_;
        const uint32_t f = flags_[s];
        const uint32_t st = f & 0xF;
        const uint64_t d = f >> 4;
        if (st == kOcc) {
          if (std::memcmp(&keys_[s * K_], ck.data(), K_) == 0) {
            std::memcpy(&vals_[s * V_], cv.data(), V_);
            break;
          }
          if (p > d) {
            std::memcpy(tk.data(), &keys_[s * K_], K_);
            std::memcpy(tv.data(), &vals_[s * V_], V_);
            std::memcpy(&keys_[s * K_], ck.data(), K_);
            std::memcpy(&vals_[s * V_], cv.data(), V_);
            flags_[s] = (static_cast<uint32_t>(p) << 4) | kOcc;
            std::swap(ck, tk);
            std::swap(cv, tv);
            home = hash_key(ck.data(), K_, mode_) & mask_;
            p = d + 1;
            continue;
          }
          ++p;
          continue;
        }
        std::memcpy(&keys_[s * K_], ck.data(), K_);
        std::memcpy(&vals_[s * V_], cv.data(), V_);
        flags_[s] = (static_cast<uint32_t>(p) << 4) | kOcc;
        ++added;
        break;
      }
    }
    len_ += added;
    *out_new = added;
    return FM_OK;
  }

  fm_status bulk_remove(const void* keys, uint64_t n, uint64_t* out_removed) override {
    const uint8_t* q = static_cast<const uint8_t*>(keys);
    uint64_t removed = 0;
    for (uint64_t i = 0; i < n; ++i) {
      const int64_t s = find(q + i * K_);
      if (s >= 0) {
        flags_[s] = kTomb;
        ++removed;
      }
    }
    len_ -= removed;
    *out_removed = removed;
    return FM_OK;
  }

  fm_status clear() override {
    std::fill(flags_.begin(), flags_.end(), kEmpty);
    len_ = 0;
    return FM_OK;
  }

  fm_status export_slots(uint32_t* flags, void* keys, void* values) override {
    std::memcpy(flags, flags_.data(), cap_ * 4);
    for (uint64_t i = 0; i < cap_; ++i) {
      const bool occ = (flags_[i] & 0xF) == kOcc;
      if (keys) {
        uint8_t* k = static_cast<uint8_t*>(keys) + i * K_;
        if (occ) std::memcpy(k, &keys_[i * K_], K_);
        else std::memset(k, 0, K_);
      }
      if (values) {
        uint8_t* v = static_cast<uint8_t*>(values) + i * V_;
        if (occ) std::memcpy(v, &vals_[i * V_], V_);
        else std::memset(v, 0, V_);
      }
    }
    return FM_OK;
  }

  fm_status import_slots(const uint32_t* flags, const void* keys, const void* values, uint64_t len) override {
    std::memcpy(flags_.data(), flags, cap_ * 4);
    std::memcpy(keys_.data(), keys, cap_ * K_);
    std::memcpy(vals_.data(), values, cap_ * V_);
    len_ = len;
    return FM_OK;
  }

 private:
  // rayon_cpu.rs:91-120 (between ops no slot is INSERTING).
  int64_t find(const uint8_t* q) const {
    const uint64_t home = hash_key(q, K_, mode_) & mask_;
    for (uint64_t p = 0; p <= mask_; ++p) {
      const uint64_t s = (home + p) & mask_;
      const uint32_t f = flags_[s];
      const uint32_t st = f & 0xF;
      if (st == kEmpty) return -1;
      if (st == kOcc) {
        if (p > (f >> 4)) return -1;
        if (std::memcmp(&keys_[s * K_], q, K_) == 0) return static_cast<int64_t>(s);
      }
    }
    return -1;
  }

  uint64_t cap_ = 0, mask_ = 0, len_ = 0;
  uint32_t K_ = 0, V_ = 0, mode_ = 0;
  unsigned threads_ = 1;
  std::vector<uint8_t> keys_, vals_;
  std::vector<uint32_t> flags_;
};

}  // namespace

std::unique_ptr<Backend> make_cpu_backend(const fm_builder& b, fm_status* st, std::string* err) {
  std::unique_ptr<CpuBackend> c(new CpuBackend());
  *st = c->init(b, err);
  if (*st != FM_OK) return nullptr;
  return c;
}

}  // namespace fm
