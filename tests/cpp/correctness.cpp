// correctness.cpp -- the reference's tests/correctness.rs restated in C++
// against the flash_map:: facade.  Usage: correctness cpu|gpu
// ("gpu" builds every map with require_gpu(), so no test can pass on a
// fallback).  Prints one line per test and exits non-zero on failure.
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../flash-map_amd/host/flashmap.hpp"

using namespace flash_map;
using B8 = std::array<uint8_t, 8>;

static bool g_gpu = false;
static int g_failed = 0;

#define CHECK(cond)                                                                 \
  do {                                                                              \
    if (!(cond)) throw std::runtime_error(std::string("check failed: ") + #cond);   \
  } while (0)

template <class K, class V>
FlashMap<K, V> make(uint64_t cap, HashStrategy h = HashStrategy::Identity) {
  FlashMapBuilder b(cap);
  b.hash_strategy(h).quiet();
  if (g_gpu) b.require_gpu();
  else b.force_cpu();
  FlashMap<K, V> m = b.build<K, V>();
  CHECK(m.backend() == (g_gpu ? Backend::Gpu : Backend::Cpu));
  return m;
}

static B8 le(uint64_t x) {
  B8 b;
  std::memcpy(b.data(), &x, 8);
  return b;
}

template <class T>
static std::vector<T> range_map(uint64_t lo, uint64_t hi, std::function<T(uint64_t)> f) {
  std::vector<T> v;
  for (uint64_t i = lo; i < hi; ++i) v.push_back(f(i));
  return v;
}

static std::vector<std::pair<B8, B8>> pairs(uint64_t lo, uint64_t hi, std::function<uint64_t(uint64_t)> val) {
  return range_map<std::pair<B8, B8>>(lo, hi, [&](uint64_t i) { return std::make_pair(le(i), le(val(i))); });
}
static std::vector<B8> keys(uint64_t lo, uint64_t hi) { return range_map<B8>(lo, hi, le); }

// correctness.rs:4-22
static void insert_and_get() {
  auto m = make<B8, B8>(1024);
  CHECK(m.bulk_insert(pairs(0, 100, [](uint64_t i) { return i * 100; })) == 100);
  CHECK(m.len() == 100);
  auto r = m.bulk_get(keys(0, 100));
  for (uint64_t i = 0; i < 100; ++i) CHECK(r[i] && *r[i] == le(i * 100));
}

// :25-39
static void get_missing_keys() {
  auto m = make<B8, B8>(1024);
  m.bulk_insert(pairs(0, 10, [](uint64_t i) { return i; }));
  for (auto& r : m.bulk_get(keys(100, 110))) CHECK(!r);
}

// :42-65
static void update_existing_keys() {
  auto m = make<B8, B8>(1024);
  CHECK(m.bulk_insert(pairs(0, 10, [](uint64_t i) { return i; })) == 10);
  CHECK(m.bulk_insert(pairs(0, 10, [](uint64_t i) { return i + 1000; })) == 0);
  CHECK(m.len() == 10);
  auto r = m.bulk_get(keys(0, 10));
  for (uint64_t i = 0; i < 10; ++i) CHECK(r[i] && *r[i] == le(i + 1000));
}

// :68-95
static void remove_keys() {
  auto m = make<B8, B8>(1024);
  m.bulk_insert(pairs(0, 100, [](uint64_t i) { return i; }));
  CHECK(m.len() == 100);
  CHECK(m.bulk_remove(keys(0, 50)) == 50);
  CHECK(m.len() == 50);
  for (auto& r : m.bulk_get(keys(0, 50))) CHECK(!r);
  for (auto& r : m.bulk_get(keys(50, 100))) CHECK(r.has_value());
}

// :98-124
static void insert_after_remove_reuses_tombstones() {
  auto m = make<B8, B8>(64);
  m.bulk_insert(pairs(0, 10, [](uint64_t i) { return i; }));
  m.bulk_remove(keys(0, 10));
  CHECK(m.len() == 0);
  CHECK(m.bulk_insert(pairs(100, 110, [](uint64_t i) { return i * 2; })) == 10);
  auto r = m.bulk_get(keys(100, 110));
  for (uint64_t i = 0; i < 10; ++i) CHECK(r[i] && *r[i] == le((100 + i) * 2));
}

// :127-145
static void clear_resets_map() {
  auto m = make<B8, B8>(1024);
  m.bulk_insert(pairs(0, 100, [](uint64_t i) { return i; }));
  CHECK(m.len() == 100);
  m.clear();
  CHECK(m.len() == 0 && m.is_empty());
  for (auto& r : m.bulk_get(keys(0, 100))) CHECK(!r);
}

// :148-164
static void murmur3_hash_strategy() {
  auto m = make<B8, B8>(1024, HashStrategy::Murmur3);
  m.bulk_insert(pairs(0, 100, [](uint64_t i) { return i * 7; }));
  auto r = m.bulk_get(keys(0, 100));
  for (uint64_t i = 0; i < 100; ++i) CHECK(r[i] && *r[i] == le(i * 7));
}

// :167-191
static void large_keys_and_values() {
  using K = std::array<uint8_t, 32>;
  using V = std::array<uint8_t, 64>;
  auto m = make<K, V>(4096);
  std::vector<std::pair<K, V>> ps;
  for (uint64_t i = 0; i < 500; ++i) {
    K k{};
    V v{};
    std::memcpy(k.data(), &i, 8);
    const uint64_t x = i * 42;
    std::memcpy(v.data(), &x, 8);
    ps.emplace_back(k, v);
  }
  m.bulk_insert(ps);
  CHECK(m.len() == 500);
  std::vector<K> ks;
  for (auto& p : ps) ks.push_back(p.first);
  auto r = m.bulk_get(ks);
  for (size_t i = 0; i < 500; ++i) CHECK(r[i] && *r[i] == ps[i].second);
}

// :194-199
static void load_factor_and_capacity() {
  auto m = make<B8, B8>(1024);
  CHECK(m.load_factor() == 0.0);
  CHECK(m.capacity() >= 1024 && (m.capacity() & (m.capacity() - 1)) == 0);
}

// :202-211
static void empty_operations_are_no_ops() {
  auto m = make<B8, B8>(64);
  CHECK(m.bulk_insert({}) == 0);
  CHECK(m.bulk_get({}).empty());
  CHECK(m.bulk_remove({}) == 0);
}

// :214-227
static void table_full_error() {
  auto m = make<B8, B8>(16);
  m.bulk_insert(pairs(0, 16, [](uint64_t i) { return i; }));
  bool threw = false;
  try {
    m.bulk_insert({{le(99), le(99)}});
  } catch (const TableFull& e) {
    threw = e.occupied == 16 && e.capacity == 16 && e.load_factor == 100.0;
  }
  CHECK(threw);
}

// :230-236
static void debug_formatting() {
  auto m = make<B8, B8>(256);
  const std::string s = m.debug_string();
  CHECK(s.find("FlashMap") != std::string::npos && s.find("len") != std::string::npos &&
        s.find("capacity") != std::string::npos);
}

// :239-250
static void u64_key_value_types() {
  auto m = make<uint64_t, uint64_t>(1024);
  std::vector<std::pair<uint64_t, uint64_t>> ps;
  for (uint64_t i = 0; i < 50; ++i) ps.emplace_back(i, i * i);
  m.bulk_insert(ps);
  std::vector<uint64_t> ks;
  for (uint64_t i = 0; i < 50; ++i) ks.push_back(i);
  auto r = m.bulk_get(ks);
  for (uint64_t i = 0; i < 50; ++i) CHECK(r[i] && *r[i] == i * i);
}

// :257-285
static void robin_hood_high_load_correctness() {
  for (uint64_t pct : {80, 90, 95}) {
    const uint64_t n = 1024 * pct / 100;
    auto m = make<B8, B8>(1024);
    CHECK(m.bulk_insert(pairs(0, n, [](uint64_t i) { return i * 7; })) == n);
    CHECK(m.len() == n);
    auto r = m.bulk_get(keys(0, n));
    for (uint64_t i = 0; i < n; ++i) CHECK(r[i] && *r[i] == le(i * 7));
  }
}

// :288-309
static void robin_hood_miss_exits_early() {
  const uint64_t n = 1024 * 90 / 100;
  auto m = make<B8, B8>(1024);
  m.bulk_insert(pairs(0, n, [](uint64_t i) { return i; }));
  for (auto& r : m.bulk_get(keys(10000, 10500))) CHECK(!r);
}

// :312-362
static void robin_hood_insert_remove_reinsert() {
  const uint64_t n = 512 * 80 / 100;
  auto m = make<B8, B8>(512);
  m.bulk_insert(pairs(0, n, [](uint64_t i) { return i; }));
  CHECK(m.bulk_remove(keys(0, n / 2)) == n / 2);
  CHECK(m.bulk_insert(pairs(10000, 10000 + n / 2, [](uint64_t i) { return i * 3; })) == n / 2);
  auto r = m.bulk_get(keys(n / 2, n));
  for (uint64_t i = 0; i < n - n / 2; ++i) CHECK(r[i] && *r[i] == le(n / 2 + i));
  r = m.bulk_get(keys(10000, 10000 + n / 2));
  for (uint64_t i = 0; i < n / 2; ++i) CHECK(r[i] && *r[i] == le((10000 + i) * 3));
}

// :365-392
static void robin_hood_update_at_high_load() {
  const uint64_t n = 1024 * 45 / 100;
  auto m = make<B8, B8>(1024);
  m.bulk_insert(pairs(0, n, [](uint64_t i) { return i; }));
  CHECK(m.bulk_insert(pairs(0, n, [](uint64_t i) { return i + 999; })) == 0);
  CHECK(m.len() == n);
  auto r = m.bulk_get(keys(0, n));
  for (uint64_t i = 0; i < n; ++i) CHECK(r[i] && *r[i] == le(i + 999));
}

// AsyncFlashMap (async_map.rs): concurrent gets, exclusive mutations.
static void async_wrapper() {
  AsyncFlashMap<uint64_t, uint64_t> am(make<uint64_t, uint64_t>(1024));
  std::vector<std::pair<uint64_t, uint64_t>> ps;
  std::vector<uint64_t> ks;
  for (uint64_t i = 0; i < 100; ++i) {
    ps.emplace_back(i, i * 3);
    ks.push_back(i);
  }
  CHECK(am.bulk_insert(ps).get() == 100);
  std::vector<std::future<std::vector<std::optional<uint64_t>>>> gets;
  for (int t = 0; t < 8; ++t) gets.push_back(am.bulk_get(ks));
  for (auto& f : gets) {
    auto r = f.get();
    for (uint64_t i = 0; i < 100; ++i) CHECK(r[i] && *r[i] == i * 3);
  }
  CHECK(am.bulk_remove({1, 2, 3}).get() == 3);
  CHECK(am.len() == 97);
  am.clear().get();
  CHECK(am.is_empty());
}

int main(int argc, char** argv) {
  g_gpu = argc > 1 && std::string(argv[1]) == "gpu";
  const std::vector<std::pair<const char*, void (*)()>> tests = {
      {"insert_and_get", insert_and_get},
      {"get_missing_keys", get_missing_keys},
      {"update_existing_keys", update_existing_keys},
      {"remove_keys", remove_keys},
      {"insert_after_remove_reuses_tombstones", insert_after_remove_reuses_tombstones},
      {"clear_resets_map", clear_resets_map},
      {"murmur3_hash_strategy", murmur3_hash_strategy},
      {"large_keys_and_values", large_keys_and_values},
      {"load_factor_and_capacity", load_factor_and_capacity},
      {"empty_operations_are_no_ops", empty_operations_are_no_ops},
      {"table_full_error", table_full_error},
      {"debug_formatting", debug_formatting},
      {"u64_key_value_types", u64_key_value_types},
      {"robin_hood_high_load_correctness", robin_hood_high_load_correctness},
      {"robin_hood_miss_exits_early", robin_hood_miss_exits_early},
      {"robin_hood_insert_remove_reinsert", robin_hood_insert_remove_reinsert},
      {"robin_hood_update_at_high_load", robin_hood_update_at_high_load},
      {"async_wrapper", async_wrapper},
  };
  for (auto& [name, fn] : tests) {
    std::printf("test %s ... ", name);
    std::fflush(stdout);
    try {
      fn();
      std::printf("ok\n");
    } catch (const std::exception& e) {
      std::printf("FAILED: %s\n", e.what());
      ++g_failed;
    }
    std::fflush(stdout);
  }
  std::printf("%zu tests, %d failed (%s backend)\n", tests.size(), g_failed, g_gpu ? "gpu" : "cpu");
  return g_failed ? 1 : 0;
}
