// fm_hash.h -- key hashing shared by the GPU kernels and the host code.
//
// Bit-exact with the reference's src/hash.rs:30-86 (device twins in
// src/kernels.cu:42-96).  Wrapping 64-bit integer arithmetic only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fm {

// src/hash.rs:31-39: first min(8, key_size) bytes as a little-endian u64.
__host__ __device__ inline uint64_t identity_hash(const uint8_t* key, uint32_t key_size) {
  uint64_t h = 0;
  const uint32_t n = key_size < 8 ? key_size : 8;
  for (uint32_t i = 0; i < n; ++i) h |= static_cast<uint64_t>(key[i]) << (8 * i);
  return h;
}

// One 8-byte chunk of the murmur-style body: src/hash.rs:52-57.
__host__ __device__ inline uint64_t murmur_chunk(uint64_t h, uint64_t k) {
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  h ^= k;
  return h * 5u + 0x52dce729u;
}

// Tail and length avalanche: src/hash.rs:60-76.
__host__ __device__ inline uint64_t murmur_finish(uint64_t h, uint64_t rem, uint32_t key_size) {
  if (rem != 0 || key_size % 8 != 0) {
    rem *= 0xff51afd7ed558ccdull;
    rem ^= rem >> 33;
    h ^= rem;
  }
  h ^= static_cast<uint64_t>(key_size);
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}

// src/hash.rs:42-77 over a byte pointer (any alignment).
__host__ __device__ inline uint64_t murmur3_hash(const uint8_t* key, uint32_t key_size) {
  uint64_t h = 0x9e3779b97f4a7c15ull;
  const uint32_t chunks = key_size / 8;
  for (uint32_t c = 0; c < chunks; ++c) {
    uint64_t k = 0;
    for (uint32_t i = 0; i < 8; ++i) k |= static_cast<uint64_t>(key[c * 8 + i]) << (8 * i);
    h = murmur_chunk(h, k);
  }
  uint64_t rem = 0;
  for (uint32_t i = chunks * 8; i < key_size; ++i)
    rem |= static_cast<uint64_t>(key[i]) << (8 * (i - chunks * 8));
  return murmur_finish(h, rem, key_size);
}

// Murmur over a key already held as little-endian 32-bit words (key_size a
// multiple of 8 is the common case; any multiple of 4 is handled).
__host__ __device__ inline uint64_t murmur3_words(const uint32_t* w, uint32_t key_size) {
  uint64_t h = 0x9e3779b97f4a7c15ull;
  const uint32_t chunks = key_size / 8;
  for (uint32_t c = 0; c < chunks; ++c)
    h = murmur_chunk(h, static_cast<uint64_t>(w[2 * c]) | (static_cast<uint64_t>(w[2 * c + 1]) << 32));
  const uint64_t rem = (key_size % 8) ? static_cast<uint64_t>(w[2 * chunks]) : 0;
  return murmur_finish(h, rem, key_size);
}

// src/hash.rs:80-86; mode 0 = Identity, 1 = Murmur3 (hash.rs:22-27).
__host__ __device__ inline uint64_t hash_key(const uint8_t* key, uint32_t key_size, uint32_t mode) {
  return mode == 0 ? identity_hash(key, key_size) : murmur3_hash(key, key_size);
}

}  // namespace fm
