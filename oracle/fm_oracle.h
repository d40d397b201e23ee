/*
 * fm_oracle.h -- CPU oracle for the FlashMap bulk-op path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a sequential restatement of the
 * reference's CPU backend (src/rayon_cpu.rs) and hash functions
 * (src/hash.rs), applied one key at a time in batch index order.  Only the
 * test-suite, the smoke check and the benchmark's CPU-baseline leg may load
 * it, and only as the checker.  The shipped library never links it.
 *
 * Semantics pinned by: the reference's own known-answer tests
 * (src/hash.rs:92-104), the 17 closed-form tests of tests/correctness.rs,
 * published MurmurHash3 fmix64 (via the SMHasher verification value of
 * MurmurHash3_x64_128), an independent pure-Python restatement
 * (oracle/pyoracle.py), structural invariants and brute force on 16-slot
 * tables (tests/test_oracle.py).
 */
#ifndef FM_ORACLE_H
#define FM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Slot flag encoding, src/rayon_cpu.rs:22-41 (kernels.cu:23-30). */
#define OR_FLAG_EMPTY 0u
#define OR_FLAG_OCCUPIED 1u
#define OR_FLAG_TOMBSTONE 2u
#define OR_FLAG_INSERTING 3u

/* Status codes mirror FlashMapError (src/error.rs:4-39). */
#define OR_OK 0
#define OR_ERR_TABLE_FULL 5

uint64_t or_identity_hash(const uint8_t* key, uint32_t key_size);
uint64_t or_murmur3_hash(const uint8_t* key, uint32_t key_size);
uint64_t or_hash_key(const uint8_t* key, uint32_t key_size, uint32_t mode);
/* Batch helper: hashes n keys of key_size bytes. */
void or_hash_many(const uint8_t* keys, uint64_t n, uint32_t key_size,
                  uint32_t mode, uint64_t* out);

typedef struct or_map {
    uint8_t* keys;    /* cap * key_size */
    uint8_t* values;  /* cap * value_size */
    uint32_t* flags;  /* cap */
    uint64_t capacity;
    uint64_t mask;
    uint64_t len;
    uint32_t key_size;
    uint32_t value_size;
    uint32_t hash_mode;
} or_map;

/* capacity rounding: max(c,16).next_power_of_two(), rayon_cpu.rs:63-77 */
uint64_t or_round_capacity(uint64_t requested);
or_map* or_create(uint64_t capacity, uint32_t key_size, uint32_t value_size,
                  uint32_t hash_mode);
void or_destroy(or_map* m);

/* rayon_cpu.rs:82-126 */
void or_bulk_get(const or_map* m, const uint8_t* query, uint64_t n,
                 uint8_t* out_values, uint8_t* out_found);
/* rayon_cpu.rs:133-270.  Returns OR_OK or OR_ERR_TABLE_FULL (no mutation). */
int or_bulk_insert(or_map* m, const uint8_t* keys, const uint8_t* values,
                   uint64_t n, uint64_t* out_new);
/* rayon_cpu.rs:273-328 */
uint64_t or_bulk_remove(or_map* m, const uint8_t* keys, uint64_t n);
/* rayon_cpu.rs:347-353 */
void or_clear(or_map* m);
uint64_t or_len(const or_map* m);
uint64_t or_capacity(const or_map* m);

/* Logical slot array: raw flags, and keys/values (bytes of non-OCCUPIED
 * slots are zeroed in the export so exports compare bytewise). */
void or_export(const or_map* m, uint32_t* flags, uint8_t* keys,
               uint8_t* values);
/* Loads a logical slot array (flags, keys, values) and sets len. */
void or_import(or_map* m, const uint32_t* flags, const uint8_t* keys,
               const uint8_t* values, uint64_t len);

/* ---- workload generators (SplitMix64, Steele/Lea/Flood 2014) ---- */
uint64_t or_splitmix64_next(uint64_t* state);

/* ---- multi-threaded restatement of rayon_cpu.rs (CPU baseline) ---- */
typedef struct or_par_map or_par_map;
or_par_map* or_par_create(uint64_t capacity, uint32_t key_size,
                          uint32_t value_size, uint32_t hash_mode,
                          int threads);
void or_par_destroy(or_par_map* m);
void or_par_bulk_get(or_par_map* m, const uint8_t* query, uint64_t n,
                     uint8_t* out_values, uint8_t* out_found);
int or_par_bulk_insert(or_par_map* m, const uint8_t* keys,
                       const uint8_t* values, uint64_t n, uint64_t* out_new);
uint64_t or_par_bulk_remove(or_par_map* m, const uint8_t* keys, uint64_t n);
uint64_t or_par_len(const or_par_map* m);
int or_par_threads(const or_par_map* m);
int or_par_copy_from(or_par_map* dst, const or_par_map* src, int structure_only);

#ifdef __cplusplus
}
#endif
#endif
