/*
 * fm_oracle.c -- sequential CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the reference's CPU path: src/hash.rs (hashing) and
 * src/rayon_cpu.rs (get / insert / remove / clear) with every batch applied
 * one key at a time in batch index order.  "The CPU path" is defined as that
 * in-order application; with one worker thread rayon's par_iter visits items
 * in that order.  Loaded only by tests, the smoke check and the benchmark's
 * CPU-baseline leg, always as the checker.
 */
#include "fm_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Hashing: src/hash.rs:30-86                                          */
/* ------------------------------------------------------------------ */

/* src/hash.rs:31-39 -- first min(8, len) bytes, little-endian. */
uint64_t or_identity_hash(const uint8_t* key, uint32_t key_size) {
    uint64_t h = 0;
    uint32_t n = key_size < 8 ? key_size : 8;
    for (uint32_t i = 0; i < n; i++) h |= ((uint64_t)key[i]) << (i * 8);
    return h;
}

/* src/hash.rs:42-77 -- wrapping u64 arithmetic. */
uint64_t or_murmur3_hash(const uint8_t* key, uint32_t key_size) {
    uint64_t h = 0x9e3779b97f4a7c15ull;
    uint32_t chunks = key_size / 8;
    for (uint32_t c = 0; c < chunks; c++) {
        uint64_t k = 0;
        for (uint32_t i = 0; i < 8; i++)
            k |= ((uint64_t)key[c * 8 + i]) << (i * 8);
        k *= 0xff51afd7ed558ccdull;
        k ^= k >> 33;
        k *= 0xc4ceb9fe1a85ec53ull;
        k ^= k >> 33;
        h ^= k;
        h = h * 5u + 0x52dce729u;
    }
    uint64_t rem = 0;
    for (uint32_t i = chunks * 8; i < key_size; i++)
        rem |= ((uint64_t)key[i]) << ((i - chunks * 8) * 8);
    if (rem != 0 || key_size % 8 != 0) { /* hash.rs:64 */
        rem *= 0xff51afd7ed558ccdull;
        rem ^= rem >> 33;
        h ^= rem;
    }
    h ^= (uint64_t)key_size;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

/* src/hash.rs:80-86; mode numbering from hash.rs:22-27. */
uint64_t or_hash_key(const uint8_t* key, uint32_t key_size, uint32_t mode) {
    return mode == 0 ? or_identity_hash(key, key_size)
                     : or_murmur3_hash(key, key_size);
}

void or_hash_many(const uint8_t* keys, uint64_t n, uint32_t key_size,
                  uint32_t mode, uint64_t* out) {
    for (uint64_t i = 0; i < n; i++)
        out[i] = or_hash_key(keys + i * key_size, key_size, mode);
}

/* ------------------------------------------------------------------ */
/* Flag helpers: src/rayon_cpu.rs:22-41                                */
/* ------------------------------------------------------------------ */
static inline uint32_t get_flag(uint32_t f) { return f & 0xFu; }
static inline uint64_t get_dist(uint32_t f) { return (uint64_t)(f >> 4); }
/* ((dist as u32) << 4) | (flag & 0xF): the u32 cast truncates. */
static inline uint32_t make_flag(uint32_t flag, uint64_t dist) {
    return (uint32_t)((uint32_t)dist << 4) | (flag & 0xFu);
}

/* ------------------------------------------------------------------ */
/* Construction: src/rayon_cpu.rs:63-77                                */
/* ------------------------------------------------------------------ */
uint64_t or_round_capacity(uint64_t requested) {
    uint64_t c = requested < 16 ? 16 : requested;
    uint64_t p = 1;
    while (p < c) p <<= 1;
    return p;
}

or_map* or_create(uint64_t capacity, uint32_t key_size, uint32_t value_size,
                  uint32_t hash_mode) {
    or_map* m = (or_map*)calloc(1, sizeof(or_map));
    if (!m) return NULL;
    m->capacity = or_round_capacity(capacity);
    m->mask = m->capacity - 1;
    m->key_size = key_size;
    m->value_size = value_size;
    m->hash_mode = hash_mode;
    m->keys = (uint8_t*)calloc(m->capacity, key_size ? key_size : 1);
    m->values = (uint8_t*)calloc(m->capacity, value_size ? value_size : 1);
    m->flags = (uint32_t*)calloc(m->capacity, sizeof(uint32_t));
    if (!m->keys || !m->values || !m->flags) {
        or_destroy(m);
        return NULL;
    }
    return m;
}

void or_destroy(or_map* m) {
    if (!m) return;
    free(m->keys);
    free(m->values);
    free(m->flags);
    free(m);
}

/* ------------------------------------------------------------------ */
/* bulk_get: src/rayon_cpu.rs:82-126 (probe loop :91-120)               */
/* ------------------------------------------------------------------ */
static int64_t probe_find(const or_map* m, const uint8_t* q) {
    const uint32_t K = m->key_size;
    uint64_t slot = or_hash_key(q, K, m->hash_mode) & m->mask;
    uint64_t p = 0;
    while (p <= m->mask) {
        uint64_t idx = (slot + p) & m->mask;
        uint32_t f = m->flags[idx];
        uint32_t flag = get_flag(f);
        uint64_t dist = get_dist(f);
        if (flag == OR_FLAG_EMPTY) return -1;
        /* FLAG_INSERTING never exists between sequential ops (:101-105). */
        if (flag == OR_FLAG_OCCUPIED) {
            if (p > dist) return -1; /* Robin Hood early exit (:108-110) */
            if (memcmp(m->keys + idx * K, q, K) == 0) return (int64_t)idx;
        }
        p += 1; /* tombstone or different key */
    }
    return -1;
}

void or_bulk_get(const or_map* m, const uint8_t* query, uint64_t n,
                 uint8_t* out_values, uint8_t* out_found) {
    const uint32_t K = m->key_size, V = m->value_size;
    for (uint64_t i = 0; i < n; i++) {
        int64_t idx = probe_find(m, query + i * K);
        if (idx >= 0) {
            memcpy(out_values + i * V, m->values + (uint64_t)idx * V, V);
            out_found[i] = 1;
        } else {
            memset(out_values + i * V, 0, V);
            out_found[i] = 0;
        }
    }
}

/* ------------------------------------------------------------------ */
/* bulk_insert: src/rayon_cpu.rs:133-270                               */
/* ------------------------------------------------------------------ */
int or_bulk_insert(or_map* m, const uint8_t* keys, const uint8_t* values,
                   uint64_t n, uint64_t* out_new) {
    const uint32_t K = m->key_size, V = m->value_size;
    /* TableFull pre-check, MAX_LOAD_FACTOR = 1.0 (:134-142): no mutation. */
    if (m->len + n > m->capacity) {
        if (out_new) *out_new = 0;
        return OR_ERR_TABLE_FULL;
    }
    uint8_t* cur_key = (uint8_t*)malloc(K ? K : 1);
    uint8_t* cur_val = (uint8_t*)malloc(V ? V : 1);
    uint8_t* tmp_key = (uint8_t*)malloc(K ? K : 1);
    uint8_t* tmp_val = (uint8_t*)malloc(V ? V : 1);
    uint64_t num_new = 0;
    for (uint64_t i = 0; i < n; i++) {
        memcpy(cur_key, keys + i * K, K);
        memcpy(cur_val, values + i * V, V);
        uint64_t home = or_hash_key(cur_key, K, m->hash_mode) & m->mask;
        uint64_t p = 0;
        for (;;) {
            if (p > m->mask) break; /* carried pair dropped (:167-169) */
            uint64_t idx = (home + p) & m->mask;
            uint32_t f = m->flags[idx];
            uint32_t flag = get_flag(f);
            uint64_t dist = get_dist(f);
            if (flag == OR_FLAG_OCCUPIED) {
                if (memcmp(m->keys + idx * K, cur_key, K) == 0) {
                    /* update in place, not counted (:176-182) */
                    memcpy(m->values + idx * V, cur_val, V);
                    break;
                }
                if (p > dist) {
                    /* Robin Hood: evict resident, carry it on (:184-218) */
                    memcpy(tmp_key, m->keys + idx * K, K);
                    memcpy(tmp_val, m->values + idx * V, V);
                    memcpy(m->keys + idx * K, cur_key, K);
                    memcpy(m->values + idx * V, cur_val, V);
                    m->flags[idx] = make_flag(OR_FLAG_OCCUPIED, p);
                    memcpy(cur_key, tmp_key, K);
                    memcpy(cur_val, tmp_val, V);
                    home = or_hash_key(cur_key, K, m->hash_mode) & m->mask;
                    p = dist + 1;
                    continue;
                }
                p += 1;
                continue;
            }
            if (flag == OR_FLAG_EMPTY || flag == OR_FLAG_TOMBSTONE) {
                /* claim (:228-254) */
                memcpy(m->keys + idx * K, cur_key, K);
                memcpy(m->values + idx * V, cur_val, V);
                m->flags[idx] = make_flag(OR_FLAG_OCCUPIED, p);
                num_new += 1;
                break;
            }
            p += 1; /* unreachable between sequential ops */
        }
    }
    free(cur_key);
    free(cur_val);
    free(tmp_key);
    free(tmp_val);
    m->len += num_new; /* :267-269 */
    if (out_new) *out_new = num_new;
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* bulk_remove: src/rayon_cpu.rs:273-328                               */
/* ------------------------------------------------------------------ */
uint64_t or_bulk_remove(or_map* m, const uint8_t* keys, uint64_t n) {
    const uint32_t K = m->key_size;
    uint64_t removed = 0;
    for (uint64_t i = 0; i < n; i++) {
        int64_t idx = probe_find(m, keys + i * K);
        if (idx >= 0) {
            m->flags[idx] = OR_FLAG_TOMBSTONE; /* raw 2: dist bits cleared */
            removed += 1;
        }
    }
    m->len -= removed; /* :325-327 */
    return removed;
}

/* src/rayon_cpu.rs:347-353 -- flags only; stale bytes stay. */
void or_clear(or_map* m) {
    memset(m->flags, 0, m->capacity * sizeof(uint32_t));
    m->len = 0;
}

uint64_t or_len(const or_map* m) { return m->len; }
uint64_t or_capacity(const or_map* m) { return m->capacity; }

void or_export(const or_map* m, uint32_t* flags, uint8_t* keys,
               uint8_t* values) {
    const uint32_t K = m->key_size, V = m->value_size;
    memcpy(flags, m->flags, m->capacity * sizeof(uint32_t));
    for (uint64_t i = 0; i < m->capacity; i++) {
        int occ = get_flag(m->flags[i]) == OR_FLAG_OCCUPIED;
        if (keys) {
            if (occ) memcpy(keys + i * K, m->keys + i * K, K);
            else memset(keys + i * K, 0, K);
        }
        if (values) {
            if (occ) memcpy(values + i * V, m->values + i * V, V);
            else memset(values + i * V, 0, V);
        }
    }
}

void or_import(or_map* m, const uint32_t* flags, const uint8_t* keys,
               const uint8_t* values, uint64_t len) {
    memcpy(m->flags, flags, m->capacity * sizeof(uint32_t));
    memcpy(m->keys, keys, m->capacity * m->key_size);
    memcpy(m->values, values, m->capacity * m->value_size);
    m->len = len;
}

/* SplitMix64 (Steele, Lea, Flood 2014; Vigna's public-domain version). */
uint64_t or_splitmix64_next(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
