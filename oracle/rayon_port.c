/*
 * rayon_port.c -- multi-threaded restatement of src/rayon_cpu.rs
 * (TEST / BASELINE INFRASTRUCTURE ONLY).
 *
 * Same algorithm as the reference's Rayon backend: AtomicU32 flags, CAS
 * claims (AcqRel / Relaxed), Release publication, Acquire loads, one unit of
 * work per key, keys handed out to a persistent worker pool in chunks (rayon's
 * work-stealing pool is persistent too).  Used as the CPU baseline timed
 * beside the GPU.  Like the reference, concurrent inserts of duplicate keys or
 * evict-vs-update races are order-dependent, so its results are compared with
 * the sequential oracle only where order cannot matter.
 */
#include "fm_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define PAR_CHUNK 256

struct or_par_map {
    uint8_t* keys;
    uint8_t* values;
    _Atomic uint32_t* flags;
    uint64_t capacity, mask;
    _Atomic uint64_t len;
    uint32_t key_size, value_size, hash_mode;
    /* pool */
    int threads;
    pthread_t* tids;
    pthread_mutex_t mu;
    pthread_cond_t cv_start, cv_done;
    uint64_t generation;
    int pending;
    int shutdown;
    /* current job */
    int job;  /* 0 get, 1 insert, 2 remove */
    const uint8_t* in_keys;
    const uint8_t* in_values;
    uint8_t* out_values;
    uint8_t* out_found;
    uint64_t n;
    _Atomic uint64_t next;
    _Atomic uint64_t counter;
};

static inline uint32_t get_flag(uint32_t f) { return f & 0xFu; }
static inline uint64_t get_dist(uint32_t f) { return (uint64_t)(f >> 4); }
static inline uint32_t make_flag(uint32_t flag, uint64_t dist) {
    return (uint32_t)((uint32_t)dist << 4) | (flag & 0xFu);
}
static inline void spin_hint(void) {
#if defined(__x86_64__)
    __builtin_ia32_pause();
#endif
}

/* rayon_cpu.rs:86-121 */
static void par_get_one(or_par_map* m, uint64_t i) {
    const uint32_t K = m->key_size, V = m->value_size;
    const uint8_t* q = m->in_keys + i * K;
    uint64_t slot = or_hash_key(q, K, m->hash_mode) & m->mask;
    uint64_t p = 0;
    while (p <= m->mask) {
        uint64_t idx = (slot + p) & m->mask;
        uint32_t f = atomic_load_explicit(&m->flags[idx], memory_order_acquire);
        uint32_t flag = get_flag(f);
        uint64_t dist = get_dist(f);
        if (flag == OR_FLAG_EMPTY) break;
        if (flag == OR_FLAG_INSERTING) { spin_hint(); continue; }
        if (flag == OR_FLAG_OCCUPIED) {
            if (p > dist) break;
            if (memcmp(m->keys + idx * K, q, K) == 0) {
                memcpy(m->out_values + i * V, m->values + idx * V, V);
                m->out_found[i] = 1;
                return;
            }
        }
        p += 1;
    }
    memset(m->out_values + i * V, 0, V);
    m->out_found[i] = 0;
}

/* rayon_cpu.rs:151-263 (thread-local carry buffers, :155-156) */
static void par_insert_one(or_par_map* m, uint64_t i, uint8_t* cur_key,
                           uint8_t* cur_val, uint8_t* tk, uint8_t* tv) {
    const uint32_t K = m->key_size, V = m->value_size;
    memcpy(cur_key, m->in_keys + i * K, K);
    memcpy(cur_val, m->in_values + i * V, V);
    uint64_t home = or_hash_key(cur_key, K, m->hash_mode) & m->mask;
    uint64_t p = 0;
    for (;;) {
        if (p > m->mask) return;
        uint64_t idx = (home + p) & m->mask;
        uint32_t f = atomic_load_explicit(&m->flags[idx], memory_order_acquire);
        uint32_t flag = get_flag(f);
        uint64_t dist = get_dist(f);
        if (flag == OR_FLAG_OCCUPIED) {
            if (memcmp(m->keys + idx * K, cur_key, K) == 0) {
                memcpy(m->values + idx * V, cur_val, V);
                return;
            }
            if (p > dist) {
                uint32_t expect = f;
                if (atomic_compare_exchange_strong_explicit(
                        &m->flags[idx], &expect, make_flag(OR_FLAG_INSERTING, p),
                        memory_order_acq_rel, memory_order_relaxed)) {
                    memcpy(tk, m->keys + idx * K, K);
                    memcpy(tv, m->values + idx * V, V);
                    memcpy(m->keys + idx * K, cur_key, K);
                    memcpy(m->values + idx * V, cur_val, V);
                    atomic_store_explicit(&m->flags[idx],
                                          make_flag(OR_FLAG_OCCUPIED, p),
                                          memory_order_release);
                    memcpy(cur_key, tk, K);
                    memcpy(cur_val, tv, V);
                    home = or_hash_key(cur_key, K, m->hash_mode) & m->mask;
                    p = dist + 1;
                    continue;
                }
                spin_hint();
                continue;
            }
            p += 1;
            continue;
        }
        if (flag == OR_FLAG_EMPTY || flag == OR_FLAG_TOMBSTONE) {
            uint32_t expect = f;
            if (atomic_compare_exchange_strong_explicit(
                    &m->flags[idx], &expect, make_flag(OR_FLAG_INSERTING, p),
                    memory_order_acq_rel, memory_order_relaxed)) {
                memcpy(m->keys + idx * K, cur_key, K);
                memcpy(m->values + idx * V, cur_val, V);
                atomic_store_explicit(&m->flags[idx],
                                      make_flag(OR_FLAG_OCCUPIED, p),
                                      memory_order_release);
                atomic_fetch_add_explicit(&m->counter, 1, memory_order_relaxed);
                return;
            }
            spin_hint();
            continue;
        }
        if (flag == OR_FLAG_INSERTING) { spin_hint(); continue; }
        p += 1;
    }
}

/* rayon_cpu.rs:276-321 */
static void par_remove_one(or_par_map* m, uint64_t i) {
    const uint32_t K = m->key_size;
    const uint8_t* q = m->in_keys + i * K;
    uint64_t slot = or_hash_key(q, K, m->hash_mode) & m->mask;
    uint64_t p = 0;
    while (p <= m->mask) {
        uint64_t idx = (slot + p) & m->mask;
        uint32_t f = atomic_load_explicit(&m->flags[idx], memory_order_acquire);
        uint32_t flag = get_flag(f);
        uint64_t dist = get_dist(f);
        if (flag == OR_FLAG_EMPTY) return;
        if (flag == OR_FLAG_INSERTING) { spin_hint(); continue; }
        if (flag == OR_FLAG_OCCUPIED) {
            if (p > dist) return;
            if (memcmp(m->keys + idx * K, q, K) == 0) {
                uint32_t expect = f;
                if (atomic_compare_exchange_strong_explicit(
                        &m->flags[idx], &expect, OR_FLAG_TOMBSTONE,
                        memory_order_acq_rel, memory_order_relaxed))
                    atomic_fetch_add_explicit(&m->counter, 1,
                                              memory_order_relaxed);
                return;
            }
        }
        p += 1;
    }
}

static void run_job(or_par_map* m) {
    uint8_t *ck = NULL, *cv = NULL, *tk = NULL, *tv = NULL;
    if (m->job == 1) {
        ck = (uint8_t*)malloc(m->key_size + 1);
        cv = (uint8_t*)malloc(m->value_size + 1);
        tk = (uint8_t*)malloc(m->key_size + 1);
        tv = (uint8_t*)malloc(m->value_size + 1);
    }
    for (;;) {
        uint64_t b = atomic_fetch_add(&m->next, PAR_CHUNK);
        if (b >= m->n) break;
        uint64_t e = b + PAR_CHUNK < m->n ? b + PAR_CHUNK : m->n;
        for (uint64_t i = b; i < e; i++) {
            if (m->job == 0) par_get_one(m, i);
            else if (m->job == 1) par_insert_one(m, i, ck, cv, tk, tv);
            else par_remove_one(m, i);
        }
    }
    free(ck); free(cv); free(tk); free(tv);
}

static void* worker(void* arg) {
    or_par_map* m = (or_par_map*)arg;
    uint64_t seen = 0;
    for (;;) {
        pthread_mutex_lock(&m->mu);
        while (m->generation == seen && !m->shutdown)
            pthread_cond_wait(&m->cv_start, &m->mu);
        if (m->shutdown) { pthread_mutex_unlock(&m->mu); return NULL; }
        seen = m->generation;
        pthread_mutex_unlock(&m->mu);
        run_job(m);
        pthread_mutex_lock(&m->mu);
        if (--m->pending == 0) pthread_cond_signal(&m->cv_done);
        pthread_mutex_unlock(&m->mu);
    }
}

static void dispatch(or_par_map* m) {
    atomic_store(&m->next, 0);
    atomic_store(&m->counter, 0);
    pthread_mutex_lock(&m->mu);
    m->pending = m->threads - 1;
    m->generation++;
    pthread_cond_broadcast(&m->cv_start);
    pthread_mutex_unlock(&m->mu);
    run_job(m); /* the caller participates, as rayon's install does */
    pthread_mutex_lock(&m->mu);
    while (m->pending > 0) pthread_cond_wait(&m->cv_done, &m->mu);
    pthread_mutex_unlock(&m->mu);
}

or_par_map* or_par_create(uint64_t capacity, uint32_t key_size,
                          uint32_t value_size, uint32_t hash_mode,
                          int threads) {
    or_par_map* m = (or_par_map*)calloc(1, sizeof(or_par_map));
    if (!m) return NULL;
    m->capacity = or_round_capacity(capacity);
    m->mask = m->capacity - 1;
    m->key_size = key_size;
    m->value_size = value_size;
    m->hash_mode = hash_mode;
    m->keys = (uint8_t*)calloc(m->capacity, key_size ? key_size : 1);
    m->values = (uint8_t*)calloc(m->capacity, value_size ? value_size : 1);
    m->flags = (_Atomic uint32_t*)calloc(m->capacity, sizeof(uint32_t));
    if (!m->keys || !m->values || !m->flags) { or_par_destroy(m); return NULL; }
    m->threads = threads < 1 ? 1 : threads;
    pthread_mutex_init(&m->mu, NULL);
    pthread_cond_init(&m->cv_start, NULL);
    pthread_cond_init(&m->cv_done, NULL);
    m->tids = (pthread_t*)calloc((size_t)m->threads, sizeof(pthread_t));
    for (int t = 1; t < m->threads; t++)
        pthread_create(&m->tids[t], NULL, worker, m);
    return m;
}

void or_par_destroy(or_par_map* m) {
    if (!m) return;
    if (m->tids) {
        pthread_mutex_lock(&m->mu);
        m->shutdown = 1;
        pthread_cond_broadcast(&m->cv_start);
        pthread_mutex_unlock(&m->mu);
        for (int t = 1; t < m->threads; t++) pthread_join(m->tids[t], NULL);
        free(m->tids);
        pthread_mutex_destroy(&m->mu);
        pthread_cond_destroy(&m->cv_start);
        pthread_cond_destroy(&m->cv_done);
    }
    free(m->keys);
    free(m->values);
    free((void*)m->flags);
    free(m);
}

void or_par_bulk_get(or_par_map* m, const uint8_t* query, uint64_t n,
                     uint8_t* out_values, uint8_t* out_found) {
    m->job = 0; m->in_keys = query; m->n = n;
    m->out_values = out_values; m->out_found = out_found;
    dispatch(m);
}

int or_par_bulk_insert(or_par_map* m, const uint8_t* keys,
                       const uint8_t* values, uint64_t n, uint64_t* out_new) {
    uint64_t len = atomic_load_explicit(&m->len, memory_order_relaxed);
    if (len + n > m->capacity) { if (out_new) *out_new = 0; return OR_ERR_TABLE_FULL; }
    m->job = 1; m->in_keys = keys; m->in_values = values; m->n = n;
    dispatch(m);
    uint64_t added = atomic_load(&m->counter);
    atomic_fetch_add_explicit(&m->len, added, memory_order_relaxed);
    if (out_new) *out_new = added;
    return OR_OK;
}

uint64_t or_par_bulk_remove(or_par_map* m, const uint8_t* keys, uint64_t n) {
    m->job = 2; m->in_keys = keys; m->n = n;
    dispatch(m);
    uint64_t removed = atomic_load(&m->counter);
    atomic_fetch_sub_explicit(&m->len, removed, memory_order_relaxed);
    return removed;
}

uint64_t or_par_len(const or_par_map* m) {
    return atomic_load_explicit(&((or_par_map*)m)->len, memory_order_relaxed);
}
int or_par_threads(const or_par_map* m) { return m->threads; }

/* Copies src's table into dst (same shape): flags, keys and len; the value
 * bytes too unless structure_only (no probe branches on a value, so a timing
 * loop may leave them).  Lets a timing loop restart every repetition from
 * one preloaded state. */
int or_par_copy_from(or_par_map* dst, const or_par_map* src, int structure_only) {
    if (dst->capacity != src->capacity || dst->key_size != src->key_size ||
        dst->value_size != src->value_size || dst->hash_mode != src->hash_mode)
        return -1;
    memcpy((void*)dst->flags, (const void*)src->flags, dst->capacity * sizeof(uint32_t));
    memcpy(dst->keys, src->keys, dst->capacity * (dst->key_size ? dst->key_size : 1));
    if (!structure_only)
        memcpy(dst->values, src->values, dst->capacity * (dst->value_size ? dst->value_size : 1));
    atomic_store(&dst->len, atomic_load(&((or_par_map*)src)->len));
    return 0;
}
