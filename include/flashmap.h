/*
 * flashmap.h -- C ABI of the MI355X FlashMap backend.
 *
 * Drop-in boundary for the reference's FlashMap<K,V> (src/lib.rs).  K and V
 * are erased to byte sizes; keys and values are raw little-endian bytes
 * (bytemuck views, src/gpu.rs:120, :212-217).  Each entry point cites the
 * reference interface it replaces.  All functions are thread-safe per handle
 * (calls on one handle serialise on an internal mutex; src/async_map.rs:55
 * runs concurrent gets on one map).
 *
 * Results of every bulk op are bit-exact with the reference's CPU path
 * applied in batch index order (src/rayon_cpu.rs), including duplicate keys
 * inside a batch and tombstone effects.
 */
#ifndef FLASHMAP_H
#define FLASHMAP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FM_ABI_VERSION 1

typedef struct fm_map fm_map;

/* One status per FlashMapError variant (src/error.rs:4-39), same order. */
typedef enum fm_status {
  FM_OK = 0,
  FM_ERR_DEVICE_INIT = 1,    /* CudaInit        error.rs:5-6   */
  FM_ERR_GPU_ALLOC = 2,      /* GpuAlloc        error.rs:8-9   */
  FM_ERR_KERNEL_LAUNCH = 3,  /* KernelLaunch    error.rs:11-12 */
  FM_ERR_TRANSFER = 4,       /* Transfer        error.rs:14-15 */
  FM_ERR_TABLE_FULL = 5,     /* TableFull{..}   error.rs:17-24 */
  FM_ERR_ZERO_CAPACITY = 6,  /* ZeroCapacity    error.rs:26-27 */
  FM_ERR_NO_BACKEND = 7,     /* NoBackend       error.rs:29-30 */
  FM_ERR_GPU_REQUIRED = 8,   /* GpuRequired     error.rs:32-33 */
  FM_ERR_LOCK_POISONED = 9,  /* LockPoisoned    error.rs:35-36 */
  FM_ERR_ASYNC_JOIN = 10,    /* AsyncJoin       error.rs:38-39 */
  /* No reference counterpart: null pointer, size mismatch, misaligned
   * device buffer, key/value size outside the supported range. */
  FM_ERR_INVALID_ARGUMENT = 11
} fm_status;

/* HashStrategy (src/hash.rs:3-27); numbering = HashStrategy::to_mode. */
typedef enum fm_hash_strategy {
  FM_HASH_IDENTITY = 0,
  FM_HASH_MURMUR3 = 1
} fm_hash_strategy;

/* enum FlashMapBackend (src/lib.rs:98-103). */
typedef enum fm_backend_kind {
  FM_BACKEND_NONE = 0,
  FM_BACKEND_GPU = 1, /* FlashMapBackend::Gpu   */
  FM_BACKEND_CPU = 2  /* FlashMapBackend::Rayon */
} fm_backend_kind;

/* FlashMapBuilder (src/lib.rs:377-411). */
typedef struct fm_builder {
  uint64_t capacity;      /* FlashMapBuilder::new(capacity)       */
  uint32_t key_size;      /* size_of::<K>()                       */
  uint32_t value_size;    /* size_of::<V>()                       */
  uint32_t hash_strategy; /* .hash_strategy() lib.rs:396-399      */
  int32_t device_id;      /* .device_id()     lib.rs:402-405      */
  uint32_t force_cpu;     /* .force_cpu()     lib.rs:408-411      */
  uint32_t require_gpu;   /* no reference counterpart: fail with the GPU
                             error instead of falling back (tests, benches) */
  uint32_t quiet;         /* no reference counterpart: suppress the
                             fallback message of lib.rs:433-435          */
} fm_builder;

/* FlashMapBuilder::new (lib.rs:386-393): Identity, device 0, no force. */
void fm_builder_init(fm_builder* b, uint64_t capacity, uint32_t key_size,
                     uint32_t value_size);
/* FlashMapBuilder::build (lib.rs:414-449): GPU first, then the CPU backend,
 * printing "[flash-map] GPU unavailable (...), using CPU backend" on
 * fallback.  Capacity rounds to max(c,16).next_power_of_two() on every
 * backend (rayon_cpu.rs:64).  c == 0 gives FM_ERR_ZERO_CAPACITY on the GPU
 * tier (gpu.rs:50-52), which then falls back like any GPU error. */
fm_status fm_build(const fm_builder* b, fm_map** out);
/* FlashMap::with_capacity (lib.rs:110-112). */
fm_status fm_with_capacity(uint64_t capacity, uint32_t key_size,
                           uint32_t value_size, fm_map** out);
/* Drop for FlashMap / GpuFlashMap (frees all device memory). */
void fm_destroy(fm_map* m);

/* ---- host-facing API (lib.rs:123-207; gpu.rs:111-374) ---------------- */

/* FlashMap::bulk_get (lib.rs:124-131, gpu.rs:111-191).  keys: n*K bytes;
 * out_values: n*V bytes (misses zeroed); out_found: n bytes (1 hit, 0 miss;
 * the Option discriminant). */
fm_status fm_bulk_get(fm_map* m, const void* keys, uint64_t n,
                      void* out_values, uint8_t* out_found);
/* FlashMap::bulk_insert (lib.rs:143-150, gpu.rs:193-273).  keys n*K bytes,
 * values n*V bytes (the (K,V) tuples split into two arrays as gpu.rs:212-217
 * does).  TableFull when len + n > capacity, before any mutation.
 * out_new: number of newly occupied slots. */
fm_status fm_bulk_insert(fm_map* m, const void* keys, const void* values,
                         uint64_t n, uint64_t* out_new);
/* FlashMap::bulk_remove (lib.rs:155-162, gpu.rs:275-332). */
fm_status fm_bulk_remove(fm_map* m, const void* keys, uint64_t n,
                         uint64_t* out_removed);
/* len / capacity / load_factor / clear (lib.rs:165-207). */
uint64_t fm_len(const fm_map* m);
uint64_t fm_capacity(const fm_map* m);
double fm_load_factor(const fm_map* m);
fm_status fm_clear(fm_map* m);
/* Which FlashMapBackend arm was built (lib.rs:357-362). */
fm_backend_kind fm_backend(const fm_map* m);
uint32_t fm_key_size(const fm_map* m);
uint32_t fm_value_size(const fm_map* m);
/* impl Debug for FlashMap (lib.rs:353-370):
 * FlashMap { backend: "GPU", len: N, capacity: C, load_factor: "x.y%" } */
fm_status fm_debug_string(const fm_map* m, char* buf, size_t buf_len);

/* Error detail of the calling thread's last failed call (Display text of
 * the FlashMapError, error.rs:4-39).  Returns the untruncated length. */
size_t fm_last_error_message(char* buf, size_t buf_len);
/* TableFull fields of the calling thread's last FM_ERR_TABLE_FULL
 * (error.rs:17-24; load_factor in percent as gpu.rs:204). */
fm_status fm_last_table_full(uint64_t* occupied, uint64_t* capacity,
                             double* load_factor_pct);

/* ---- device-resident API (lib.rs:213-350, gpu.rs:376-712) -------------
 * GPU backend only; the CPU backend returns FM_ERR_GPU_REQUIRED
 * (lib.rs:235).  Device buffers come from fm_alloc_device / fm_upload (or
 * any device allocation on the map's device, 16-byte aligned).  Work is
 * ordered on the map's stream.  Calls that return a count return once the
 * op's last kernel has handed the count to host memory: the caller's
 * buffers are read and the op's table writes are visible to later work on
 * the map's stream, which runs after them.  Work on any other stream (or a
 * host read of the caller's buffers after a device write by the op) must
 * call fm_synchronize first: that kernel may still be exiting and its
 * writes are not yet released beyond the map's stream. */

/* FlashMap::device (lib.rs:218-224): the hipStream_t of the map. */
void* fm_device_stream(fm_map* m);
int32_t fm_device_id(const fm_map* m);
/* FlashMap::alloc_device (lib.rs:251-257, gpu.rs:401-406): zeroed. */
fm_status fm_alloc_device(fm_map* m, uint64_t bytes, void** out_dptr);
/* Drop for CudaSlice. */
fm_status fm_free_device(fm_map* m, void* dptr);
/* upload_keys / upload_values (lib.rs:231-247, gpu.rs:386-399). */
fm_status fm_upload(fm_map* m, const void* host, uint64_t bytes,
                    void** out_dptr);
/* download (lib.rs:261-267, gpu.rs:409-413). */
fm_status fm_download(fm_map* m, const void* dptr, uint64_t bytes,
                      void* host);
/* bulk_get_device (lib.rs:274-284, gpu.rs:422-478).  The caller passes the
 * output buffers (count*V and count bytes) instead of receiving fresh
 * allocations; asynchronous on the map's stream. */
fm_status fm_bulk_get_device(fm_map* m, const void* d_keys, uint64_t count,
                             void* d_out_values, uint8_t* d_out_found);
/* bulk_get_values_device (lib.rs:291-301, gpu.rs:484-536): misses zeroed. */
fm_status fm_bulk_get_values_device(fm_map* m, const void* d_keys,
                                    uint64_t count, void* d_out_values);
/* bulk_insert_device (lib.rs:307-318, gpu.rs:541-611). */
fm_status fm_bulk_insert_device(fm_map* m, const void* d_keys,
                                const void* d_values, uint64_t count,
                                uint64_t* out_new);
/* No reference counterpart: fm_bulk_insert_device for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same results as fm_bulk_insert_device; it calls no
 * hipMalloc, hipFree or allocating routine: its per-call buffers live in one
 * 32 MiB device array per GPU, held until the process ends, so count is at
 * most FM_NOALLOC_MAX_KEYS (FM_ERR_INVALID_ARGUMENT beyond).  Every such call
 * on a GPU shares that GPU's array: calls that overlap on one GPU (on maps
 * used from different threads) race.  The GPU backend only (GpuRequired on
 * the CPU arm). */
#define FM_NOALLOC_MAX_KEYS (1ull << 20)
fm_status fm_bulk_insert_device_noalloc(fm_map* m, const void* d_keys,
                                        const void* d_values, uint64_t count,
                                        uint64_t* out_new);
/* bulk_insert_device_uncounted (lib.rs:325-338, gpu.rs:618-672): no
 * TableFull check, no readback, len += count. */
fm_status fm_bulk_insert_device_uncounted(fm_map* m, const void* d_keys,
                                          const void* d_values,
                                          uint64_t count);
/* No reference counterpart (the device family has no remove): device-key
 * variant of bulk_remove, same semantics. */
fm_status fm_bulk_remove_device(fm_map* m, const void* d_keys, uint64_t count,
                                uint64_t* out_removed);
/* recount (lib.rs:344-350, gpu.rs:677-712): len = number of OCCUPIED slots. */
fm_status fm_recount(fm_map* m, uint64_t* out_len);
/* Waits for all work queued on the map's stream. */
fm_status fm_synchronize(fm_map* m);

/* ---- logical slot array (test support; no reference counterpart) ------
 * flags: capacity u32 (dist<<4 | state, kernels.cu:13-30); keys/values:
 * capacity*K / capacity*V bytes, zeroed where the slot is not OCCUPIED. */
fm_status fm_export_slots(fm_map* m, uint32_t* flags, void* keys,
                          void* values);
fm_status fm_import_slots(fm_map* m, const uint32_t* flags, const void* keys,
                          const void* values, uint64_t len);

/* No reference counterpart: copies src's table into dst, which must have the
 * same device, capacity, K, V and hash strategy.  Asynchronous on dst's
 * stream (src is quiesced first).  FM_COPY_ALL copies flags, keys, values,
 * bitmaps and len; FM_COPY_STRUCTURE leaves dst's value bytes as they are
 * (values never steer control flow or addresses, so a structure copy resets
 * the work a later op will do; used between benchmark steps). */
#define FM_COPY_ALL 0u
#define FM_COPY_STRUCTURE 1u
fm_status fm_copy_table(fm_map* dst, fm_map* src, uint32_t parts);

/* ---- per-kernel timing (no reference counterpart) ----------------------
 * When enabled, every kernel the map launches is bracketed by hipEvents on
 * the map's stream; fm_profile_read waits for the stream and reports, per
 * kernel name, the launch count and summed device milliseconds since the
 * last fm_profile_enable. */
typedef struct fm_kernel_stat {
  char name[32];
  uint64_t launches;
  double total_ms;
} fm_kernel_stat;
fm_status fm_profile_enable(fm_map* m, int enable);
fm_status fm_profile_read(fm_map* m, fm_kernel_stat* out, uint32_t max,
                          uint32_t* count);
/* Limits the timing events to the launches of one kernel (its name as
 * fm_profile_read reports it; NULL or "" times every kernel again), so a
 * timed loop pays for two event records per launch of that kernel only.
 * Unknown names are FM_ERR_INVALID_ARGUMENT.  Applies to the launches
 * made after the call. */
fm_status fm_profile_only(fm_map* m, const char* kernel_name);

/* ---- launch tuning (no reference counterpart) --------------------------
 * group: lanes that probe one key together (1,2,4,8,16);
 * block: threads per workgroup (64..1024, multiple of 64). */
typedef struct fm_tuning {
  uint32_t get_group;
  uint32_t get_block;
  uint32_t remove_group;
  uint32_t insert_block;
  uint32_t insert_coop_arc; /* closures whose barrier-to-barrier arc is longer
                               are walked by a whole wave (0: all of them) */
  uint32_t insert_fused;    /* must be 0: kept for the layout of the struct
                               (a cooperative single-launch insert measured
                               slower than the kernel sequence and was
                               removed); anything else is
                               FM_ERR_INVALID_ARGUMENT */
  uint32_t insert_dedup;    /* 1: fold later copies of a key within an insert
                               batch into its first copy before claiming slots
                               (exact; pays off for batches with hot keys;
                               always on inside fm_serve_log); 0 (default) */
} fm_tuning;
fm_status fm_get_tuning(const fm_map* m, fm_tuning* out);
fm_status fm_set_tuning(fm_map* m, const fm_tuning* t);
/* Get autotuner (no reference counterpart; bulk_get launches 8 queries of
 * 32 lanes per 256-thread block, kernels.cu:365-468 / gpu.rs:141-167).
 * Times fm_bulk_get_device on count keys at d_keys for every get_group in
 * {1,2,4,8,16} x get_block in {64,...,1024} (median wall time of `reps`
 * synchronised calls, 1..64, after one untimed call) and sets the fastest;
 * the table is not changed.  GPU maps only. */
fm_status fm_autotune_get(fm_map* m, const void* d_keys, uint64_t count,
                          uint32_t reps, fm_tuning* out_best,
                          double* out_get_us);

/* ---- serving loop (no reference counterpart) ---------------------------
 * Closest reference analogue: AsyncFlashMap (src/async_map.rs:32-106), which
 * runs each call on a blocking thread under an RwLock with no ordering
 * guarantee.  fm_serve_log replays a log of requests against one GPU map as
 * a continuous-batching server: requests are admitted in arrival order as
 * device staging memory frees up, adjacent requests of the same kind are
 * coalesced into one device batch, and the host->device copy, the table op
 * and the device->host copy of consecutive batches overlap on three streams.
 * Results equal applying the requests one by one in arrival order (TableFull
 * included: an insert that would overflow the table fails alone).  The map
 * is used exclusively for the duration of the call. */
typedef enum fm_op { FM_OP_GET = 0, FM_OP_INSERT = 1, FM_OP_REMOVE = 2 } fm_op;

typedef struct fm_serve_request {
  uint32_t op;          /* fm_op */
  uint32_t reserved;
  uint64_t n;           /* keys in the request */
  const void* keys;     /* host, n*K */
  const void* values;   /* host, n*V (insert) */
  void* out_values;     /* host, n*V (get; misses zeroed) */
  uint8_t* out_found;   /* host, n (get), nullable */
  uint8_t* out_flags;   /* host, n, nullable: insert 1 = claimed a free
                           slot, remove 1 = removed */
  double arrival_s;     /* offered arrival time, seconds from the start */
  /* results */
  uint64_t count;       /* get: hits; insert: new slots; remove: removed */
  int32_t status;       /* FM_OK or FM_ERR_TABLE_FULL */
  int32_t batch;        /* device batch the request ran in */
  double admit_s;       /* when its staging was granted */
  double done_s;        /* when its results were in host memory */
} fm_serve_request;

typedef struct fm_serve_config {
  uint64_t staging_bytes;  /* device + pinned staging ring (each) */
  uint64_t max_batch_keys; /* coalescing limit per device batch */
  uint32_t realtime;       /* 1: release requests at arrival_s; 0: all at t=0 */
  uint32_t reserved;
} fm_serve_config;

typedef struct fm_serve_stats {
  uint64_t requests, batches, keys;
  double wall_s;           /* first arrival to last completion */
  double p50_ms, p99_ms, mean_ms, max_ms; /* done_s - arrival_s */
  uint64_t max_inflight_batches;
  uint64_t staging_waits;  /* admissions that waited for staging */
} fm_serve_stats;

fm_status fm_serve_log(fm_map* m, fm_serve_request* reqs, uint64_t n_reqs,
                       const fm_serve_config* cfg, fm_serve_stats* stats);

/* Pinned host memory (no reference counterpart).  A host buffer (keys,
 * values, outputs) that lies inside one allocation from fm_host_alloc is
 * moved by DMA directly: by the serving loop between a request and the
 * device staging ring, and by the host-buffer calls (fm_bulk_*) of 1 MiB or
 * more between the caller and the device.  Other buffers of that size go
 * through pinned staging by a host copy.  Results are the same either way.
 * fm_host_free takes exactly a pointer returned by fm_host_alloc. */
fm_status fm_host_alloc(uint64_t bytes, void** out);
fm_status fm_host_free(void* p);

/* ---- table groups: many independent maps, fused launches ---------------
 * No reference counterpart (the reference launches once per call per map,
 * gpu.rs:141-167).  A group holds n independent tables with one key/value
 * size and hash strategy on one device; each op runs over all tables in a
 * few persistent launches, and fm_group_step replays one captured hipGraph
 * for insert -> get -> remove on every table.  Per table the results equal n
 * separate FlashMaps applying the same batches in the same order, TableFull
 * included (that table's insert batch is skipped, status FM_ERR_TABLE_FULL).
 *
 * Batches are concatenated in table order in device buffers; counts[t] is
 * the number of keys for table t (0 allowed).  Outputs per key follow the
 * same concatenation; per-table outputs are host arrays of n entries. */
typedef struct fm_group fm_group;

typedef struct fm_group_tuning {
  uint32_t group;   /* lanes per key in get/remove probes: 1,2,4,8,16     */
  uint32_t block;   /* threads per workgroup = keys per chunk: 64..512    */
  uint32_t grid;    /* persistent workgroups per launch: 64..8192         */
  uint32_t coop_arc; /* closures with a longer barrier-to-barrier arc are
                        walked by a whole wave (0: all of them)          */
} fm_group_tuning;

fm_status fm_group_create(uint32_t n_tables, const uint64_t* capacities,
                          uint32_t key_size, uint32_t value_size,
                          uint32_t hash_strategy, int32_t device_id,
                          fm_group** out);
void fm_group_destroy(fm_group* g);
uint32_t fm_group_tables(const fm_group* g);
uint64_t fm_group_len(const fm_group* g, uint32_t table);
uint64_t fm_group_capacity(const fm_group* g, uint32_t table);
void* fm_group_stream(fm_group* g);
fm_status fm_group_alloc_device(fm_group* g, uint64_t bytes, void** out);
fm_status fm_group_free_device(fm_group* g, void* p);
fm_status fm_group_upload(fm_group* g, const void* host, uint64_t bytes,
                          void** out);
fm_status fm_group_download(fm_group* g, const void* dptr, uint64_t bytes,
                            void* host);
/* bulk_insert on every table (out_new, out_status: n entries each). */
fm_status fm_group_bulk_insert_device(fm_group* g, const void* d_keys,
                                      const void* d_values,
                                      const uint64_t* counts,
                                      uint64_t* out_new, int32_t* out_status);
/* bulk_get on every table (outputs per key, misses zeroed). */
fm_status fm_group_bulk_get_device(fm_group* g, const void* d_keys,
                                   const uint64_t* counts,
                                   void* d_out_values, uint8_t* d_out_found);
/* bulk_remove on every table (out_removed: n entries). */
fm_status fm_group_bulk_remove_device(fm_group* g, const void* d_keys,
                                      const uint64_t* counts,
                                      uint64_t* out_removed);
/* insert -> get -> remove on every table as one hipGraph launch (captured
 * on first use and re-captured when buffers or tuning change). */
fm_status fm_group_step_device(fm_group* g, const void* d_ins_keys,
                               const void* d_ins_values,
                               const uint64_t* ins_counts,
                               const void* d_get_keys,
                               const uint64_t* get_counts, void* d_out_values,
                               uint8_t* d_out_found, const void* d_rem_keys,
                               const uint64_t* rem_counts, uint64_t* out_new,
                               int32_t* out_status, uint64_t* out_removed);
fm_status fm_group_export_slots(fm_group* g, uint32_t table, uint32_t* flags,
                                void* keys, void* values);
fm_status fm_group_copy(fm_group* dst, fm_group* src);
fm_status fm_group_get_tuning(const fm_group* g, fm_group_tuning* out);
fm_status fm_group_set_tuning(fm_group* g, const fm_group_tuning* t);
/* Autotuner (no reference counterpart).  Times fm_group_step_device on the
 * given batches for candidate tunings and sets the fastest on g.  Each run
 * steps a scratch group restored from g's tables first, so g's tables are
 * left unchanged; every candidate must give the same per-table counts and
 * statuses (else FM_ERR_KERNEL_LAUNCH).  candidates/n_candidates: a list to
 * try in full, or NULL/0 for a coordinate descent from g's current tuning
 * over group {1,2,4,8,16}, block {64..512}, grid {256..8192} and coop_arc
 * {0,16,64,256,4096}.  Per candidate: one untimed step, then the median wall
 * time of `reps` (1..64) steps.  scratch: a group of g's shape, or NULL to
 * allocate one for the call (as much device memory as g).  d_out_values /
 * d_out_found receive one step's get results on g's current state. */
fm_status fm_group_autotune(fm_group* g, fm_group* scratch,
                            const void* d_ins_keys, const void* d_ins_values,
                            const uint64_t* ins_counts, const void* d_get_keys,
                            const uint64_t* get_counts, void* d_out_values,
                            uint8_t* d_out_found, const void* d_rem_keys,
                            const uint64_t* rem_counts,
                            const fm_group_tuning* candidates,
                            uint32_t n_candidates, uint32_t reps,
                            fm_group_tuning* out_best, double* out_step_us,
                            uint32_t* out_evaluated);

#ifdef __cplusplus
}
#endif
#endif /* FLASHMAP_H */
