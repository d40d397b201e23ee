// fm_kernels.h -- gfx950 device code for FlashMap bulk ops.
//
// Table layout in HBM (structure of arrays, one set per map):
//   flags[cap]      u32   dist<<4 | state  (EMPTY 0, OCCUPIED 1, TOMBSTONE 2;
//                         the same word as src/kernels.cu:13-30)
//   keys[cap*K]     bytes
//   vals[cap*V]     bytes
//   freebits[cap/64]  u64  bit set <=> slot is EMPTY or TOMBSTONE
//   emptybits[cap/64] u64  bit set <=> slot is EMPTY
//   owner[cap]      u32   per-slot scratch of inserts: the batch index of
//                         the key that dry-claimed the slot (read only at
//                         slots dry-claimed by the same batch, never reset;
//                         replay_arc_canon also keeps its final records in
//                         the entries of the arc it replays)
//   rowner[cap]     u32   per-slot scratch of removes (conflict detection),
//                         0xFFFFFFFF between ops
//
// Exactness contract: every op produces exactly the logical slot array that
// the reference's CPU path (src/rayon_cpu.rs) produces when it applies the
// batch one key at a time in index order.
//
//  * get: read-only; any schedule that takes the first match before the first
//    terminator is exact (rayon_cpu.rs:86-121).
//  * insert: a sequential insert touches only slots h..e, where e is at or
//    before the first free slot at/after its home h (walks move forward one
//    slot at a time; an eviction restarts the carried key at the next slot,
//    rayon_cpu.rs:211-216).  A "dry" parking pass in which every batch key
//    claims the first free bit at/after its home yields, order-independently,
//    a superset of the slots the real inserts consume.  Slots still free after
//    it ("barriers") therefore stay free through the whole sequential batch,
//    and cut the ring into closures that no sequential walk crosses.  Keys
//    are grouped by closure (stable radix sort on the closure's barrier) and
//    each closure is replayed serially in batch order by one lane.
//  * remove: removes never create EMPTY or OCCUPIED slots.  Each key's first
//    event (match / terminator) on the pre-batch table is final unless an
//    earlier key of the batch tombstones that very slot (a duplicate key, or
//    the terminator of a later key).  Such conflicts are detected with a
//    per-slot atomicMin of the matching key index; conflicting batches are
//    replayed exactly, grouped by EMPTY-delimited runs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fm_hash.h"

namespace fm {
namespace dev {

constexpr uint32_t kEmpty = 0u;
constexpr uint32_t kOccupied = 1u;
constexpr uint32_t kTombstone = 2u;
constexpr uint32_t kMaxGenericSize = 256u;  // FM_MAX_KV_SIZE, kernels.cu:32
constexpr uint32_t kNoOwner = 0xFFFFFFFFu;

// Status bits written by kernels.
constexpr uint32_t kStatusDryFail = 1u;   // dry pass found no free slot
constexpr uint32_t kStatusConflict = 2u;  // remove batch needs exact replay
constexpr uint32_t kStatusBudget = 4u;    // an insert walk exceeded cap+2 steps
constexpr uint32_t kStatusBigPending = 16u;  // host copy only: k_ins_big and the settle still to run
constexpr uint32_t kStatusConflictAll = 32u; // remove conflict with no EMPTY slot: the whole batch is one run

// Event kinds of a probe (get / remove).
constexpr uint32_t kEvNone = 0u;      // probed all cap slots
constexpr uint32_t kEvMatch = 1u;     // OCCUPIED, equal key
constexpr uint32_t kEvTermOcc = 2u;   // OCCUPIED with probe > dist (RH exit)
constexpr uint32_t kEvTermEmpty = 3u; // EMPTY

struct Table {
  uint8_t* keys;
  uint8_t* vals;
  uint32_t* flags;
  unsigned long long* freebits;
  unsigned long long* emptybits;
  uint32_t* owner;
  uint32_t* rowner;  // single-map removes (table groups keep theirs in owner)
  uint64_t cap;
  uint64_t mask;
  uint32_t K;
  uint32_t V;
  uint32_t mode;
  // Per-op, not table state: when set, inserts write 1/0 per batch key for
  // "claimed a free slot" and removes 1/0 for "tombstoned a slot".
  uint8_t* outcome;
  // Per-op: when set, a copy of freebits that removes keep equal to it (the
  // claim bitmap, so the next insert needs no copy).
  unsigned long long* freecopy;
  // Per-op insert deduplication (null when off): keys are bucketed by home;
  // dd_first[b] = smallest batch index in bucket b, dd_last[b] = largest
  // index of a later copy of that key; dd_bucket[i] = bucket of key i, with
  // kDupKey set when key i is a later copy of its bucket's first key.
  uint32_t* dd_first;
  uint32_t* dd_last;
  uint32_t* dd_bucket;
  uint64_t dd_smask;
  // k_ins_big replays an arc with at least this many members from its
  // groups of equal homes (replay_arc_canon) when the arc is canonical; 0:
  // never.
  uint32_t canon_min;
  // Per-op insert, when set: the bitmaps the map holds after this batch.
  // The dry pass copies freebits into nfree and nq and emptybits into
  // nempty; every real claim then clears its bit in all three (claim_next)
  // as the replay makes it, while freebits, emptybits and the claim bitmap
  // stay as the batch found them for the closure elections.  No settle pass.
  // The three are one block: nempty == nfree + words, nq == nfree + 2 words.
  unsigned long long* nfree;
  unsigned long long* nempty;
  unsigned long long* nq;
};

// A real claim of slot s in the next bitmaps (see Table::nfree).
__device__ __forceinline__ void claim_next(const Table& t, uint64_t s) {
  const unsigned long long m = ~(1ull << (s & 63));
  atomicAnd(&t.nfree[s >> 6], m);
  atomicAnd(&t.nempty[s >> 6], m);
  atomicAnd(&t.nq[s >> 6], m);
}

constexpr uint32_t kDupKey = 0x80000000u;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// Insert deduplication.  With no removes inside a batch, every later copy of
// a key only overwrites the value that its first copy inserted or matched
// (the first copy leaves no free slot between the key's home and the key, and
// later inserts keep it reachable), so the batch is applied exactly as its
// first copies, each carrying the value of its key's last copy; later copies
// claim nothing.  Only copies of each bucket's first key are folded (keys
// that merely share a bucket stay separate batch entries).
__device__ __forceinline__ bool dd_is_copy(const Table& t, uint64_t i) {
  return t.dd_first && (t.dd_bucket[i] & kDupKey);
}
__device__ __forceinline__ uint64_t dd_value_index(const Table& t, uint64_t i) {
  if (!t.dd_first) return i;
  const uint32_t b = t.dd_bucket[i] & ~kDupKey;
  if (t.dd_first[b] != i) return i;
  const uint64_t last = t.dd_last[b];
  return last > i ? last : i;
}

struct Counters {
  uint32_t status;
  uint32_t nbig;  // closures queued for the wave replay
  unsigned long long count;
  uint32_t done;      // completion tickets of a settle pass
  uint32_t nsmall;    // k_ins_big: work ticket
};

__device__ __forceinline__ uint32_t make_flag(uint32_t state, uint64_t dist) {
  return (static_cast<uint32_t>(dist) << 4) | (state & 0xFu);
}

// ---------------------------------------------------------------------------
// Fixed-size byte blobs held in registers (keys and values).
// N > 0: N % 4 == 0, vector loads of the widest width that divides N
// (buffers are aligned to that width).  N == 0: runtime size, bytes.
// ---------------------------------------------------------------------------
template <int N>
struct Blob {
  static_assert(N % 4 == 0, "static blob sizes are multiples of 4");
  static constexpr int W = N / 4;
  uint32_t w[W];

  __device__ __forceinline__ void load(const uint8_t* __restrict__ p, uint32_t) {
    if constexpr (N % 16 == 0) {
#pragma unroll
      for (int i = 0; i < N / 16; ++i) {
        const uint4 v = reinterpret_cast<const uint4*>(p)[i];
        w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
      }
    } else if constexpr (N % 8 == 0) {
#pragma unroll
      for (int i = 0; i < N / 8; ++i) {
        const uint2 v = reinterpret_cast<const uint2*>(p)[i];
        w[2 * i] = v.x; w[2 * i + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int i = 0; i < W; ++i) w[i] = reinterpret_cast<const uint32_t*>(p)[i];
    }
  }
  __device__ __forceinline__ void store(uint8_t* __restrict__ p, uint32_t) const {
    if constexpr (N % 16 == 0) {
#pragma unroll
      for (int i = 0; i < N / 16; ++i)
        reinterpret_cast<uint4*>(p)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
    } else if constexpr (N % 8 == 0) {
#pragma unroll
      for (int i = 0; i < N / 8; ++i) reinterpret_cast<uint2*>(p)[i] = make_uint2(w[2 * i], w[2 * i + 1]);
    } else {
#pragma unroll
      for (int i = 0; i < W; ++i) reinterpret_cast<uint32_t*>(p)[i] = w[i];
    }
  }
  __device__ __forceinline__ bool eq(const uint8_t* __restrict__ p, uint32_t n) const {
    Blob<N> o;
    o.load(p, n);
    uint32_t diff = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) diff |= w[i] ^ o.w[i];
    return diff == 0;
  }
  __device__ __forceinline__ bool eqb(const Blob<N>& o) const {
    uint32_t diff = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) diff |= w[i] ^ o.w[i];
    return diff == 0;
  }
  __device__ __forceinline__ uint64_t hash(uint32_t, uint32_t mode) const {
    if (mode == 0) {
      if constexpr (N >= 8) return static_cast<uint64_t>(w[0]) | (static_cast<uint64_t>(w[1]) << 32);
      else return static_cast<uint64_t>(w[0]);
    }
    return murmur3_words(w, N);
  }
};

template <>
struct Blob<0> {
  uint8_t b[kMaxGenericSize];
  __device__ __forceinline__ void load(const uint8_t* __restrict__ p, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) b[i] = p[i];
  }
  __device__ __forceinline__ void store(uint8_t* __restrict__ p, uint32_t n) const {
    for (uint32_t i = 0; i < n; ++i) p[i] = b[i];
  }
  __device__ __forceinline__ bool eq(const uint8_t* __restrict__ p, uint32_t n) const {
    for (uint32_t i = 0; i < n; ++i)
      if (b[i] != p[i]) return false;
    return true;
  }
  __device__ __forceinline__ bool eqb(const Blob<0>& o) const {
    for (uint32_t i = 0; i < kMaxGenericSize; ++i)
      if (b[i] != o.b[i]) return false;
    return true;
  }
  __device__ __forceinline__ uint64_t hash(uint32_t n, uint32_t mode) const { return hash_key(b, n, mode); }
};

// ---------------------------------------------------------------------------
// Bitmap helpers.  Bits at positions >= cap in the last word are always 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long load_word(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Keeps two loads issued together: an empty asm statement that reads the
// loaded registers, so the compiler cannot move the second load into the one
// branch that uses it (which would add a dependent round trip there).
template <int N>
__device__ __forceinline__ void keep_loaded(uint32_t f, const Blob<N>& b) {
  asm volatile("" ::"v"(f));
#pragma unroll
  for (int i = 0; i < Blob<N>::W; ++i) asm volatile("" ::"v"(b.w[i]));
}


// First set bit at or after `start`, cyclically; cap if there is none.
__device__ inline uint64_t bm_next_set(const unsigned long long* bm, uint64_t start, uint64_t cap) {
  const uint64_t words = (cap + 63) >> 6;
  uint64_t w = start >> 6;
  unsigned long long word = bm[w] & (~0ull << (start & 63));
  for (uint64_t it = 0; it <= words; ++it) {
    if (word) return (w << 6) + static_cast<uint64_t>(__ffsll(word) - 1);
    w = (w + 1 == words) ? 0 : w + 1;
    word = bm[w];
  }
  return cap;
}

// bm_next_set over a bitmap that other workgroups of the same launch modify
// with atomics (every word read with a device-scope load).
__device__ inline uint64_t bm_next_set_agent(const unsigned long long* bm, uint64_t start, uint64_t cap) {
  const uint64_t words = (cap + 63) >> 6;
  uint64_t w = start >> 6;
  unsigned long long word = load_word(&bm[w]) & (~0ull << (start & 63));
  for (uint64_t it = 0; it <= words; ++it) {
    if (word) return (w << 6) + static_cast<uint64_t>(__ffsll(word) - 1);
    w = (w + 1 == words) ? 0 : w + 1;
    word = load_word(&bm[w]);
  }
  return cap;
}

// Dry-pass claim: atomically clear the first set bit at/after `start`
// (cyclically) and return its position.  Linearisable: bits are only ever
// cleared while the pass runs, so the bit won is the first free one at the
// instant of the successful atomic.  `budget` bounds the loop.
__device__ inline bool bm_claim(unsigned long long* bm, uint64_t start, uint64_t cap, uint64_t budget,
                                uint64_t* out) {
  const uint64_t words = (cap + 63) >> 6;
  uint64_t w = start >> 6;
  unsigned long long m = ~0ull << (start & 63);
  // The start bit first, with no read before it: the atomic's old word shows
  // the word's other free bits when the start slot is already taken.
  unsigned long long cur;
  {
    const unsigned long long bit = 1ull << (start & 63);
    const unsigned long long old = atomicAnd(&bm[w], ~bit);
    if (old & bit) {
      *out = start;
      return true;
    }
    cur = old & m;
  }
  for (uint64_t it = 0; it < budget; ++it) {
    while (cur) {
      const int b = __ffsll(cur) - 1;
      const unsigned long long bit = 1ull << b;
      const unsigned long long old = atomicAnd(&bm[w], ~bit);
      if (old & bit) {
        *out = (w << 6) + static_cast<uint64_t>(b);
        return true;
      }
      cur = old & m;
      if (++it >= budget) return false;
    }
    m = ~0ull;
    w = (w + 1 == words) ? 0 : w + 1;
    cur = load_word(&bm[w]);
  }
  return false;
}

// ---------------------------------------------------------------------------
// Cooperative probe by a group of G lanes (G power of two, 1..64): lane l of
// the group inspects slot home+base+l; wave-wide ballots find the first
// event (match or terminator) of each group.  Semantics: rayon_cpu.rs:91-120
// (probe at most cap slots; EMPTY ends; OCCUPIED with p > dist ends; an
// equal key at OCCUPIED matches; tombstones and other keys continue).
// Home filter: an equal key sits at displacement p == dist (stored distances
// equal the displacement mod cap), so keys are compared only there.
// ---------------------------------------------------------------------------
template <int KS, int G>
__device__ __forceinline__ void group_probe(const Table& t, const Blob<KS>& q, uint32_t K, uint64_t home,
                                            uint32_t gl, uint32_t gshift, uint32_t* kind, uint64_t* slot) {
  constexpr unsigned long long kGroupMask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
  for (uint64_t base = 0;; base += G) {
    const uint64_t p = base + gl;
    const bool valid = p <= t.mask;
    const uint64_t idx = (home + p) & t.mask;
    const uint32_t f = valid ? t.flags[idx] : 0u;
    const uint32_t st = f & 0xFu;
    const uint64_t d = f >> 4;
    const bool occ = st == kOccupied;
    bool match = false;
    if (valid && occ && d == p) match = q.eq(t.keys + idx * K, K);
    const bool term = !valid || st == kEmpty || (occ && p > d);
    const uint32_t my_kind = match ? kEvMatch : (!valid ? kEvNone : (st == kEmpty ? kEvTermEmpty : kEvTermOcc));
    if constexpr (G == 1) {
      if (match || term) {
        *kind = my_kind;
        *slot = idx;
        return;
      }
    } else {
      const unsigned long long ev = ((__ballot(match) | __ballot(term)) >> gshift) & kGroupMask;
      if (ev) {
        const int first = __ffsll(ev) - 1;
        *kind = static_cast<uint32_t>(__shfl(static_cast<int>(my_kind), first, G));
        *slot = static_cast<uint64_t>(__shfl(static_cast<long long>(idx), first, G));
        return;
      }
    }
  }
}

// Group-cooperative copy of V bytes (src == nullptr writes zeros).
template <int VS, int G>
__device__ __forceinline__ void group_copy(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, uint32_t V,
                                           uint32_t gl) {
  if constexpr (VS != 0 && VS % 16 == 0) {
#pragma unroll
    for (uint32_t c = gl; c < VS / 16; c += G)
      reinterpret_cast<uint4*>(dst)[c] = src ? reinterpret_cast<const uint4*>(src)[c] : make_uint4(0, 0, 0, 0);
  } else if constexpr (VS != 0 && VS % 8 == 0) {
#pragma unroll
    for (uint32_t c = gl; c < VS / 8; c += G)
      reinterpret_cast<uint2*>(dst)[c] = src ? reinterpret_cast<const uint2*>(src)[c] : make_uint2(0, 0);
  } else if constexpr (VS != 0 && VS % 4 == 0) {
#pragma unroll
    for (uint32_t c = gl; c < VS / 4; c += G)
      reinterpret_cast<uint32_t*>(dst)[c] = src ? reinterpret_cast<const uint32_t*>(src)[c] : 0u;
  } else {
    for (uint32_t c = gl; c < V; c += G) dst[c] = src ? src[c] : 0;
  }
}

// Adds per-thread counts into *dst with one atomic per wave.
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint32_t v) {
  unsigned long long sum = v;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if ((threadIdx.x & 63) == 0 && sum) atomicAdd(dst, sum);
}

// Adds per-thread counts into *dst with one atomic per workgroup: atomics
// on one address serialise (about 14 ns each on MI355X), so counters are
// reduced in the workgroup first.  Called by every thread of the workgroup.
// Workgroup sum of v, valid in thread 0.
__device__ inline unsigned long long block_sum(unsigned long long v) {
  __shared__ unsigned long long part[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();  // part[] may still be read by a previous call
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long sum = 0;
  if (threadIdx.x == 0)
    for (uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w) sum += part[w];
  return sum;
}

__device__ inline void block_add(unsigned long long* dst, unsigned long long v) {
  const unsigned long long sum = block_sum(v);
  if (threadIdx.x == 0 && sum) atomicAdd(dst, sum);
}

// A workgroup barrier over LDS only: the waves' global stores and atomics
// stay in flight (__syncthreads waits for every one of them first).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// block_add for a kernel whose waves end with global stores or atomics in
// flight: the reduction waits for LDS only.
__device__ inline void block_add_nowait(unsigned long long* dst, unsigned long long v) {
  __shared__ unsigned long long part[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  lds_barrier();  // part[] may still be read by a previous call
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  lds_barrier();
  if (threadIdx.x == 0) {
    unsigned long long sum = 0;
    for (uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w) sum += part[w];
    if (sum) atomicAdd(dst, sum);
  }
}

// ---------------------------------------------------------------------------
// bulk_get (rayon_cpu.rs:82-126).  One group of G lanes per query.
// out_found may be null (bulk_get_values_device, kernels.cu:481-556).
// ---------------------------------------------------------------------------
template <int KS, int VS, int G>
__global__ void __launch_bounds__(1024)
k_get(Table t, const uint8_t* __restrict__ qkeys, uint64_t n, uint8_t* __restrict__ out_vals,
      uint8_t* __restrict__ out_found) {
  const uint32_t K = KS ? KS : t.K;
  const uint32_t V = VS ? VS : t.V;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t gl = lane & (G - 1);
  const uint32_t gshift = lane & ~(G - 1u);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / G;
  for (uint64_t qi = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; qi < n; qi += stride) {
    Blob<KS> q;
    q.load(qkeys + qi * K, K);
    const uint64_t home = q.hash(K, t.mode) & t.mask;
    uint32_t kind;
    uint64_t slot;
    group_probe<KS, G>(t, q, K, home, gl, gshift, &kind, &slot);
    const bool hit = kind == kEvMatch;
    group_copy<VS, G>(out_vals + qi * V, hit ? t.vals + slot * V : nullptr, V, gl);
    if (out_found && gl == 0) out_found[qi] = hit ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// bulk_insert, exact (see header).
// ---------------------------------------------------------------------------

// Pass 1: hash every key and dry-claim one free bit of Q (a copy of
// freebits) at/after its home.  With `dslot` given, the claimed slot is
// recorded per key and the key's index is written to owner[slot].
// Deduplication pass 1: each key's bucket, and the first index per bucket.
template <int KS>
__global__ void __launch_bounds__(1024)
k_ins_dedup(Table t, const uint8_t* __restrict__ in_keys, uint64_t n) {
  const uint32_t K = KS ? KS : t.K;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    Blob<KS> k;
    k.load(in_keys + i * K, K);
    const uint32_t b = static_cast<uint32_t>((k.hash(K, t.mode) & t.mask) & t.dd_smask);
    t.dd_bucket[i] = b;
    atomicMin(&t.dd_first[b], static_cast<uint32_t>(i));
  }
}

// Deduplication reset (sort path; the leader path resets in k_ins_bits).
static __global__ void __launch_bounds__(1024)
k_dedup_reset(Table t, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t b = t.dd_bucket[i] & ~kDupKey;
    t.dd_first[b] = kNoOwner;
    t.dd_last[b] = 0;
  }
}

template <int KS>
__global__ void __launch_bounds__(1024)
k_ins_dry(Table t, const uint8_t* __restrict__ in_keys, uint64_t n, unsigned long long* __restrict__ Q,
          uint32_t* __restrict__ home_out, uint32_t* __restrict__ dslot, Counters* ctr) {
  const uint32_t K = KS ? KS : t.K;
  const uint64_t budget = ((t.cap + 63) >> 6) + n + 2;
  if (t.nfree) {  // the next bitmaps start as the pre-batch ones
    const uint64_t words = (t.cap + 63) >> 6;
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; w < words;
         w += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
      const unsigned long long f = t.freebits[w];
      t.nfree[w] = f;
      t.nq[w] = f;
      t.nempty[w] = t.emptybits[w];
    }
  }
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    Blob<KS> k;
    k.load(in_keys + i * K, K);
    const uint64_t h = k.hash(K, t.mode) & t.mask;
    if (home_out) home_out[i] = static_cast<uint32_t>(h);
    if (t.dd_first) {  // a later copy of its bucket's first key claims nothing
      const uint32_t b = t.dd_bucket[i];
      const uint32_t f = t.dd_first[b];
      if (f != i && k.eq(in_keys + static_cast<uint64_t>(f) * K, K)) {
        atomicMax(&t.dd_last[b], static_cast<uint32_t>(i));
        t.dd_bucket[i] = b | kDupKey;
        if (dslot) dslot[i] = kNoSlot;
        if (t.outcome) t.outcome[i] = 0;
        continue;
      }
    }
    uint64_t s;
    if (!bm_claim(Q, h, t.cap, budget, &s)) {
      atomicOr(&ctr->status, kStatusDryFail);
      if (dslot) dslot[i] = 0;  // never read: the leader kernels stop on kStatusDryFail
      continue;
    }
    if (dslot) {
      dslot[i] = static_cast<uint32_t>(s);
      t.owner[s] = static_cast<uint32_t>(i);
    }
  }
}

// Pass 2: closure id = first barrier (bit still set in Q) at/after home.
// With no barrier (the batch fills every free slot) or a failed dry pass,
// every key gets the same id and one lane replays the whole batch in order.
static __global__ void __launch_bounds__(1024)
k_ins_closure(uint64_t cap, const unsigned long long* __restrict__ Q, const uint32_t* __restrict__ home, uint64_t n,
              uint32_t* __restrict__ cid, uint32_t* __restrict__ idx, const Counters* ctr) {
  const bool failed = (ctr->status & kStatusDryFail) != 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    cid[i] = failed ? static_cast<uint32_t>(cap) : static_cast<uint32_t>(bm_next_set(Q, home[i], cap));
    idx[i] = static_cast<uint32_t>(i);
  }
}

// Sequential insert of one pair into the table (rayon_cpu.rs:157-262 with
// the batch applied in order).  Returns 1 if a free slot was claimed.
// Every iteration advances the walk by one slot (++p, or an eviction that
// restarts the carried key at the next slot), so with a free slot anywhere
// the walk ends within cap steps.  Only a table with no free slot at all
// (reachable through the uncounted device insert) can exceed the budget; the
// walk then stops and flags kStatusBudget instead of circling the ring.
// kPublishBits: clear the claimed slot's bitmap bits here.  The leader path
// passes false: there t.freebits is the pre-batch set other lanes are still
// scanning, and k_ins_bits publishes the claims after the replay.
// insert_walk is the same walk for a pair already in registers, entered at
// probe position p (the caller has established that positions 0..p-1 are
// OCCUPIED with dist >= position and hold no equal key); it reports the
// slot it claimed in *at.
template <int KS, int VS, bool kPublishBits>
__device__ inline uint32_t insert_walk(const Table& t, Blob<KS> ck, Blob<VS> cv, uint64_t home, uint64_t p, uint32_t K,
                                       uint32_t V, Counters* ctr, uint64_t* at) {
  for (uint64_t step = 0;; ++step) {
    if (p > t.mask) return 0;  // carried pair dropped (rayon_cpu.rs:167-169)
    if (step > t.cap + 1) {
      atomicOr(&ctr->status, kStatusBudget);
      return 0;
    }
    const uint64_t idx = (home + p) & t.mask;
    const uint32_t f = t.flags[idx];
    // Static key sizes: the slot's key is loaded with its flag word (one
    // round trip for the compare or the eviction instead of two).
    Blob<KS> rk;
    if constexpr (KS > 0) {
      rk.load(t.keys + idx * K, K);
      keep_loaded(f, rk);
    }
    const uint32_t st = f & 0xFu;
    const uint64_t d = f >> 4;
    if (st == kOccupied) {
      // equal keys share a home, hence sit at displacement p == d
      bool same;
      if constexpr (KS > 0) same = d == p && ck.eqb(rk);
      else same = d == p && ck.eq(t.keys + idx * K, K);
      if (same) {
        cv.store(t.vals + idx * V, V);  // update in place, not new
        return 0;
      }
      if (p > d) {  // Robin Hood eviction: swap and carry the resident on
        Blob<KS> ek;
        Blob<VS> ev;
        if constexpr (KS > 0) ek = rk;
        else ek.load(t.keys + idx * K, K);
        ev.load(t.vals + idx * V, V);
        ck.store(t.keys + idx * K, K);
        cv.store(t.vals + idx * V, V);
        t.flags[idx] = make_flag(kOccupied, p);
        ck = ek;
        cv = ev;
        home = ck.hash(K, t.mode) & t.mask;
        p = d + 1;
        continue;
      }
      ++p;
      continue;
    }
    // EMPTY or TOMBSTONE: claim.
    ck.store(t.keys + idx * K, K);
    cv.store(t.vals + idx * V, V);
    t.flags[idx] = make_flag(kOccupied, p);
    if constexpr (kPublishBits) {
      const unsigned long long bit = 1ull << (idx & 63);
      atomicAnd(&t.freebits[idx >> 6], ~bit);
      atomicAnd(&t.emptybits[idx >> 6], ~bit);
    }
    *at = idx;
    return 1;
  }
}

template <int KS, int VS, bool kPublishBits = true>
__device__ inline uint32_t insert_one(const Table& t, const uint8_t* __restrict__ kin, const uint8_t* __restrict__ vin,
                                      uint32_t K, uint32_t V, Counters* ctr, uint64_t* at = nullptr) {
  Blob<KS> ck;
  Blob<VS> cv;
  ck.load(kin, K);
  cv.load(vin, V);
  const uint64_t home = ck.hash(K, t.mode) & t.mask;
  uint64_t where = 0;
  const uint32_t c = insert_walk<KS, VS, kPublishBits>(t, ck, cv, home, 0, K, V, ctr, &where);
  if (at) *at = where;
  return c;
}

// Pass 3: one lane per closure replays its keys in batch order.
template <int KS, int VS>
__global__ void __launch_bounds__(1024)
k_ins_replay(Table t, const uint8_t* __restrict__ in_keys, const uint8_t* __restrict__ in_vals, uint64_t n,
             const uint32_t* __restrict__ cid_s, const uint32_t* __restrict__ idx_s, Counters* ctr) {
  const uint32_t K = KS ? KS : t.K;
  const uint32_t V = VS ? VS : t.V;
  unsigned long long claimed = 0;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x;; base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (base >= n) break;
    const uint64_t s = base + threadIdx.x;
    if (s < n && (s == 0 || cid_s[s] != cid_s[s - 1])) {
      const uint32_t c = cid_s[s];
      for (uint64_t u = s; u < n && cid_s[u] == c; ++u) {
        const uint64_t i = idx_s[u];
        if (dd_is_copy(t, i)) continue;
        const uint32_t c1 = insert_one<KS, VS>(t, in_keys + i * K, in_vals + dd_value_index(t, i) * V, K, V, ctr);
        claimed += c1;
        if (t.outcome) t.outcome[i] = static_cast<uint8_t>(c1);
      }
    }
  }
  block_add(&ctr->count, claimed);
}

// ---------------------------------------------------------------------------
// Sort-free closure replay, used when a barrier provably exists (len + n <
// cap, so the dry pass leaves a free slot).  P = t.freebits (pre-batch free
// slots; untouched until k_ins_bits), Q = free slots after the dry pass,
// N = P & ~Q = dry-claimed slots.  Every closure is an arc between two
// barriers; its N slots are exactly its keys' dry claims and owner[] maps
// each N slot to its key.  The key whose dry slot is the first N slot of the
// arc (the nearest P slot below it is a barrier) leads the closure.
// ---------------------------------------------------------------------------
constexpr int kLeaderMax = 32;  // members one leader lane orders and replays
constexpr int kBigMax = 4096;   // members a workgroup sorts in LDS at a time

__device__ __forceinline__ bool bit_at(const unsigned long long* bm, uint64_t i) {
  return (bm[i >> 6] >> (i & 63)) & 1ull;
}

// Highest set bit strictly below `start`, cyclically (`start` itself when it
// is the only set bit).
__device__ inline uint64_t bm_prev_set(const unsigned long long* bm, uint64_t start, uint64_t cap) {
  const uint64_t words = (cap + 63) >> 6;
  uint64_t w = start >> 6;
  unsigned long long word = bm[w] & ((1ull << (start & 63)) - 1ull);
  for (uint64_t it = 0; it <= words; ++it) {
    if (word) return (w << 6) + static_cast<uint64_t>(63 - __clzll(word));
    w = (w == 0) ? words - 1 : w - 1;
    word = bm[w];
  }
  return start;
}

// Orders LDS accesses between the lanes of one wave (a wave's LDS operations
// complete in issue order; the fences keep the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// The same sequential insert as insert_one, walked by a whole wave: the 64
// lanes load 64 consecutive slots (flag + key) in one round trip, the walk's
// decisions for those slots are taken in order on wave-uniform state
// (readlane of lane j = slot j), and all moves of the window are applied
// afterwards (every slot is written at most once per walk; the walk only
// reads slots ahead of the ones it has written).  Evicted residents' values
// are loaded once per window; the pair carried across windows lives in
// `carry` (LDS).  Only for the leader path (a free slot exists, so the walk
// ends within cap steps and never revisits a slot) and sizes that are
// multiples of 4.  Called by all 64 lanes of a single-wave workgroup.
// ---------------------------------------------------------------------------
template <int KS, int VS>
__device__ uint32_t insert_one_wave(const Table& t, const uint8_t* __restrict__ kin, const uint8_t* __restrict__ vin,
                                    uint32_t* carry, Counters* ctr) {
  static_assert(KS % 4 == 0 && VS % 4 == 0 && KS > 0 && VS > 0, "static word sizes");
  constexpr int KW = KS / 4, VW = VS / 4;
  const int lane = threadIdx.x & 63;
  for (int w = lane; w < KW; w += 64) carry[w] = reinterpret_cast<const uint32_t*>(kin)[w];
  for (int w = lane; w < VW; w += 64) carry[KW + w] = reinterpret_cast<const uint32_t*>(vin)[w];
  wave_sync();
  Blob<KS> ck;  // carried key, identical in every lane
#pragma unroll
  for (int w = 0; w < KW; ++w) ck.w[w] = carry[w];
  uint64_t home = ck.hash(KS, t.mode) & t.mask;
  uint64_t p = 0, pos = home, steps = 0;
  for (;;) {
    const uint64_t s = (pos + lane) & t.mask;
    const uint32_t f = t.flags[s];
    Blob<KS> rk;  // read with the flag (one round trip); unused unless occupied
    rk.load(t.keys + s * KS, KS);
    keep_loaded(f, rk);
    const bool occ = (f & 0xFu) == kOccupied;
    auto eq_ballot = [&]() {
      uint32_t diff = occ ? 0u : 1u;
#pragma unroll
      for (int w = 0; w < KW; ++w) diff |= rk.w[w] ^ ck.w[w];
      return static_cast<unsigned long long>(__ballot(diff == 0));
    };
    unsigned long long eqm = eq_ballot();
    int my_src = -2;       // new content of my slot: -2 none, -1 carry buffer, k: resident of lane k
    uint32_t my_flag = 0;  // new flag when my_src != -2 and not an update
    bool my_update = false;
    int c_src = -1;
    int end = 0;  // 0 window exhausted, 1 done (update/claim/drop)
    uint32_t claimed = 0;
    unsigned long long evm = 0;
    for (int j = 0; j < 64; ++j) {
      if (p > t.mask) {  // carried pair dropped (rayon_cpu.rs:167-169)
        end = 1;
        break;
      }
      if (++steps > t.cap + 1) {
        if (lane == 0) atomicOr(&ctr->status, kStatusBudget);
        end = 1;
        break;
      }
      const uint32_t fj = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(f), j));
      const uint32_t stj = fj & 0xFu;
      const uint64_t dj = fj >> 4;
      if (stj == kOccupied) {
        if (dj == p && ((eqm >> j) & 1ull)) {  // equal key: update its value
          if (lane == j) {
            my_src = c_src;
            my_update = true;
          }
          end = 1;
          break;
        }
        if (p > dj) {  // Robin Hood eviction: carried pair in, resident carried on
          if (lane == j) {
            my_src = c_src;
            my_flag = make_flag(kOccupied, p);
          }
          evm |= 1ull << j;
          c_src = j;
#pragma unroll
          for (int w = 0; w < KW; ++w)
            ck.w[w] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rk.w[w]), j));
          eqm = eq_ballot();
          home = ck.hash(KS, t.mode) & t.mask;
          p = dj + 1;
          continue;
        }
        ++p;
        continue;
      }
      // EMPTY or TOMBSTONE: claim.
      if (lane == j) {
        my_src = c_src;
        my_flag = make_flag(kOccupied, p);
      }
      claimed = 1;
      end = 1;
      break;
    }
    // Values of residents that were carried on (read before any write).
    Blob<VS> rv;
#pragma unroll
    for (int w = 0; w < VW; ++w) rv.w[w] = 0;
    if ((evm >> lane) & 1ull) rv.load(t.vals + s * VS, VS);
    // Fetch every written slot's new key/value from its source.  All lanes
    // take part in the shuffles.
    const int src = my_src < 0 ? lane : my_src;
    Blob<KS> nk;
    Blob<VS> nv;
#pragma unroll
    for (int w = 0; w < KW; ++w) nk.w[w] = static_cast<uint32_t>(__shfl(static_cast<int>(rk.w[w]), src, 64));
#pragma unroll
    for (int w = 0; w < VW; ++w) nv.w[w] = static_cast<uint32_t>(__shfl(static_cast<int>(rv.w[w]), src, 64));
    if (my_src == -1) {
#pragma unroll
      for (int w = 0; w < KW; ++w) nk.w[w] = carry[w];
#pragma unroll
      for (int w = 0; w < VW; ++w) nv.w[w] = carry[KW + w];
    }
    if (my_src != -2) {
      nv.store(t.vals + s * VS, VS);
      if (!my_update) {
        nk.store(t.keys + s * KS, KS);
        t.flags[s] = my_flag;
      }
    }
    if (end) return claimed;
    // Window exhausted: the pair carried into the next window goes to LDS.
    wave_sync();
    if (c_src >= 0 && lane == c_src) {
#pragma unroll
      for (int w = 0; w < KW; ++w) carry[w] = rk.w[w];
#pragma unroll
      for (int w = 0; w < VW; ++w) carry[KW + w] = rv.w[w];
    }
    wave_sync();
    pos = (pos + 64) & t.mask;
  }
}

// ---------------------------------------------------------------------------
// A whole closure replayed in LDS.  A closure owns its arc: the slots
// strictly between two barriers (free slots no walk of the batch reaches),
// and its members' walks start at homes inside the arc and never leave it.
// So the arc's flags and keys are read into LDS once, every member's walk
// (the walk of insert_one_wave: 64 slots per window, decisions in slot
// order) runs on LDS, and the arc is written back once.  Values stay in
// HBM during the replay: each arc slot carries a value source instead (the
// arc-relative slot whose pre-replay value it holds, or kArcBatch | the
// batch index of the value), and the write-back copies each moved value
// once.  Values only move forward along the arc (Robin Hood evictions push
// residents to later slots), so writing back from the end of the arc down,
// each chunk of 64 slots reading all its sources before writing, reads
// every source before it is overwritten.
// ---------------------------------------------------------------------------
constexpr uint32_t kArcLdsWords = 28672;       // 112 KiB: flags, value sources and keys of one arc
constexpr uint32_t kBigPoolWords = 38912;      // 152 KiB: k_ins_big's member list and replay arrays
constexpr uint32_t kArcBatch = 0x80000000u;    // value source: a batch value, not an arc slot

template <int KS>
__host__ __device__ constexpr uint32_t arc_lds_cap() {
  return kArcLdsWords / (2 + KS / 4);
}

// Replays members lst[0..m) (batch indices into in_keys, in batch order) of
// the closure whose arc is the L slots from r0.  Called by all 64 lanes of
// a single-wave workgroup; `lds` holds kArcLdsWords words.  Returns the free
// slots claimed (in every lane).
template <int KS, int VS>
__device__ uint32_t replay_arc_lds(const Table& t, const uint8_t* __restrict__ in_keys,
                                   const uint8_t* __restrict__ in_vals, uint64_t r0, uint32_t L,
                                   const uint32_t* lst, uint32_t m, uint32_t* lds, Counters* ctr) {
  static_assert(KS % 4 == 0 && VS % 4 == 0 && KS > 0 && VS > 0, "static word sizes");
  constexpr int KW = KS / 4;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t* fl = lds;          // flags
  uint32_t* vs = lds + L;      // value sources
  uint32_t* ky = lds + 2 * L;  // keys, KW words per slot
  for (uint32_t q = lane; q < L; q += 64) {
    const uint64_t slot = (r0 + q) & t.mask;
    fl[q] = t.flags[slot];
    vs[q] = q;
    Blob<KS> k;
    k.load(t.keys + slot * KS, KS);
#pragma unroll
    for (int w = 0; w < KW; ++w) ky[q * KW + w] = k.w[w];
  }
  wave_sync();
  uint32_t claimed = 0;
  Blob<KS> mk;       // lane l: key of member (j & ~63) + l, read 64 members at a time
  uint32_t mvi = 0;  // and its value index
  for (uint32_t j = 0; j < m; ++j) {
    if ((j & 63u) == 0) {
      const uint32_t jl = j + lane;
      if (jl < m) {
        const uint64_t kl = lst[jl];
        mk.load(in_keys + kl * KS, KS);
        mvi = static_cast<uint32_t>(dd_value_index(t, kl));
      }
    }
    const uint64_t kb = lst[j];
    Blob<KS> ck;  // carried key, identical in every lane
#pragma unroll
    for (int w = 0; w < KW; ++w)
      ck.w[w] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mk.w[w]), static_cast<int>(j & 63u)));
    const uint32_t cvs =
        kArcBatch | static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mvi), static_cast<int>(j & 63u)));
    Blob<KS> ck0 = ck;  // the pair carried into the current window
    uint32_t cvs0 = cvs;
    // The walk as prefix minima.  Along a window, a carried pair that
    // entered at slot e with position p_e has position p_e + (k - e) at slot
    // k, and evicts the resident there iff that exceeds the resident's dist
    // d_k, i.e. iff d_k - k < p_e - e.  The evicted resident is carried on
    // with position d_k + 1 at slot k + 1, i.e. the same quantity d_k - k.
    // So with a_k = d_k - k per occupied slot and a = p for the pair carried
    // into the window (slot 0), the eviction points are the slots where a_k
    // is a strict new minimum of the a's before it, the pair arriving at slot
    // k is the one holding that running minimum, and its position there is
    // the minimum + k.  An equal key ends the walk only where a_k equals the
    // running minimum (position == dist); a free slot ends it; so does a
    // position past cap - 1 (the carried pair is dropped, rayon_cpu.rs:
    // 167-169).  One wave scan per window finds all of it.
    int64_t p = 0;
    uint32_t pos = static_cast<uint32_t>(((ck.hash(KS, t.mode) & t.mask) - r0) & t.mask);
    uint32_t c1 = 0;
    for (;;) {
      const uint32_t q = pos + lane;
      const bool in = q < L;
      const uint32_t f = in ? fl[q] : 0u;
      const bool occ = (f & 0xFu) == kOccupied;
      Blob<KS> rk;
#pragma unroll
      for (int w = 0; w < KW; ++w) rk.w[w] = occ ? ky[q * KW + w] : 0u;
      const uint32_t rvs = in ? vs[q] : 0u;
      const int64_t ak = (in && occ) ? static_cast<int64_t>(f >> 4) - static_cast<int64_t>(lane) : INT64_MAX;
      // inclusive running minimum (value, lane), the earlier lane kept on ties
      int64_t mv = ak;
      int mi = static_cast<int>(lane);
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int64_t v2 = __shfl_up(mv, off, 64);
        const int i2 = __shfl_up(mi, off, 64);
        if (lane >= static_cast<uint32_t>(off) && v2 <= mv) {
          mv = v2;
          mi = i2;
        }
      }
      // exclusive, with the carried-in pair (lane -1) first
      int64_t ev = __shfl_up(mv, 1, 64);
      int ei = __shfl_up(mi, 1, 64);
      if (lane == 0 || !(ev < p)) {
        ev = p;
        ei = -1;
      }
      // the pair at slot `lane` and its key
      Blob<KS> hk;
#pragma unroll
      for (int w = 0; w < KW; ++w) {
        const uint32_t x = static_cast<uint32_t>(__shfl(static_cast<int>(rk.w[w]), ei < 0 ? 0 : ei, 64));
        hk.w[w] = ei < 0 ? ck0.w[w] : x;
      }
      const uint32_t hvs_l = static_cast<uint32_t>(__shfl(static_cast<int>(rvs), ei < 0 ? 0 : ei, 64));
      const uint32_t hvs = ei < 0 ? cvs0 : hvs_l;
      const int64_t ppos = ev + static_cast<int64_t>(lane);  // position of that pair here
      bool eq = occ;
#pragma unroll
      for (int w = 0; w < KW; ++w) eq = eq && rk.w[w] == hk.w[w];
      const bool drop = ppos > static_cast<int64_t>(t.mask);
      const bool stop = drop || !in || !occ || (ak == ev && eq);
      const unsigned long long stops = __ballot(stop);
      const uint32_t E = stops ? static_cast<uint32_t>(__ffsll(stops) - 1) : 64u;
      // evictions before the end, and the end slot itself
      if (lane < E && in && occ && ak < ev) {
        vs[q] = hvs;
        fl[q] = make_flag(kOccupied, static_cast<uint64_t>(ppos));
#pragma unroll
        for (int w = 0; w < KW; ++w) ky[q * KW + w] = hk.w[w];
      }
      bool end = false;
      if (E < 64) {
        end = true;
        const uint32_t eflag = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(f), E));
        const bool edrop = (__builtin_amdgcn_readlane(static_cast<int>(drop), E)) != 0;
        if (edrop) {
          // carried pair dropped: nothing written here
        } else if (pos + E >= L) {
          if (lane == 0) atomicOr(&ctr->status, kStatusBudget);
        } else if ((eflag & 0xFu) != kOccupied) {  // EMPTY or TOMBSTONE: claim
          if (lane == E) {
            vs[q] = hvs;
            fl[q] = make_flag(kOccupied, static_cast<uint64_t>(ppos));
#pragma unroll
            for (int w = 0; w < KW; ++w) ky[q * KW + w] = hk.w[w];
          }
          c1 = 1;
        } else if (lane == E) {  // equal key: update its value
          vs[q] = hvs;
        }
      } else {  // the pair holding the window's minimum is carried into the next window
        int64_t tv = __shfl(mv, 63, 64);
        int ti = __shfl(mi, 63, 64);
        if (!(tv < p)) {
          tv = p;
          ti = -1;
        }
        if (ti >= 0) {
#pragma unroll
          for (int w = 0; w < KW; ++w) ck0.w[w] = static_cast<uint32_t>(__shfl(static_cast<int>(rk.w[w]), ti, 64));
          cvs0 = static_cast<uint32_t>(__shfl(static_cast<int>(rvs), ti, 64));
        }
        p = tv + 64;
      }
      wave_sync();
      if (end) break;
      pos += 64;
    }
    claimed += c1;
    if (lane == 0 && t.outcome) t.outcome[kb] = static_cast<uint8_t>(c1);
  }
  // Write-back, from the end of the arc down (see above).
  for (int64_t base = static_cast<int64_t>((L - 1) & ~63u); base >= 0; base -= 64) {
    const uint32_t q = static_cast<uint32_t>(base) + lane;
    const bool in = q < L;
    const uint32_t src = in ? vs[q] : q;
    const bool moved = in && src != q;
    Blob<VS> v;
    if (moved) {
      if (src & kArcBatch) v.load(in_vals + static_cast<uint64_t>(src & ~kArcBatch) * VS, VS);
      else v.load(t.vals + ((r0 + src) & t.mask) * VS, VS);
    }
    wave_sync();  // every lane's source read before any write of this chunk
    if (in) {
      const uint64_t slot = (r0 + q) & t.mask;
      t.flags[slot] = fl[q];
      Blob<KS> k;
#pragma unroll
      for (int w = 0; w < KW; ++w) k.w[w] = ky[q * KW + w];
      k.store(t.keys + slot * KS, KS);
      if (moved) v.store(t.vals + slot * VS, VS);
    }
  }
  return claimed;
}

// ---------------------------------------------------------------------------
// Canonical arcs: an arc's whole batch replay computed from its groups of
// equal homes.  An arc is canonical when its pre-batch residents lie in
// packed runs sorted by home: every OCCUPIED slot has dist 0, or follows an
// OCCUPIED slot whose dist is at least its own minus 1 (insert-only
// histories leave every arc this way; a tombstone inside a run does not).
// In a canonical run a sequential insert of key y with home h
// (rayon_cpu.rs:157-262) walks past the residents of homes <= h (comparing
// keys where the home is h), takes the place of the first resident of a
// larger home, and every later group of equal homes up to the run's end
// moves one slot on with its first member carried to its back (the carried
// pair passes the rest of its group, whose dists equal its position, and
// evicts the first resident of the next larger home); the walk ends at F_y,
// the first free slot at/after h, and the run stays canonical.  So after
// the batch:
//  * the arc holds its residents and the first copy of every new key,
//    sorted by home, at the positions parking gives them
//    (pos = max(home, previous pos + 1));
//  * a later copy of a key (of a resident or of an earlier batch key) only
//    replaces that element's value: the last copy's value stays;
//  * a group of equal homes h is its residents in slot order, then its new
//    keys appended in batch order, rotated by one (front to back) for every
//    insert y with home_y < h < F_y made after it gained its first element
//    (the F_y are a parking of the first copies in batch order).
// The wave computes F_y for every first copy in batch order (the arc's
// occupancy in registers; rotations counted per home as range additions:
// per slot inside a 64-slot block, per block across blocks), then each
// group's final order from its rotation counts, and writes every changed
// slot from the end of the arc down (every element moves towards the end,
// so a slot's old pair is read before it is overwritten).  The table is not
// written before the replay is known to be consistent; otherwise the
// function returns kNotCanon with only owner[] scratch written, and the
// caller replays the walks.  Called by all 64 lanes of a single-wave
// workgroup with lst[0..m) the members' batch indices in batch order.
// LDS after lst: rec[m], fp[max(m, 128)], hb[L], hg[L], srt[m] (16-bit).
// ---------------------------------------------------------------------------
constexpr uint32_t kNotCanon = 0xFFFFFFFFu;
constexpr uint32_t kCanonMaxL = 8192;         // the arc's occupancy: two 64-bit words per lane
constexpr uint32_t kRecLater = 0x80000000u;   // rec: a later copy of a key
constexpr uint32_t kRecFlag = 0x8000u;        // rec: first copy: its group needs its rotation count;
                                              //      later copy: its target is a resident
__host__ __device__ constexpr uint32_t canon_words(uint32_t m, uint32_t L) {
  return m + m + (m > 128 ? m : 128) + L + (L + 1) + (m + 1) / 2 + (L + 1) / 2;
}

__device__ __forceinline__ uint32_t rd_lane(uint32_t v, uint32_t l) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(l)));
}
__device__ __forceinline__ unsigned long long rd_lane64(unsigned long long v, uint32_t l) {
  return static_cast<unsigned long long>(rd_lane(static_cast<uint32_t>(v), l)) |
         (static_cast<unsigned long long>(rd_lane(static_cast<uint32_t>(v >> 32), l)) << 32);
}
__device__ __forceinline__ unsigned long long rd_any64(unsigned long long v, uint32_t src) {
  return static_cast<unsigned long long>(static_cast<uint32_t>(__shfl(static_cast<int>(v), static_cast<int>(src), 64))) |
         (static_cast<unsigned long long>(static_cast<uint32_t>(__shfl(static_cast<int>(v >> 32), static_cast<int>(src), 64)))
          << 32);
}
__device__ __forceinline__ uint32_t wave_incl_add(uint32_t v) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(v), off, 64));
    if (lane >= static_cast<uint32_t>(off)) v += y;
  }
  return v;
}
__device__ __forceinline__ int wave_incl_max(int v) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(v, off, 64);
    if (lane >= static_cast<uint32_t>(off)) v = max(v, y);
  }
  return v;
}

// Position of the first of `n` elements of home h (lane h of a 64-home
// chunk) when the elements of every home are parked in home order:
// elements before home h number s (exclusive sum), and the first one sits
// at s + max over nonempty homes h' <= h of (h' - s(h')).  `sum` and `mx`
// carry the chunk scans over.
__device__ __forceinline__ uint32_t park_first(uint32_t h, uint32_t n, uint32_t& sum, int& mx) {
  const uint32_t incl = wave_incl_add(n);
  const uint32_t s = sum + incl - n;
  const int m = max(wave_incl_max(n ? static_cast<int>(h) - static_cast<int>(s) : INT_MIN), mx);
  sum += rd_lane(incl, 63);
  mx = rd_lane(static_cast<uint32_t>(m), 63);
  return n ? s + static_cast<uint32_t>(m) : 0u;
}

template <int KS>
__device__ __forceinline__ uint32_t key_fp(const Blob<KS>& k) {
  uint32_t f = 0x9E3779B9u;
#pragma unroll
  for (int w = 0; w < KS / 4; ++w) {
    f = (f ^ k.w[w]) * 0x85EBCA6Bu;
    f ^= f >> 13;
  }
  return f;
}

// Rotations of a deque of `size` elements by r (front to back): index i -> (i - r) mod size.
__device__ __forceinline__ uint32_t rot_index(uint32_t i, uint32_t r, uint32_t size) {
  const uint32_t k = r % size;
  return i >= k ? i - k : i + size - k;
}

template <int KS, int VS>
__device__ uint32_t replay_arc_canon(const Table& t, const uint8_t* __restrict__ in_keys,
                                     const uint8_t* __restrict__ in_vals, uint64_t r0, uint32_t L,
                                     const uint32_t* lst, uint32_t m_in, uint32_t* pool) {
  static_assert(KS % 4 == 0 && VS % 4 == 0 && KS > 0 && VS > 0, "static word sizes");
  const uint32_t lane = threadIdx.x & 63u;
  // (wave-uniform: held in scalar registers, so the loops below are uniform)
  const uint32_t m = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(m_in)));
  L = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(L)));
  r0 = static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(r0))) |
       (static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(r0 >> 32))) << 32);
  uint32_t* rec = pool + m;
  uint32_t* fp = rec + m;
  uint32_t* hb = fp + (m > 128 ? m : 128);  // residents (high 16) | members' cursor (low 16)
  uint32_t* hg = hb + L;                    // Fenwick tree over homes 1..L: walks (low 16) | ends (high 16)
  uint16_t* srt = reinterpret_cast<uint16_t*>(hg + L + 1);
  uint16_t* fcn = srt + ((m + 1) & ~1u);    // per home: elements, then the first one's final position
  uint16_t* rec16 = reinterpret_cast<uint16_t*>(rec);
  const uint32_t nch = (L + 63) >> 6;
  for (uint32_t h = lane; h <= L; h += 64) {
    if (h < L) hb[h] = 0;
    hg[h] = 0;
  }
  wave_sync();
  // 1. The arc's flags: canonical check, residents per home, occupancy.
  unsigned long long occ0 = ~0ull, occ1 = ~0ull;  // lane l: words l and l + 64 (bit set: occupied or past the arc)
  bool bad = false;
  uint32_t prevf = 0;  // the flag before the chunk (the barrier before the arc: free)
  for (uint32_t c = 0; c < nch; ++c) {
    const uint32_t q = (c << 6) + lane;
    const bool in = q < L;
    const uint32_t f = in ? t.flags[(r0 + q) & t.mask] : 0u;
    uint32_t pf = static_cast<uint32_t>(__shfl_up(static_cast<int>(f), 1, 64));
    if (lane == 0) pf = prevf;
    const bool occ = in && (f & 0xFu) == kOccupied;
    const uint32_t d = f >> 4;
    if (occ) {
      const bool pocc = (pf & 0xFu) == kOccupied;
      if (d > q || (d > 0 && !(pocc && (pf >> 4) + 1 >= d))) bad = true;
      else atomicAdd(&hb[q - d], 0x10000u);
    }
    const unsigned long long word = __ballot(occ || !in);
    if (c < 64) {
      if (lane == c) occ0 = word;
    } else if (lane == c - 64) {
      occ1 = word;
    }
    prevf = rd_lane(f, 63);
  }
  if (__ballot(bad)) return kNotCanon;
  // 2. Members: homes (relative to the arc), fingerprints, counts per home.
  for (uint32_t j = lane; j < m; j += 64) {
    Blob<KS> k;
    k.load(in_keys + static_cast<uint64_t>(lst[j]) * KS, KS);
    const uint64_t h = ((k.hash(KS, t.mode) & t.mask) - r0) & t.mask;
    if (h >= L) {
      bad = true;
    } else {
      rec[j] = static_cast<uint32_t>(h);
      fp[j] = key_fp(k);
      atomicAdd(&hb[h], 1u);
    }
  }
  if (__ballot(bad)) return kNotCanon;
  wave_sync();
  // 3. Members by home (srt: batch ranks); after it hb's low half is the
  //    end of each home's range.
  {
    uint32_t sum = 0;
    for (uint32_t c = 0; c < nch; ++c) {
      const uint32_t h = (c << 6) + lane;
      const uint32_t w = h < L ? hb[h] : 0u;
      const uint32_t n = w & 0xFFFFu;
      const uint32_t incl = wave_incl_add(n);
      if (h < L) hb[h] = (w & 0xFFFF0000u) | (sum + incl - n);
      sum += rd_lane(incl, 63);
    }
  }
  wave_sync();
  for (uint32_t j = lane; j < m; j += 64)  // (a home's range in any order; sorted below)
    srt[atomicAdd(&hb[rec[j] & 0x7FFFu], 1u) & 0xFFFFu] = static_cast<uint16_t>(j);
  wave_sync();
  // 4. Later copies.  (a) One lane per home: the home's members put in
  //    batch order, and the first slot of its residents (parking of the
  //    residents by home) kept in fcn.  (b) One lane per member: its key
  //    against the group's residents, then against the group's earlier
  //    members (the first equal one is the key's first copy).  (c) One lane
  //    per home: the group's element count in fcn; first copies of groups
  //    of two or more elements need their rotation counts.
  {
    uint32_t rsum = 0;
    int rmx = INT_MIN;
    for (uint32_t c = 0; c < nch; ++c) {
      const uint32_t h = (c << 6) + lane;
      const bool in = h < L;
      const uint32_t w = in ? hb[h] : 0u;
      const uint32_t rf = park_first(h, w >> 16, rsum, rmx);
      const uint32_t mend = w & 0xFFFFu, mbeg = (in && h > 0) ? (hb[h - 1] & 0xFFFFu) : 0u;
      for (uint32_t i = mbeg + 1; i < mend; ++i) {
        const uint16_t v = srt[i];
        uint32_t k = i;
        while (k > mbeg && srt[k - 1] > v) {
          srt[k] = srt[k - 1];
          --k;
        }
        srt[k] = v;
      }
      if (in) fcn[h] = static_cast<uint16_t>(rf);
    }
  }
  wave_sync();
  for (uint32_t y = lane; y < m; y += 64) {
    const uint32_t j = srt[y];
    const uint32_t h = rec[j] & 0x7FFFu;
    const uint32_t w = hb[h];
    const uint32_t rc = w >> 16, mbeg = h > 0 ? (hb[h - 1] & 0xFFFFu) : 0u;
    const uint32_t fj = fp[j];
    uint32_t tgt = kNoOwner;
    bool tres = false;
    if (rc > 0) {
      const uint32_t rf = fcn[h];
      Blob<KS> kj;
      kj.load(in_keys + static_cast<uint64_t>(lst[j]) * KS, KS);
      for (uint32_t i = 0; i < rc && tgt == kNoOwner; ++i)
        if (kj.eq(t.keys + ((r0 + rf + i) & t.mask) * KS, KS)) {
          tgt = rf + i;
          tres = true;
        }
    }
    for (uint32_t y2 = mbeg; y2 < y && tgt == kNoOwner; ++y2) {
      const uint32_t j2 = srt[y2];
      if (fp[j2] != fj) continue;
      Blob<KS> kj;
      kj.load(in_keys + static_cast<uint64_t>(lst[j]) * KS, KS);
      if (kj.eq(in_keys + static_cast<uint64_t>(lst[j2]) * KS, KS)) tgt = j2;
    }
    if (tgt != kNoOwner) rec[j] |= kRecLater | (tres ? kRecFlag : 0u) | (tgt << 16);
  }
  wave_sync();
  for (uint32_t h = lane; h < L; h += 64) {
    const uint32_t w = hb[h];
    const uint32_t mend = w & 0xFFFFu, mbeg = h > 0 ? (hb[h - 1] & 0xFFFFu) : 0u;
    uint32_t firsts = 0;
    for (uint32_t y = mbeg; y < mend; ++y) firsts += (rec[srt[y]] & kRecLater) ? 0u : 1u;
    const uint32_t n = (w >> 16) + firsts;
    fcn[h] = static_cast<uint16_t>(n);
    if (n >= 2)
      for (uint32_t y = mbeg; y < mend; ++y) {
        const uint32_t j = srt[y];
        if (!(rec[j] & kRecLater)) rec[j] |= kRecFlag;
      }
  }
  wave_sync();
  // 5. F of every first copy: its walk's end, a parking of the first
  //    copies in batch order (one at a time; the arc's occupancy in
  //    registers), kept in rec's high half.
  bool lost = false;  // (wave-uniform)
  for (uint32_t base = 0; base < m; base += 64) {
    const uint32_t jl = base + lane;
    const uint32_t recl = jl < m ? rec[jl] : kRecLater;
    const uint32_t cnt = min(64u, m - base);
    uint32_t Fv = 0;
    uint32_t r = rd_lane(recl, 0);  // (the chunk's records are in registers before the steps)
    for (uint32_t u = 0; u < cnt; ++u) {
      const uint32_t rn = rd_lane(recl, u < 63 ? u + 1 : 63u);
      if (!(r & kRecLater)) {
        const uint32_t h = r & 0x7FFFu;
        const uint32_t w = h >> 6;
        unsigned long long word = w < 64 ? rd_lane64(occ0, w) : rd_lane64(occ1, w - 64);
        const unsigned long long fr = ~word & (~0ull << (h & 63));
        uint32_t F = (w << 6) + static_cast<uint32_t>(__ffsll(fr) - 1);
        uint32_t fw = w;
        if (!fr) {  // past the home's word: the first word after it with a free slot
          const unsigned long long f0 = lane > w ? ~occ0 : 0ull;
          const unsigned long long f1 = lane + 64 > w ? ~occ1 : 0ull;
          const unsigned long long b0 = __ballot(f0 != 0);
          const unsigned long long b1 = __ballot(f1 != 0);
          if (!b0 && !b1) {
            lost = true;
            break;
          }
          const uint32_t l = static_cast<uint32_t>(__ffsll(b0 ? b0 : b1) - 1);
          fw = l + (b0 ? 0u : 64u);
          word = fw < 64 ? rd_lane64(occ0, fw) : rd_lane64(occ1, fw - 64);
          F = (fw << 6) + static_cast<uint32_t>(__ffsll(~word) - 1);
        }
        word |= 1ull << (F & 63);
        const bool mine = lane == (fw & 63u);
        if (fw < 64) occ0 = mine ? word : occ0;
        else occ1 = mine ? word : occ1;
        Fv = lane == u ? F : Fv;
      }
      r = rn;
    }
    if (lost) break;
    if (jl < m && !(recl & kRecLater)) rec16[2 * jl + 1] = static_cast<uint16_t>(Fv);
  }
  if (lost) return kNotCanon;
  wave_sync();
  // 5b. Rotation counts, 64 first copies at a time in batch order: a walk y
  //     rotates the homes h with home_y < h < F_y.  A Fenwick tree over the
  //     homes counts, for the walks of earlier chunks, those with
  //     home_y < h (low half, added at home_y + 1) and those with F_y <= h
  //     (high half, added at F_y); their difference at h is the walks over
  //     h.  Walks earlier in the same chunk are counted from the lanes.  A
  //     first copy of a group of two or more elements keeps its count (the
  //     rotations its group had before it) in rec's high half.
  for (uint32_t base = 0; base < m; base += 64) {
    const uint32_t x = base + lane;
    const uint32_t r = x < m ? rec[x] : kRecLater;
    const bool first = !(r & kRecLater);
    const uint32_t h = r & 0x7FFFu, F = (r >> 16) & 0x7FFFu;
    const bool ask = first && (r & kRecFlag);
    uint32_t c = 0;
    if (ask) {
      uint32_t lo_sum = 0, hi_sum = 0;
      for (uint32_t i = h; i > 0; i -= i & (0u - i)) {
        const uint32_t v = hg[i];
        lo_sum += v & 0xFFFFu;
        hi_sum += v >> 16;
      }
      c = lo_sum - hi_sum;
    }
    for (uint32_t o = 0; o < 63; ++o) {
      const uint32_t ro = rd_lane(r, o);
      c += (o < lane && !(ro & kRecLater) && (ro & 0x7FFFu) < h && h < ((ro >> 16) & 0x7FFFu)) ? 1u : 0u;
    }
    wave_sync();
    if (first && F > h) {
      for (uint32_t i = h + 1; i <= L; i += i & (0u - i)) atomicAdd(&hg[i], 1u);
      for (uint32_t i = F; i <= L; i += i & (0u - i)) atomicAdd(&hg[i], 0x10000u);
    }
    if (ask) rec16[2 * x + 1] = static_cast<uint16_t>(c);
    wave_sync();
  }
  // 6. Each home's first final position (parking of the final groups).
  {
    uint32_t sum = 0;
    int mx = INT_MIN;
    for (uint32_t c = 0; c < nch; ++c) {
      const uint32_t h = (c << 6) + lane;
      const uint32_t f = h < L ? fcn[h] : 0u;
      const uint32_t p = park_first(h, f, sum, mx);
      if (h < L) fcn[h] = static_cast<uint16_t>(p);
    }
  }
  for (uint32_t q = lane; q < L; q += 64) t.owner[(r0 + q) & t.mask] = kNoOwner;
  __threadfence_block();
  wave_sync();
  // 7. Each group's final order: a record per element at its final slot in
  //    owner[]: bit 31 set: key and value of batch rank (bits 0-14, the
  //    element's last copy), else the resident at slot (bits 0-14)
  //    unchanged; bits 16-30 the home.  An element's index in its group
  //    follows the rotations after it joined (each first copy of the group
  //    brings a count; cfin is the count after the batch).  First copies
  //    one lane each, residents one lane per home.
  const auto walks_over = [&](uint32_t h) -> uint32_t {  // the walks over h of the whole batch
    uint32_t lo_sum = 0, hi_sum = 0;
    for (uint32_t i = h; i > 0; i -= i & (0u - i)) {
      const uint32_t v = hg[i];
      lo_sum += v & 0xFFFFu;
      hi_sum += v >> 16;
    }
    return lo_sum - hi_sum;
  };
  uint32_t claimed = 0;
  for (uint32_t y = lane; y < m; y += 64) {
    const uint32_t j = srt[y];
    const uint32_t r = rec[j];
    if (r & kRecLater) continue;
    ++claimed;
    const uint32_t h = r & 0x7FFFu;
    const uint32_t w = hb[h];
    const uint32_t rc = w >> 16, mend = w & 0xFFFFu, mbeg = h > 0 ? (hb[h - 1] & 0xFFFFu) : 0u;
    uint32_t e = rc;
    for (uint32_t y2 = mbeg; y2 < y; ++y2) e += (rec[srt[y2]] & kRecLater) ? 0u : 1u;
    uint32_t idx = e, size = e + 1, cprev = (r >> 16) & 0x7FFFu, vs = j;
    for (uint32_t y2 = y + 1; y2 < mend; ++y2) {
      const uint32_t j2 = srt[y2];
      const uint32_t r2 = rec[j2];
      if (r2 & kRecLater) {
        if (!(r2 & kRecFlag) && ((r2 >> 16) & 0x7FFFu) == j) vs = j2;
      } else {
        const uint32_t c2 = (r2 >> 16) & 0x7FFFu;
        idx = rot_index(idx, c2 - cprev, size);
        cprev = c2;
        ++size;
      }
    }
    if (size > 1) idx = rot_index(idx, walks_over(h) - cprev, size);
    const uint32_t pos = fcn[h] + idx;
    if (pos >= L) bad = true;
    else t.owner[(r0 + pos) & t.mask] = 0x80000000u | vs | (h << 16);
  }
  {
    uint32_t rsum = 0;
    int rmx = INT_MIN;
    for (uint32_t c = 0; c < nch; ++c) {
      const uint32_t h = (c << 6) + lane;
      const bool in = h < L;
      const uint32_t w = in ? hb[h] : 0u;
      const uint32_t rc = w >> 16;
      const uint32_t rf = park_first(h, rc, rsum, rmx);
      if (rc == 0) continue;
      const uint32_t mend = w & 0xFFFFu, mbeg = h > 0 ? (hb[h - 1] & 0xFFFFu) : 0u;
      const uint32_t cfin = walks_over(h);
      for (uint32_t e = 0; e < rc; ++e) {
        uint32_t idx = e, size = rc, cprev = 0, vs = kNoOwner;
        for (uint32_t y = mbeg; y < mend; ++y) {
          const uint32_t j = srt[y];
          const uint32_t r = rec[j];
          if (r & kRecLater) {
            if ((r & kRecFlag) && ((r >> 16) & 0x7FFFu) == rf + e) vs = j;
          } else {
            const uint32_t cj = (r >> 16) & 0x7FFFu;
            idx = rot_index(idx, cj - cprev, size);
            cprev = cj;
            ++size;
          }
        }
        idx = rot_index(idx, cfin - cprev, size);
        const uint32_t pos = fcn[h] + idx;
        if (pos >= L) {
          bad = true;
          break;
        }
        t.owner[(r0 + pos) & t.mask] = (vs != kNoOwner ? (0x80000000u | vs) : (rf + e)) | (h << 16);
      }
    }
  }
  if (__ballot(bad)) return kNotCanon;
  __threadfence_block();  // the records, read back by other lanes
  wave_sync();
  // 8. Write-back from the end of the arc down, 64 slots at a time.
  for (int64_t base = static_cast<int64_t>((L - 1) & ~63u); base >= 0; base -= 64) {
    const uint32_t q = static_cast<uint32_t>(base) + lane;
    const uint64_t slot = (r0 + q) & t.mask;
    const uint32_t rd = q < L ? t.owner[slot] : kNoOwner;
    const uint32_t id = rd & 0x7FFFu;
    const bool batch = (rd & 0x80000000u) != 0;
    const bool put = rd != kNoOwner && (batch || id != q);
    Blob<KS> k;
    Blob<VS> v;
    if (put) {
      if (batch) {
        const uint64_t bi = lst[id];
        k.load(in_keys + bi * KS, KS);
        v.load(in_vals + dd_value_index(t, bi) * VS, VS);
      } else {
        const uint64_t os = (r0 + id) & t.mask;
        k.load(t.keys + os * KS, KS);
        v.load(t.vals + os * VS, VS);
      }
    }
    wave_sync();  // every lane's source read before any write of this chunk
    if (put) {
      t.flags[slot] = make_flag(kOccupied, q - ((rd >> 16) & 0x7FFFu));
      k.store(t.keys + slot * KS, KS);
      v.store(t.vals + slot * VS, VS);
    }
  }
  if (t.outcome)
    for (uint32_t j = lane; j < m; j += 64) t.outcome[lst[j]] = (rec[j] & kRecLater) ? 0 : 1;
  // (the sum over lanes)
  uint32_t tot = claimed;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tot += static_cast<uint32_t>(__shfl_xor(static_cast<int>(tot), off, 64));
  return tot;
}

// A one-member closure's insert decided from its flags.  Key i alone in its
// arc walks home .. s at most (s: its dry slot; every slot from home up to
// s was OCCUPIED before the batch, or the dry pass would have claimed it),
// and insert_walk's decisions on those slots (pass, compare an equal-dist
// resident's key, evict) depend on the flags alone except for the key
// comparisons.  So the flags of home .. s-1 (fl[j] = flag of slot home+j,
// j < L = s - home, read with the election's bitmap words) give every
// comparison point and eviction point of the walk as if no key were equal;
// one round of reads brings the keys at those points and the residents to
// be moved, the comparisons are taken in walk order (an equal key ends the
// walk with a value update there, the pairs evicted before it moved), and
// one round of writes applies the keys and flags.  The value moves (at most
// kPlanEvicts + 1 copies of V bytes) are returned in `mv` as (source,
// destination) byte addresses in reverse walk order, for the caller to
// apply: copying them in that order, each batch of copies reading all its
// sources before writing, reads every value before it is overwritten.
// vin: the batch value of key i.  Returns false, with nothing written, when
// the walk needs more comparisons or evictions than it keeps in registers;
// the caller then walks.  *c1: 1 if s was claimed.
constexpr uint32_t kPlanSlots = 16;
constexpr uint32_t kPlanChecks = 3;
constexpr uint32_t kPlanEvicts = 2;
constexpr uint32_t kPlanMoves = kPlanEvicts + 1;

template <int KS, int VS>
__device__ inline bool insert_planned(const Table& t, const Blob<KS>& ck, const uint8_t* vin, uint64_t home,
                                      uint64_t s, uint32_t L, const uint32_t (&fl)[kPlanSlots], uint32_t* c1,
                                      uint64_t (&mv)[2 * kPlanMoves], uint32_t* nmv) {
  // A complete plan covers the walk up to s (at most kPlanChecks comparisons
  // and kPlanEvicts evictions, s among the plan's slots).  Otherwise the
  // walk can still end at an equal key among the first kPlanChecks
  // comparisons (an update, rayon_cpu.rs:176-181): the plan succeeds only
  // when it finds one there.
  const bool window = L > kPlanSlots;
  uint32_t p = 0, ne = 0, nc = 0, ep0 = 0, ep1 = 0, cmask = 0, emask = 0;
  bool ok = true;
#pragma unroll
  for (uint32_t j = 0; j < kPlanSlots; ++j) {
    if (j < L) {
      const uint32_t f = fl[j];
      const uint32_t d = f >> 4;
      if ((f & 0xFu) != kOccupied || p > t.mask) ok = false;  // (the walk would drop a pair past a lap)
      if (d == p) {  // an equal key would end the walk here
        cmask |= 1u << j;
        ++nc;
        ++p;
      } else if (p > d) {  // Robin Hood eviction: the carried pair settles, the resident is carried on
        emask |= 1u << j;
        if (ne == 0) ep0 = p;
        else if (ne == 1) ep1 = p;  // (an incomplete plan counts more evictions than it applies)
        ++ne;
        p = d + 1;
      } else {
        ++p;
      }
    }
  }
  const bool complete = !window && nc <= kPlanChecks && ne <= kPlanEvicts;
  if (!ok || p > t.mask || (!complete && nc == 0)) return false;
  const uint32_t K = KS, V = VS;
  auto slot = [&](uint32_t j) { return (home + j) & t.mask; };
  const uint32_t e0 = emask ? static_cast<uint32_t>(__ffs(emask) - 1) : 0u;
  const uint32_t e1 = ne > 1 ? static_cast<uint32_t>(__ffs(emask & (emask - 1)) - 1) : 0u;
  Blob<KS> ke0, ke1, kc0, kc1, kc2;
  uint32_t cj0 = 0, cj1 = 0, cj2 = 0;
  {
    uint32_t m = cmask;
    cj0 = m ? static_cast<uint32_t>(__ffs(m) - 1) : 0u;
    m &= m - 1;
    cj1 = m ? static_cast<uint32_t>(__ffs(m) - 1) : 0u;
    m &= m - 1;
    cj2 = m ? static_cast<uint32_t>(__ffs(m) - 1) : 0u;
  }
  if (ne > 0) ke0.load(t.keys + slot(e0) * K, K);
  if (ne > 1) ke1.load(t.keys + slot(e1) * K, K);
  if (nc > 0) kc0.load(t.keys + slot(cj0) * K, K);
  if (nc > 1) kc1.load(t.keys + slot(cj1) * K, K);
  if (nc > 2) kc2.load(t.keys + slot(cj2) * K, K);
  // comparisons in walk order; the pair carried at point j is the one that
  // the evictions before j left (0: the batch pair).  Every candidate pair
  // is compared and the result selected: selecting among the blobs
  // themselves would put them in scratch memory.
  const uint32_t c0 = static_cast<uint32_t>(__popc(emask & ((1u << cj0) - 1u)));
  const uint32_t c1_ = static_cast<uint32_t>(__popc(emask & ((1u << cj1) - 1u)));
  const uint32_t c2_ = static_cast<uint32_t>(__popc(emask & ((1u << cj2) - 1u)));
  const bool q0 = c0 == 0 ? ck.eqb(kc0) : (c0 == 1 ? ke0.eqb(kc0) : ke1.eqb(kc0));
  const bool q1 = c1_ == 0 ? ck.eqb(kc1) : (c1_ == 1 ? ke0.eqb(kc1) : ke1.eqb(kc1));
  const bool q2 = c2_ == 0 ? ck.eqb(kc2) : (c2_ == 1 ? ke0.eqb(kc2) : ke1.eqb(kc2));
  // (a comparison after more than two evictions would compare a resident
  // the plan did not load: no stop there or after it)
  uint32_t stop_j = ~0u;
  if (nc > 0 && c0 <= 2 && q0) stop_j = cj0;
  else if (nc > 1 && c1_ <= 2 && q1) stop_j = cj1;
  else if (nc > 2 && c2_ <= 2 && q2) stop_j = cj2;
  if (!complete && stop_j == ~0u) return false;
  const uint32_t applied = stop_j != ~0u ? static_cast<uint32_t>(__popc(emask & ((1u << stop_j) - 1u))) : ne;
  const uint64_t vb = reinterpret_cast<uint64_t>(t.vals);
  const uint64_t src_e0 = vb + slot(e0) * V, src_e1 = vb + slot(e1) * V;
  if (applied >= 1) {
    ck.store(t.keys + slot(e0) * K, K);
    t.flags[slot(e0)] = make_flag(kOccupied, ep0);
  }
  if (applied >= 2) {
    ke0.store(t.keys + slot(e1) * K, K);
    t.flags[slot(e1)] = make_flag(kOccupied, ep1);
  }
  // the last pair carried: to the equal key's slot (value only) or to s
  mv[0] = applied == 0 ? reinterpret_cast<uint64_t>(vin) : (applied == 1 ? src_e0 : src_e1);
  if (stop_j != ~0u) {
    mv[1] = vb + slot(stop_j) * V;
    *c1 = 0;
  } else {
    if (ne == 0) ck.store(t.keys + s * K, K);
    else if (ne == 1) ke0.store(t.keys + s * K, K);
    else ke1.store(t.keys + s * K, K);
    t.flags[s] = make_flag(kOccupied, p);
    mv[1] = vb + s * V;
    *c1 = 1;
  }
  // then the resident of e0 -> e1 (two evictions applied), and the batch
  // pair -> e0 (fixed indices: a computed one would put mv in scratch)
  mv[2] = applied >= 2 ? src_e0 : reinterpret_cast<uint64_t>(vin);
  mv[3] = applied >= 2 ? src_e1 : src_e0;
  mv[4] = reinterpret_cast<uint64_t>(vin);
  mv[5] = src_e0;
  *nmv = 1 + applied;
  return true;
}

// Applies a wave's list of V-byte value copies (lst_src/lst_dst, `total`
// entries in LDS) with all 64 lanes: V/16 lanes per copy, 16 bytes each, so
// one copy is one request per line instead of V/16.  The list is taken in
// batches of 64 copies staged through `stage` (64 * V bytes of LDS): every
// source of a batch is read before any destination of it is written, so
// lists built in reverse walk order per walk (insert_planned) read every
// value before it is overwritten.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int VS>
__device__ inline void wave_copy_values(const uint64_t* lst_src, const uint64_t* lst_dst, uint32_t total,
                                        u32x4* stage) {
  static_assert(VS % 16 == 0 && VS >= 16 && VS <= 1024, "16-byte parts");
  constexpr uint32_t kParts = VS / 16, kPerRow = 64 / kParts, kRows = kParts;  // 64 copies per batch
  typedef __attribute__((address_space(1))) const u32x4 gsrc_t;  // global memory (not generic) accesses
  typedef __attribute__((address_space(1))) u32x4 gdst_t;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t part = lane % kParts, row_copy = lane / kParts;
  for (uint32_t c0 = 0; c0 < total; c0 += kPerRow * kRows) {
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r) {  // no branch: every load in flight at once (past the end: the last copy)
      const uint32_t c = min(c0 + r * kPerRow + row_copy, total - 1);
      stage[r * 64 + lane] = reinterpret_cast<gsrc_t*>(lst_src[c])[part];
    }
    wave_sync();
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r) {
      const uint32_t c = c0 + r * kPerRow + row_copy;
      if (c < total) reinterpret_cast<gdst_t*>(lst_dst[c])[part] = stage[r * 64 + lane];
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// Dense batches: arcs replayed in registers, one wave per bitmap word.  A
// closure's arc is the run of slots between two barriers (pre-batch free
// slots that no dry claim took), and its members walk only inside it (the
// exactness contract above).  The wave of word w takes every arc that
// starts in w's 64 slots and has at most kArcSlots slots: lane = slot, the
// arc's flags, keys and value sources in registers, and the members (the
// owners of its dry claims) inserted in batch order, each walk decided by
// one wave scan, the prefix-minimum form of replay_arc_lds over a single
// window (so the result is insert_walk's, slot for slot).  The arcs that
// end inside the word are replayed together in one round, the one that
// runs into the next word in a second.  Longer arcs are queued for
// k_ins_big by the dry slot of their first member, the entry k_ins_leader
// queues.  The arcs of a round replay side by side, one member each per
// step.  Every changed slot is written once at the end, its value read
// from its pre-batch slot or its batch position before any value is
// written.  Requires cap >= 256 (the window's three bitmap words differ).
// ---------------------------------------------------------------------------
constexpr uint32_t kArcSlots = 64;

// Bitonic sort of one 64-bit key per lane across the wave, ascending by lane.
__device__ __forceinline__ void wave_sort64(unsigned long long& sk, uint32_t lane) {
#pragma unroll
  for (uint32_t size = 2; size <= 64; size <<= 1) {
#pragma unroll
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      const unsigned long long other = __shfl_xor(sk, static_cast<int>(stride), 64);
      const bool keep_min = ((lane & stride) == 0) == ((lane & size) == 0);
      sk = keep_min ? (other < sk ? other : sk) : (other > sk ? other : sk);
    }
  }
}

constexpr uint32_t kVsBatch = 0x100u;  // value source: the batch value of the member in lane (vs & 63)

// One round: lanes are the slots base .. base+63, `act` marks the slots of
// the arcs replayed, `starts` their first slots and `mem` their dry-claimed
// slots.  Every arc replays its j-th member (in batch order) in step j, so
// a round takes as many steps as its largest arc has members.  Returns the
// slots claimed (the same in every lane).  kPacked: the arcs come from
// anywhere in the table, packed side by side into the 64 lanes, and lane l
// holds slot `lane_slot` (consecutive within an arc; `base` unused) and, for
// a member, its batch index `lane_owner` (owner[] is not read).
template <int KS, int VS, bool kPacked = false>
__device__ __forceinline__ uint32_t arcs_round(const Table& t, const uint8_t* __restrict__ in_keys,
                                               const uint8_t* __restrict__ in_vals, uint64_t base,
                                               unsigned long long act, unsigned long long starts,
                                               unsigned long long mem, Counters* ctr,
                                               unsigned long long& claims, uint64_t lane_slot = 0,
                                               uint32_t lane_owner = 0, uint32_t* __restrict__ mk_lds = nullptr) {
  constexpr int KW = KS / 4;
  const uint32_t lane = threadIdx.x & 63u;
  const bool in = (act >> lane) & 1ull;
  const bool is_mem = (mem >> lane) & 1ull;
  const uint64_t q = kPacked ? lane_slot : (base + lane) & t.mask;
  // this lane's arc: first lane a0, length alen, members mcount, and the
  // sorted position of its first member (members of earlier arcs sort first)
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const unsigned long long sm = starts & upto;
  const uint32_t a0 = sm ? 63u - static_cast<uint32_t>(__clzll(sm)) : 0u;
  uint32_t alen;
  if constexpr (kPacked) {  // arcs abut: this one ends where the next starts (or the packed lanes end)
    const unsigned long long nxt = a0 < 63 ? (starts & (~0ull << (a0 + 1))) : 0ull;
    alen = nxt ? static_cast<uint32_t>(__ffsll(nxt) - 1) - a0 : static_cast<uint32_t>(__popcll(act)) - a0;
  } else {
    const unsigned long long rest = ~(act >> a0);
    alen = rest ? static_cast<uint32_t>(__ffsll(rest) - 1) : 64u - a0;
  }
  const unsigned long long amask = (alen == 64 ? ~0ull : ((1ull << alen) - 1ull)) << a0;
  const uint32_t mcount = static_cast<uint32_t>(__popcll(mem & amask));
  const uint32_t mst = static_cast<uint32_t>(__popcll(mem & ((1ull << a0) - 1ull)));
  // fresh: every slot of the round's arcs was free before the batch (the
  // arcs hold only their members' dry claims), so no slot is read
  const bool fresh = !kPacked && act == mem;
  uint32_t f = 0, o = 0;
  Blob<KS> rk;  // the slot's key
#pragma unroll
  for (int w = 0; w < KW; ++w) rk.w[w] = 0;
  if (in) {
    if (!fresh) {
      f = t.flags[q];
      rk.load(t.keys + q * KS, KS);
    }
    if (is_mem) o = kPacked ? lane_owner : t.owner[q];
    keep_loaded<KS>(f, rk);
  }
  Blob<KS> mk;  // a member's key, in the lane of its dry slot
  uint32_t hp = 0, vi = 0;
  uint64_t mh = 0;  // its hash
  // (packed: the slot held by lane a0, the arc's first)
  const uint64_t org = kPacked ? (static_cast<uint64_t>(static_cast<uint32_t>(__shfl(static_cast<int>(q), a0, 64))) |
                                  (static_cast<uint64_t>(static_cast<uint32_t>(__shfl(static_cast<int>(q >> 32), a0, 64))) << 32))
                               : 0ull;
  if (is_mem) {
    mk.load(in_keys + static_cast<uint64_t>(o) * KS, KS);
    vi = static_cast<uint32_t>(dd_value_index(t, o));
    mh = mk.hash(KS, t.mode);
    hp = kPacked ? a0 + static_cast<uint32_t>(((mh & t.mask) - org) & t.mask)
                 : static_cast<uint32_t>(((mh & t.mask) - base) & t.mask);  // home lane
    if (hp > lane || hp < a0) {  // cannot happen: a member's home lies in its arc, at or before its dry slot
      atomicOr(&ctr->status, kStatusBudget);
      hp = lane;
    }
  }
  if constexpr (!kPacked) {  // members' keys in this wave's LDS (read back in order by the same wave)
#pragma unroll
    for (int w = 0; w < KW; ++w) mk_lds[w * 64 + lane] = mk.w[w];
  }
  if constexpr (!kPacked) {
    if (fresh) {
      // Insertion in batch order into free slots leaves an arc's members
      // sorted by home, one to a slot from the arc's first.  Within one
      // home's group the order follows the walks: member y parks at the
      // first slot at or after its home that no earlier member took (F_y),
      // joins its own home's group at the end (it passes the pairs of its
      // home, equal distances, and evicts the next group's first), and
      // rotates every group of a home h with home_y < h < F_y (that group's
      // first pair is carried past the others of its home to the group's
      // end).  Groups are replayed as queues of member lanes, 6 bits each,
      // in the lane of their first pair in (home, batch) order.  A later
      // copy of a key updates the first copy instead: equal hashes in one
      // home's group (or a group of more than 8) send the round to the walk
      // below.
      unsigned long long fk = is_mem ? (static_cast<unsigned long long>(hp) << 38) |
                                           (static_cast<unsigned long long>(o) << 6) | lane
                                     : (~0ull << 6) | lane;
      wave_sort64(fk, lane);
      const uint32_t nm = static_cast<uint32_t>(__popcll(mem));
      const uint32_t shp = static_cast<uint32_t>(fk >> 38) & 63u;
      const uint64_t smh = static_cast<uint64_t>(__shfl(static_cast<long long>(mh), static_cast<int>(fk & 63u), 64));
      bool clash = false;
#pragma unroll
      for (uint32_t d = 1; d <= 8; ++d) {
        const uint32_t php = static_cast<uint32_t>(__shfl_up(static_cast<int>(shp), d, 64));
        const uint64_t pmh = static_cast<uint64_t>(__shfl_up(static_cast<long long>(smh), d, 64));
        clash = clash || (lane < nm && lane >= d && php == shp && (d == 8 || pmh == smh));
      }
      if (!__ballot(clash)) {
        const uint32_t p1 = static_cast<uint32_t>(__shfl_up(static_cast<int>(shp), 1, 64));
        const unsigned long long leaders = __ballot(lane < nm && (lane == 0 || p1 != shp));
        const bool lead = (leaders >> lane) & 1ull;
        unsigned long long bk = is_mem ? (static_cast<unsigned long long>(o) << 6) | lane : ~0ull;
        wave_sort64(bk, lane);  // members in batch order
        unsigned long long gq = 0, taken = 0;
        uint32_t gn = 0;
        for (uint32_t j = 0; j < nm; ++j) {
          const uint32_t yl = rd_lane(static_cast<uint32_t>(bk), j) & 63u;
          const uint32_t hy = rd_lane(hp, yl);
          const unsigned long long fr = ~taken & (~0ull << hy);
          const uint32_t F = fr ? static_cast<uint32_t>(__ffsll(fr) - 1) : 64u;
          taken |= fr ? (1ull << F) : 0ull;
          if (lead) {
            if (shp == hy) {
              gq |= static_cast<unsigned long long>(yl) << (6u * gn);
              ++gn;
            } else if (hy < shp && shp < F && gn > 1) {
              gq = (gq >> 6) | ((gq & 63ull) << (6u * (gn - 1)));
            }
          }
        }
        if (taken != act) atomicOr(&ctr->status, kStatusBudget);  // cannot happen: parked outside the arcs
        // slot lane takes the pair at its rank among the arcs' slots: the
        // (rank - group start)-th of its group's queue
        const uint32_t r = static_cast<uint32_t>(__popcll(act & ((1ull << lane) - 1ull)));
        const unsigned long long below = leaders & ((2ull << r) - 1ull);  // r = 63: every lane
        const uint32_t g0 = below ? 63u - static_cast<uint32_t>(__clzll(below)) : 0u;
        const uint32_t gh = static_cast<uint32_t>(__shfl(static_cast<int>(shp), static_cast<int>(r), 64));
        const unsigned long long gqr = static_cast<unsigned long long>(__shfl(static_cast<long long>(gq), static_cast<int>(g0), 64));
        const int src = static_cast<int>((gqr >> (6u * ((r - g0) & 7u))) & 63ull);
        Blob<KS> ck;
#pragma unroll
        for (int w = 0; w < KW; ++w) {
      if constexpr (kPacked) ck.w[w] = static_cast<uint32_t>(__shfl(static_cast<int>(mk.w[w]), src, 64));
      else ck.w[w] = mk_lds[w * 64 + src];
    }
        const uint32_t vj = static_cast<uint32_t>(__shfl(static_cast<int>(vi), src, 64));
        const uint32_t oj = static_cast<uint32_t>(__shfl(static_cast<int>(o), src, 64));
        if (in) {
          Blob<VS> v;
          v.load(in_vals + static_cast<uint64_t>(vj) * VS, VS);
          t.flags[q] = make_flag(kOccupied, lane - gh);
          ck.store(t.keys + q * KS, KS);
          v.store(t.vals + q * VS, VS);
          if (t.outcome) t.outcome[oj] = 1;
        }
        claims = act;
        return nm;
      }
    }
  }
  bool occ = in && (f & 0xFu) == kOccupied;
  uint32_t dist = f >> 4;
  uint32_t vs = lane;  // value source: the pre-batch value of slot lane, or kVsBatch | member lane
  bool dirty = false;
  // members by (arc, batch index); other lanes last
  unsigned long long sk = is_mem ? (static_cast<unsigned long long>(a0) << 38) |
                                       (static_cast<unsigned long long>(o) << 6) | lane
                                 : (~0ull << 6) | lane;
  wave_sort64(sk, lane);
  // As in replay_arc_lds: with a_k = dist_k - (k - h) per occupied slot k
  // at or after the member's home h, the eviction points are the strict new
  // minima of the a's (the carried pair counting as 0 before slot h), the
  // pair arriving at k is the one holding the running minimum and its
  // position there is the minimum + (k - h).  The walk ends at the first free
  // slot, or at an equal key where a_k equals the running minimum.  Scan keys
  // are (63 - a0) << 21 | (a + 64) << 7 | (lane + 1): a lane never takes the
  // minimum of an earlier arc (larger high bits), the minimum keeps the
  // earlier lane on ties, and the carried pair (index 0) comes first.
  // Distances above 127 are clipped: such a resident's a is above every
  // minimum that matters.
  constexpr uint32_t kCarried = 64u << 7;
  const uint32_t hib = (63u - a0) << 21;
  uint32_t claimed = 0;
  claims = 0;  // the round's claims (free slots taken), lane l = slot base + l
  for (uint32_t j = 0;; ++j) {
    const bool live = in && j < mcount;  // this lane's arc inserts a member in this step
    if (!__ballot(live)) break;
    const int pos = static_cast<int>(live ? mst + j : 0u);
    const uint32_t slo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(sk)), pos, 64));
    const uint32_t shi = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(sk >> 32)), pos, 64));
    const int src = static_cast<int>(slo & 63u);
    const uint32_t oj = (slo >> 6) | (shi << 26);
    Blob<KS> ck;  // the member's key
#pragma unroll
    for (int w = 0; w < KW; ++w) {
      if constexpr (kPacked) ck.w[w] = static_cast<uint32_t>(__shfl(static_cast<int>(mk.w[w]), src, 64));
      else ck.w[w] = mk_lds[w * 64 + src];
    }
    const uint32_t h = static_cast<uint32_t>(__shfl(static_cast<int>(hp), src, 64));
    const bool inw = live && lane >= h;
    const int ak = static_cast<int>(dist < 127u ? dist : 127u) - static_cast<int>(lane - h);
    uint32_t pk = (inw && occ) ? hib | (static_cast<uint32_t>(ak + 64) << 7) | (lane + 1) : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(pk), off, 64));
      if (lane >= off && y < pk) pk = y;
    }
    uint32_t ex = static_cast<uint32_t>(__shfl_up(static_cast<int>(pk), 1, 64));
    if (lane <= h || ex > (hib | kCarried)) ex = hib | kCarried;
    const int ev = static_cast<int>((ex >> 7) & 255u) - 64;
    const int ei = static_cast<int>(ex & 127u) - 1;  // -1: the carried pair
    const int eis = ei < 0 ? 0 : ei;
    Blob<KS> hk;  // key of the pair arriving here
#pragma unroll
    for (int w = 0; w < KW; ++w) {
      const uint32_t x = static_cast<uint32_t>(__shfl(static_cast<int>(rk.w[w]), eis, 64));
      hk.w[w] = ei < 0 ? ck.w[w] : x;
    }
    const uint32_t hvs_l = static_cast<uint32_t>(__shfl(static_cast<int>(vs), eis, 64));
    const uint32_t hvs = ei < 0 ? (kVsBatch | static_cast<uint32_t>(src)) : hvs_l;
    const uint32_t ppos = static_cast<uint32_t>(ev + static_cast<int>(lane - h));
    bool eq = occ;
#pragma unroll
    for (int w = 0; w < KW; ++w) eq = eq && rk.w[w] == hk.w[w];
    const bool stop = inw && (!occ || (ak == ev && eq));
    const unsigned long long stops = __ballot(stop);
    const unsigned long long occs = __ballot(occ);
    const unsigned long long mine = live ? (stops & (~0ull << h)) : 0ull;
    const uint32_t E = mine ? static_cast<uint32_t>(__ffsll(mine) - 1) : 64u;  // this arc's walk ends here
    const bool eok = live && E < a0 + alen;
    if (live && !eok && lane == h) atomicOr(&ctr->status, kStatusBudget);  // cannot happen: left the arc
    if (eok && lane < E && inw && occ && ak < ev) {  // evictions on the way
      rk = hk;
      vs = hvs;
      dist = ppos;
      dirty = true;
    }
    const bool efree = eok && !((occs >> E) & 1ull);
    if (eok && lane == E) {
      if (efree) {  // claim the free slot
        occ = true;
        dist = ppos;
        rk = hk;
      }
      vs = hvs;  // claimed, or the equal key's value updated
      dirty = true;
      if (t.outcome) t.outcome[oj] = efree ? 1 : 0;
    }
    const unsigned long long took = __ballot(eok && lane == E && efree);
    claimed += static_cast<uint32_t>(__popcll(took));
    claims |= took;
  }
  // Values: every source read before any write.  Each store waits for the
  // load that fetched its data, and one load instruction covers every lane,
  // so no lane overwrites a slot another lane has still to read.
  const uint32_t vsrc = vs & 63u;
  const uint32_t vi_src = static_cast<uint32_t>(__shfl(static_cast<int>(vi), static_cast<int>(vsrc), 64));
  const bool vmove = in && dirty && vs != lane;
  uint64_t src_slot = (base + vsrc) & t.mask;
  if constexpr (kPacked)
    src_slot = static_cast<uint64_t>(static_cast<uint32_t>(__shfl(static_cast<int>(q), static_cast<int>(vsrc), 64))) |
               (static_cast<uint64_t>(static_cast<uint32_t>(__shfl(static_cast<int>(q >> 32), static_cast<int>(vsrc), 64)))
                << 32);
  Blob<VS> v;
  if (vmove) {
    if (vs & kVsBatch) v.load(in_vals + static_cast<uint64_t>(vi_src) * VS, VS);
    else v.load(t.vals + src_slot * VS, VS);
  }
  if (in && dirty) {
    t.flags[q] = make_flag(kOccupied, dist);
    rk.store(t.keys + q * KS, KS);
    if (vmove) v.store(t.vals + q * VS, VS);
  }
  return claimed;
}

// Clears claim bits in the next bitmaps: mask m of the 64 slots from slot
// b (b may start inside a word; cap is a multiple of 64).  The three next
// bitmaps are one block (nempty = nfree + words, nq = nfree + 2 * words), so
// only t.nfree is read: the dense kernel's SGPR budget decides whether two
// of its 1024-thread workgroups fit on a CU.
__device__ inline void claim_next_mask(const Table& t, uint64_t b, unsigned long long m) {
  if (!m) return;
  const uint64_t words = t.cap >> 6;
  const uint32_t sh = static_cast<uint32_t>(b & 63u);
  const uint64_t w0 = b >> 6, w1 = w0 + 1 == words ? 0 : w0 + 1;
  const unsigned long long lo = m << sh, hi = sh ? m >> (64u - sh) : 0ull;
  unsigned long long* nb = t.nfree;
  if (lo) {
    atomicAnd(&nb[w0], ~lo);
    atomicAnd(&nb[words + w0], ~lo);
    atomicAnd(&nb[2 * words + w0], ~lo);
  }
  if (hi) {
    atomicAnd(&nb[w1], ~hi);
    atomicAnd(&nb[words + w1], ~hi);
    atomicAnd(&nb[2 * words + w1], ~hi);
  }
}

// The first dry-claimed slot at or after start (cyclic) that comes before
// the next barrier, or kNoSlot when a barrier comes first.
__device__ inline uint32_t arcs_first_member(const unsigned long long* P, const unsigned long long* Q,
                                             uint64_t start, uint64_t words) {
  uint64_t w = start >> 6;
  unsigned long long m = ~0ull << (start & 63);
  for (uint64_t it = 0; it <= words; ++it) {
    const unsigned long long p = P[w], qq = Q[w];
    const unsigned long long cb = p & ~qq & m, bb = p & qq & m;
    if (cb && (!bb || __ffsll(cb) < __ffsll(bb))) return static_cast<uint32_t>((w << 6) + __ffsll(cb) - 1);
    if (bb) return kNoSlot;
    m = ~0ull;
    w = w + 1 == words ? 0 : w + 1;
  }
  return kNoSlot;
}

// At most 72 SGPRs (78 with VCC and the other reserved ones): two of its
// 1024-thread workgroups then fit on a CU (measured on MI355X: 20 waves per
// CU at 78 SGPRs, 11 at 88, and the kernel 81 against 124 us on 2^20 slots).
template <int KS, int VS>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_num_sgpr(72), amdgpu_waves_per_eu(8)))
k_ins_arcs(Table t, const uint8_t* __restrict__ in_keys, const uint8_t* __restrict__ in_vals,
           const unsigned long long* __restrict__ Q, uint32_t* __restrict__ big, Counters* ctr) {
  static_assert(KS > 0 && VS > 0 && KS % 4 == 0 && VS % 4 == 0, "static sizes");
  __shared__ uint32_t mk_lds[1024 / 64 * (KS / 4) * 64];  // per wave: the members' keys, word-major
  const unsigned long long* P = t.freebits;
  const uint64_t words = t.cap >> 6;
  const uint64_t w = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;  // wave-uniform
  uint32_t claimed = 0;
  if (w < words && !(ctr->status & kStatusDryFail)) {
    const uint64_t wp = w == 0 ? words - 1 : w - 1, wn = w + 1 == words ? 0 : w + 1;
    const unsigned long long p0 = P[wp], p1 = P[w], p2 = P[wn];
    const unsigned long long q0 = Q[wp], q1 = Q[w], q2 = Q[wn];
    const unsigned long long c1 = p1 & ~q1, c2 = p2 & ~q2;              // dry claims
    const unsigned long long b0 = p0 & q0, b1 = p1 & q1, b2 = p2 & q2;  // barriers
    unsigned long long st = ~b1 & ((b1 << 1) | (b0 >> 63));            // arcs starting here
    unsigned long long act = 0, mem = 0, starts_in = 0;
    uint32_t cross = 64, cross_len = 0;
    while (st) {
      const uint32_t k = static_cast<uint32_t>(__ffsll(st) - 1);
      st &= st - 1;
      const unsigned long long z1 = b1 & (~0ull << k);
      if (z1) {  // ends inside the word
        const uint32_t x = static_cast<uint32_t>(__ffsll(z1) - 1);
        const unsigned long long bits = (1ull << x) - (1ull << k);
        if (c1 & bits) {
          act |= bits;
          mem |= c1 & bits;
          starts_in |= 1ull << k;
        }
        continue;
      }
      const unsigned long long cw = c1 & (~0ull << k);
      uint32_t s = kNoSlot;
      if (b2) {
        const uint32_t x = static_cast<uint32_t>(__ffsll(b2) - 1);
        const unsigned long long c2b = c2 & ((1ull << x) - 1ull);
        if (!(cw | c2b)) continue;  // no member
        if (64u - k + x <= kArcSlots) {
          cross = k;
          cross_len = 64u - k + x;
          continue;
        }
        s = cw ? static_cast<uint32_t>((w << 6) + __ffsll(cw) - 1) : static_cast<uint32_t>((wn << 6) + __ffsll(c2b) - 1);
      } else if (cw) {
        s = static_cast<uint32_t>((w << 6) + __ffsll(cw) - 1);
      } else if (c2) {
        s = static_cast<uint32_t>((wn << 6) + __ffsll(c2) - 1);
      } else {
        s = arcs_first_member(P, Q, (wn + 1 == words ? 0 : wn + 1) << 6, words);
      }
      if (s != kNoSlot && (threadIdx.x & 63u) == 0) big[atomicAdd(&ctr->nbig, 1u)] = s;  // longer than kArcSlots
    }
    uint32_t got = 0;
    // claims in the next bitmaps go out after both rounds: a store or atomic
    // still in flight holds up the second round's loads (one counter)
    unsigned long long cl1 = 0, cl2 = 0;
    const bool nb = t.nfree != nullptr;
    uint32_t* mkl = mk_lds + (threadIdx.x >> 6) * (KS / 4) * 64;
    if (act) got += arcs_round<KS, VS>(t, in_keys, in_vals, w << 6, act, starts_in, mem, ctr, cl1, 0, 0, mkl);
    if (cross < 64) {
      const unsigned long long a2 = cross_len == 64 ? ~0ull : ((1ull << cross_len) - 1ull);
      const unsigned long long m2 = (cross == 0 ? c1 : ((c1 >> cross) | (c2 << (64u - cross)))) & a2;
      got += arcs_round<KS, VS>(t, in_keys, in_vals, (w << 6) + cross, a2, 1ull, m2, ctr, cl2, 0, 0, mkl);
    }
    if (nb && (threadIdx.x & 63u) == 0) {
      claim_next_mask(t, w << 6, cl1);
      claim_next_mask(t, (w << 6) + cross, cl2);
    }
    if ((threadIdx.x & 63u) == 0) claimed = got;
  }
  block_add_nowait(&ctr->count, claimed);
}

// The last workgroup of an insert writes the op's final counters to
// host_out (pinned host memory) and zeroes them for the next op.  host_out->
// done then takes the op's sequence number with a system-scope release, so a
// host that polls it reads the other fields complete without waiting for the
// kernel to end.  The release writes back this XCD's L2 once per op.
__device__ inline void hand_to_host(Counters* ctr, Counters* host_out, uint32_t seq, unsigned long long count) {
  volatile Counters* h = host_out;
  h->status = ld_agent(&ctr->status);
  h->nbig = ld_agent(&ctr->nbig);
  h->count = count;
  h->nsmall = 0;
  *ctr = Counters{};
  __hip_atomic_store(&host_out->done, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Completion tickets of the leader's hand-over ride in the high bits of
// Counters::count (a count stays below 2^40).
constexpr uint32_t kTicketShift = 40;
constexpr unsigned long long kTicket = 1ull << kTicketShift;

// The leader's hand-over (last workgroup; every other one has added its
// count).  No closure queued for k_ins_big: the count is final and goes to
// the host now; the device counters are zeroed here (zero_final: claims go
// to the next bitmaps, no settle pass follows) or by the settle pass after
// the leader (k_ins_bits with `early`).  Closures queued: the host copy
// carries kStatusBigPending and the device keeps the count (tickets taken
// out) and the queue for k_ins_big and the hand-over after it.
// host_out->done takes `seq` with a system-scope release, as hand_to_host.
__device__ inline void hand_early(Counters* ctr, Counters* host_out, uint32_t seq, unsigned long long count,
                                  bool zero_final) {
  volatile Counters* h = host_out;
  const uint32_t nbig = ld_agent(&ctr->nbig);
  const uint32_t status = ld_agent(&ctr->status);
  if (!nbig && zero_final) *ctr = Counters{};
  else __hip_atomic_store(&ctr->count, count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  h->status = nbig ? (status | kStatusBigPending) : status;
  h->nbig = nbig;
  h->count = count;
  h->nsmall = 0;
  __hip_atomic_store(&host_out->done, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Leaders order and replay closures of up to leader_max (<= kLeaderMax)
// keys whose arc (barrier to barrier) is at most `coop_arc` slots; other
// closures are queued for k_ins_big.  With host_out (counted sparse
// inserts) every workgroup adds its claims and a completion ticket to
// ctr->count in one atomic, and the last one hands the count to the host
// (hand_early) before the settle pass runs: the settle reads no batch input,
// so the host can return while it finishes on the map's stream.  With PLAN (sparse batches, where
// nearly every key is a one-member closure) every lane first reads its own
// key and the flags of its home .. dry slot together with the election's
// bitmap words, so a one-member closure is decided and applied with two
// further rounds of memory traffic (insert_planned).  The instance without
// it (dense batches when k_ins_arcs is not used) holds no plan registers.
// One wave's step of the leader pass: lane i's key (dry slot sd, kNoSlot
// for none) leads its closure when the slot before it (in the free-slot
// bitmap) is a barrier, and then inserts the closure's members: a planned
// one-member insert, the packed replay, or walks; closures past the
// leader's reach go to push_big(s).  LDS: lst (kLeaderMax x 64 words,
// column per lane), mv_src/mv_dst (64 x kPlanMoves), mv_stage
// (64 x VS/16 16-byte parts), arc_at/arc_src (64 each): the wave's own.
// Returns the free slots this lane claimed (packed replays count on lane 0).
template <int KS, int VS, bool PLAN, typename PushBig>
__device__ inline unsigned long long leader_wave(const Table& t, const uint8_t* __restrict__ in_keys,
                                                 const uint8_t* __restrict__ in_vals, uint64_t i, uint32_t sd,
                                                 const unsigned long long* __restrict__ Q, Counters* ctr,
                                                 uint32_t coop_arc, uint32_t leader_max, uint32_t* lst,
                                                 uint64_t* mv_src, uint64_t* mv_dst, u32x4* mv_stage,
                                                 uint64_t* arc_at, uint8_t* arc_src, PushBig push_big) {
  // planned one-member inserts for static sizes; values of 16-byte multiples
  // are copied by the whole wave, smaller ones by each lane
  constexpr bool kPlan = PLAN && KS > 0 && KS % 4 == 0 && VS > 0 && VS % 4 == 0;
  constexpr bool kCoop = kPlan && VS % 16 == 0;
  const uint32_t K = KS ? KS : t.K;
  const uint32_t V = VS ? VS : t.V;
  const unsigned long long* P = t.freebits;
  const uint64_t words = (t.cap + 63) >> 6;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t* my = lst + lane;
  unsigned long long claimed = 0;
  uint64_t mv[2 * kPlanMoves];
  uint32_t nmv = 0;
  bool packed = false;  // this lane's closure goes to the packed replay below
  uint64_t arc0 = 0;
  uint32_t alen = 0;
  unsigned long long dmask = 0;
  if (sd != kNoSlot) {
    const uint64_t s = sd;
    Blob<KS> ck;
    uint64_t home = 0;
    uint32_t L = ~0u;
    uint32_t fl[kPlanSlots];
    if (kPlan) {
      ck.load(in_keys + i * K, K);
      home = ck.hash(K, t.mode) & t.mask;
      L = static_cast<uint32_t>(min<uint64_t>((s - home) & t.mask, kPlanSlots + 1));
#pragma unroll
      for (uint32_t j = 0; j < kPlanSlots; ++j) fl[j] = j < L ? t.flags[(home + j) & t.mask] : 0u;
    }
    // The free-slot and claim words at the dry slot, read together (and
    // with the flags above); most closures begin, and end, inside that word.
    uint64_t cw = s >> 6;
    unsigned long long cp = P[cw], cq = Q[cw];
    if (kPlan) {
#pragma unroll
      for (uint32_t j = 0; j < kPlanSlots; ++j) asm volatile("" ::"v"(fl[j]));
    }
    const unsigned long long below = cp & ((1ull << (s & 63)) - 1ull);
    uint64_t b;
    unsigned long long qb;
    if (below) {
      b = (cw << 6) + static_cast<uint64_t>(63 - __clzll(below));
      qb = cq;
    } else {
      b = bm_prev_set(P, s, t.cap);
      qb = Q[b >> 6];
    }
    if (b != s && ((qb >> (b & 63)) & 1ull)) {
      uint32_t m = 0;
      uint64_t pos = s;
      bool long_arc = false;
      const uint64_t a_first = (b + 1) & t.mask;
      for (;;) {
        // next pre-batch free slot at/after pos (cyclic), from the cached
        // word pair or the next ones
        if ((pos >> 6) != cw) {
          cw = pos >> 6;
          cp = P[cw];
          cq = Q[cw];
        }
        unsigned long long bits = cp & (~0ull << (pos & 63));
        for (uint64_t it = 0; !bits && it <= words; ++it) {  // b is set: found within a lap
          cw = (cw + 1 == words) ? 0 : cw + 1;
          cp = P[cw];
          cq = Q[cw];
          bits = cp;
        }
        const uint64_t x = (cw << 6) + static_cast<uint64_t>(__ffsll(bits) - 1);
        if ((cq >> (x & 63)) & 1ull) {  // the closure's barrier
          alen = static_cast<uint32_t>(min<uint64_t>((x - b - 1) & t.mask, 0xFFFFFFFFull));
          long_arc = alen > coop_arc;
          break;
        }
        if (m == leader_max) {
          m = kLeaderMax + 1;
          break;
        }
        my[m * 64] = (x == s) ? static_cast<uint32_t>(i) : t.owner[x];  // owner[s] == i (k_ins_dry)
        const uint64_t dx = (x - a_first) & t.mask;
        if (dx < 64) dmask |= 1ull << dx;  // the members' dry slots in the arc, in slot order
        ++m;
        pos = (x + 1 == t.cap) ? 0 : x + 1;
      }
      uint32_t c1 = 0;
      if (m > kLeaderMax || long_arc) {
        push_big(s);
      } else if (kPlan && m == 1 &&
                 insert_planned<KS, VS>(t, ck, in_vals + dd_value_index(t, i) * V, home, s, L, fl, &c1, mv, &nmv)) {
        claimed += c1;
        if (c1 && t.nfree) claim_next(t, s);
        if (t.outcome) t.outcome[i] = static_cast<uint8_t>(c1);
      } else if (m == 1 && !(kPlan)) {  // key i alone, dense batch: walked from its own pair
        uint64_t at = 0;
        c1 = insert_one<KS, VS, false>(t, in_keys + i * K, in_vals + dd_value_index(t, i) * V, K, V, ctr, &at);
        claimed += c1;
        if (c1 && t.nfree) claim_next(t, at);
        if (t.outcome) t.outcome[i] = static_cast<uint8_t>(c1);
      } else if (kPlan && alen <= kArcSlots && t.cap >= 256) {
        packed = true;  // (replayed with the wave's other such closures after this step)
        arc0 = (b + 1) & t.mask;
      } else {
        for (uint32_t a = 1; a < m; ++a) {  // insertion sort by batch index
          const uint32_t v = my[a * 64];
          uint32_t c = a;
          while (c > 0 && my[(c - 1) * 64] > v) {
            my[c * 64] = my[(c - 1) * 64];
            --c;
          }
          my[c * 64] = v;
        }
        for (uint32_t j = 0; j < m; ++j) {
          const uint64_t k = my[j * 64];
          uint64_t at = 0;
          c1 = insert_one<KS, VS, false>(t, in_keys + k * K, in_vals + dd_value_index(t, k) * V, K, V, ctr, &at);
          claimed += c1;
          if (c1 && t.nfree) claim_next(t, at);
          if (t.outcome) t.outcome[k] = static_cast<uint8_t>(c1);
        }
      }
    }
  }
  if constexpr (kPlan) {
    // Closures that the plan does not cover (several members, or a walk
    // past the plan's reach) whose arcs hold at most 64 slots: packed side
    // by side into the 64 lanes (arc after arc in lane order, as many as
    // fit) and replayed together by wave scans, one member per arc and
    // step (arcs_round), instead of walked slot by slot.
    unsigned long long todo = __ballot(packed);
    while (todo) {
      const bool mine = (todo >> lane) & 1ull;
      const uint32_t len = mine ? alen : 0u;
      const uint32_t incl = wave_incl_add(len);
      const bool take = mine && incl <= 64;
      const unsigned long long took = __ballot(take);
      const uint32_t first = incl - len;  // this arc's first lane
      if (take) {
        arc_at[first] = arc0;
        arc_src[first] = static_cast<uint8_t>(lane);
      }
      unsigned long long starts = 0;
      for (unsigned long long tk = took; tk; tk &= tk - 1)
        starts |= 1ull << rd_lane(first, static_cast<uint32_t>(__ffsll(tk) - 1));
      const uint32_t total = rd_lane(incl, static_cast<uint32_t>(63 - __clzll(took)));
      wave_sync();
      const unsigned long long act = total >= 64 ? ~0ull : ((1ull << total) - 1ull);
      const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
      const unsigned long long sm = starts & upto;
      const uint32_t a0 = sm ? 63u - static_cast<uint32_t>(__clzll(sm)) : 0u;
      const uint32_t src = arc_src[a0], off = lane - a0;
      const uint64_t slot = (arc_at[a0] + off) & t.mask;
      // the arc's members: the leader lane's dry-slot mask and batch indices (slot order)
      const unsigned long long dm = rd_any64(dmask, src);
      const bool is_mem = lane < total && ((dm >> off) & 1ull);
      const uint32_t o = is_mem ? lst[static_cast<uint32_t>(__popcll(dm & ((1ull << off) - 1ull))) * 64 + src] : 0u;
      const unsigned long long mem = __ballot(is_mem);
      unsigned long long claims = 0;
      claimed += arcs_round<KS, VS, true>(t, in_keys, in_vals, 0, act, starts, mem, ctr, claims, slot, o) *
                 (lane == 0 ? 1u : 0u);
      if (((claims >> lane) & 1ull) && t.nfree) claim_next(t, slot);
      wave_sync();
      todo &= ~took;
    }
  }
  if constexpr (kPlan && !kCoop) {
    // the planned inserts' value copies, each lane its own: every source
    // read before the first write
    Blob<VS> a, b, c;
    if (nmv > 0) a.load(reinterpret_cast<const uint8_t*>(mv[0]), V);
    if (nmv > 1) b.load(reinterpret_cast<const uint8_t*>(mv[2]), V);
    if (nmv > 2) c.load(reinterpret_cast<const uint8_t*>(mv[4]), V);
    if (nmv > 0) a.store(reinterpret_cast<uint8_t*>(mv[1]), V);
    if (nmv > 1) b.store(reinterpret_cast<uint8_t*>(mv[3]), V);
    if (nmv > 2) c.store(reinterpret_cast<uint8_t*>(mv[5]), V);
  }
  if constexpr (kCoop) {
    // the planned inserts' value copies, by the whole wave (wave-uniform)
    uint32_t incl = nmv;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), off, 64));
      if (lane >= static_cast<uint32_t>(off)) incl += y;
    }
    const uint32_t total = static_cast<uint32_t>(__shfl(static_cast<int>(incl), 63, 64));
    if (total) {
      const uint32_t first = incl - nmv;
#pragma unroll
      for (uint32_t k = 0; k < kPlanMoves; ++k) {
        if (k < nmv) {
          mv_src[first + k] = mv[2 * k];
          mv_dst[first + k] = mv[2 * k + 1];
        }
      }
      wave_sync();
      wave_copy_values<VS>(mv_src, mv_dst, total, mv_stage);
      wave_sync();
    }
  }
  return claimed;
}

template <int KS, int VS, bool PLAN>
__global__ void __launch_bounds__(64)
k_ins_leader(Table t, const uint8_t* __restrict__ in_keys, const uint8_t* __restrict__ in_vals, uint64_t n,
             const unsigned long long* __restrict__ Q, const uint32_t* __restrict__ dslot, uint32_t* __restrict__ big,
             Counters* ctr, uint32_t coop_arc, uint32_t leader_max, Counters* host_out, uint32_t seq) {
  constexpr bool kPlan = PLAN && KS > 0 && KS % 4 == 0 && VS > 0 && VS % 4 == 0;
  constexpr bool kCoop = kPlan && VS % 16 == 0;
  __shared__ uint32_t lst[kLeaderMax * 64];  // column per lane: lst[j * 64 + lane]
  __shared__ uint64_t mv_src[kCoop ? 64 * kPlanMoves : 1], mv_dst[kCoop ? 64 * kPlanMoves : 1];
  __shared__ u32x4 mv_stage[kCoop ? 64 * (VS / 16) : 1];
  __shared__ uint64_t arc_at[kPlan ? 64 : 1];  // packed replay: the first slot of the arc starting in lane l
  __shared__ uint8_t arc_src[kPlan ? 64 : 1];   // ... and the lane that leads it
  const bool failed = (ctr->status & kStatusDryFail) != 0;  // block-uniform
  unsigned long long claimed = 0;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x;; base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (failed || base >= n) break;
    const uint64_t i = base + threadIdx.x;
    const uint32_t sd = i < n ? dslot[i] : kNoSlot;
    claimed += leader_wave<KS, VS, PLAN>(t, in_keys, in_vals, i, sd, Q, ctr, coop_arc, leader_max, lst, mv_src, mv_dst,
                                         mv_stage, arc_at, arc_src, [&](uint64_t s) {
                                           big[atomicAdd(&ctr->nbig, 1u)] = static_cast<uint32_t>(s);
                                         });
  }
  if (!host_out) {
    block_add(&ctr->count, claimed);
    return;
  }
  const unsigned long long sum = block_sum(claimed);
  if (threadIdx.x == 0) {
    const unsigned long long old = atomicAdd(&ctr->count, sum + kTicket);
    if ((old >> kTicketShift) == gridDim.x - 1u)
      hand_early(ctr, host_out, seq, (old + sum) & (kTicket - 1u), t.nfree != nullptr);
  }
}

// Sorts lst[0..m) ascending (distinct values below n) with one wave:
// counts per bucket of the values' top bits (B buckets, a power of two),
// offsets, a scatter into tmp[m], then each bucket (one lane each) by
// insertion, and the result copied back.
__device__ inline void sort_by_buckets(uint32_t* lst, uint32_t m, uint64_t n, uint32_t* tmp, uint32_t* bc,
                                       uint32_t B) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t nb = 64u - static_cast<uint32_t>(__clzll(static_cast<unsigned long long>(n > 2 ? n - 1 : 1)));
  const uint32_t lb = static_cast<uint32_t>(__ffs(B) - 1);
  const uint32_t sh = nb > lb ? nb - lb : 0u;
  for (uint32_t b = lane; b < B; b += 64) bc[b] = 0;
  wave_sync();
  for (uint32_t j = lane; j < m; j += 64) atomicAdd(&bc[lst[j] >> sh], 1u);
  wave_sync();
  uint32_t sum = 0;
  for (uint32_t b0 = 0; b0 < B; b0 += 64) {
    const uint32_t v = bc[b0 + lane];
    const uint32_t incl = wave_incl_add(v);
    bc[b0 + lane] = sum + incl - v;
    sum += rd_lane(incl, 63);
  }
  wave_sync();
  for (uint32_t j = lane; j < m; j += 64) {
    const uint32_t v = lst[j];
    tmp[atomicAdd(&bc[v >> sh], 1u)] = v;
  }
  wave_sync();
  for (uint32_t b = lane; b < B; b += 64) {  // bc[b]: the end of bucket b
    const uint32_t beg = b ? bc[b - 1] : 0u, end = bc[b];
    for (uint32_t i = beg + 1; i < end; ++i) {
      const uint32_t v = tmp[i];
      uint32_t k = i;
      while (k > beg && tmp[k - 1] > v) {
        tmp[k] = tmp[k - 1];
        --k;
      }
      tmp[k] = v;
    }
  }
  wave_sync();
  for (uint32_t j = lane; j < m; j += 64) lst[j] = tmp[j];
  wave_sync();
}

// Replays one closure with a whole wave (64 active lanes).  The arc runs
// from `s`, the dry slot of the closure's first member, up to its barrier
// (the next bit still set in Q).  The members' batch indices are gathered
// from owner[] at the arc's dry-claimed slots (P & ~Q), sorted (bitonic, in
// windows of `lcap` indices when there are more), and inserted in batch
// order: a canonical arc of at least t.canon_min members from its groups
// (replay_arc_canon), an arc that fits in LDS by wave scans
// (replay_arc_lds), any other by walks of the whole wave (static sizes) or
// of lane 0.  With static sizes `lst` starts a pool of `pool_words` LDS
// words that the first two use.  The arc's real claims are cleared in the
// next bitmaps (Table::nfree) when the map keeps them.  Returns the free
// slots claimed (in every lane).
template <int KS, int VS>
__device__ uint32_t replay_arc(const Table& t, const uint8_t* __restrict__ in_keys,
                               const uint8_t* __restrict__ in_vals, uint64_t n, unsigned long long* Q, uint64_t s,
                               uint32_t* lst, uint32_t lcap, uint32_t* carry, uint32_t* cnt, Counters* ctr,
                               uint32_t pool_words) {
  constexpr bool kStatic = KS > 0 && VS > 0 && KS % 4 == 0 && VS % 4 == 0;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t K = KS ? KS : t.K;
  const uint32_t V = VS ? VS : t.V;
  const unsigned long long* P = t.freebits;
  const uint64_t words = (t.cap + 63) >> 6;
  long long xb = 0;
  if (lane == 0) xb = static_cast<long long>(bm_next_set_agent(Q, s, t.cap));
  const uint64_t x = static_cast<uint64_t>(__shfl(xb, 0, 64));
  const uint64_t arc = (x + t.cap - s) & t.mask;  // slots s .. x-1
  const uint64_t w0 = s >> 6;
  const uint64_t nwords = min(((s & 63) + arc + 63) >> 6, words);
  // The whole arc runs from the barrier before s (b) to x; a canonical
  // replay gathers all of its members at once (there are at most Lb).
  uint64_t b = s, Lb = 0;
  if (kStatic && pool_words) {
    long long bb = 0;
    if (lane == 0) bb = static_cast<long long>(bm_prev_set(P, s, t.cap));
    b = static_cast<uint64_t>(__shfl(bb, 0, 64));
    Lb = (x - b - 1) & t.mask;
  }
  const bool canon = kStatic && pool_words && t.canon_min && b != s && Lb <= kCanonMaxL && 2 * Lb <= pool_words;
  const uint32_t lcap0 = lcap;
  if (canon) lcap = static_cast<uint32_t>(Lb);
  uint32_t claimed = 0;
  uint64_t lo = 0, hi = ~0ull;
  bool windowed = false;
  while (lo < n) {
    uint32_t m = 0;
    // (the words that slots s .. s+arc-1 fall in, counted from w0 without
    // wrapping: a word can appear twice when the arc wraps the table)
    const uint64_t nwc = ((s & 63) + arc + 63) >> 6;
    if (kStatic && pool_words && nwc <= 128 && t.cap >= 64 && !windowed) {
      // The arc's claim words in registers (lane l: words l and l + 64 from
      // w0), then four 64-slot chunks at a time: the owners of the claimed
      // slots read together, positions from the ballots.
      const auto claim_word = [&](uint64_t i) -> unsigned long long {
        const uint64_t w = (w0 + i) % words;
        return i < nwc ? (P[w] & ~load_word(&Q[w])) : 0ull;
      };
      const unsigned long long cb0 = claim_word(lane), cb1 = claim_word(lane + 64);
      const uint32_t o = static_cast<uint32_t>(s & 63);
      const auto word_at = [&](uint32_t i) -> unsigned long long {
        return i < 64 ? rd_lane64(cb0, i) : (i < 128 ? rd_lane64(cb1, i - 64) : 0ull);
      };
      for (uint64_t c0 = 0; c0 < arc; c0 += 256) {
        uint32_t idx[4];
        bool take[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const uint32_t c = static_cast<uint32_t>(c0 >> 6) + u;
          const unsigned long long bits = o ? ((word_at(c) >> o) | (word_at(c + 1) << (64 - o))) : word_at(c);
          const uint64_t q = (static_cast<uint64_t>(c) << 6) + lane;
          take[u] = q < arc && ((bits >> lane) & 1ull);
          idx[u] = take[u] ? __hip_atomic_load(&t.owner[(s + q) & t.mask], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const unsigned long long bal = __ballot(take[u]);
          const uint32_t k = m + static_cast<uint32_t>(__popcll(bal & ((1ull << lane) - 1ull)));
          if (take[u] && k < lcap) lst[k] = idx[u];
          m += static_cast<uint32_t>(__popcll(bal));
        }
      }
      wave_sync();
    } else {
    if (lane == 0) *cnt = 0;
    wave_sync();
    for (uint64_t wi = lane; wi < nwords; wi += 64) {
      const uint64_t w = (w0 + wi) % words;
      unsigned long long bits = P[w] & ~load_word(&Q[w]);
      while (bits) {
        const int bt = __ffsll(bits) - 1;
        bits &= bits - 1;
        const uint64_t p = (w << 6) + bt;
        if (((p + t.cap - s) & t.mask) >= arc) continue;
        const uint32_t idx = __hip_atomic_load(&t.owner[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (idx < lo || idx >= hi) continue;
        const uint32_t k = atomicAdd(cnt, 1u);
        if (k < lcap) lst[k] = idx;
      }
    }
    wave_sync();
    m = *cnt;
    wave_sync();
    }
    if (m > lcap) {  // too many for one pass: windows of lcap indices
      windowed = true;
      hi = lo + lcap;
      continue;
    }
    uint32_t p2 = 1;
    while (p2 < m) p2 <<= 1;
    if (kStatic && pool_words && m > 64 && 2 * ((m + 3) & ~3u) + p2 <= pool_words) {
      sort_by_buckets(lst, m, n, lst + ((m + 3) & ~3u), lst + 2 * ((m + 3) & ~3u), p2 < 64 ? 64u : p2);
      p2 = 0;  // (sorted)
    }
    for (uint32_t k = m + lane; k < p2; k += 64) lst[k] = 0xFFFFFFFFu;
    wave_sync();
    for (uint32_t size = 2; size <= p2; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t k = lane; k < p2; k += 64) {
          const uint32_t j = k ^ stride;
          if (j > k) {
            const uint32_t a = lst[k], b2 = lst[j];
            if ((a > b2) == ((k & size) == 0)) {
              lst[k] = b2;
              lst[j] = a;
            }
          }
        }
        wave_sync();
      }
    }
    if constexpr (kStatic) {
      bool done = false;
      if (canon && !windowed && m >= t.canon_min && canon_words(m, static_cast<uint32_t>(Lb)) <= pool_words) {
        const uint32_t c = replay_arc_canon<KS, VS>(t, in_keys, in_vals, (b + 1) & t.mask, static_cast<uint32_t>(Lb),
                                                    lst, m, lst);
        wave_sync();
        if (c != kNotCanon) {
          claimed += c;
          done = true;
        }
      }
      // The arc in LDS when it fits (after the member list).
      const uint32_t arc_off = max(lcap0, (m + 3u) & ~3u);
      if (done) {
      } else if (pool_words && !windowed && b != s && Lb <= arc_lds_cap<KS>() && arc_off + kArcLdsWords <= pool_words) {
        claimed += replay_arc_lds<KS, VS>(t, in_keys, in_vals, (b + 1) & t.mask, static_cast<uint32_t>(Lb), lst, m,
                                          lst + arc_off, ctr);
      } else {
        for (uint32_t j = 0; j < m; ++j) {
          const uint64_t k = lst[j];
          const uint32_t c1 =
              insert_one_wave<KS, VS>(t, in_keys + k * K, in_vals + dd_value_index(t, k) * V, carry, ctr);
          claimed += c1;
          if (lane == 0 && t.outcome) t.outcome[k] = static_cast<uint8_t>(c1);
        }
      }
    } else {
      uint32_t c0 = 0;
      if (lane == 0) {
        for (uint32_t j = 0; j < m; ++j) {
          const uint64_t k = lst[j];
          const uint32_t c1 =
              insert_one<KS, VS, false>(t, in_keys + k * K, in_vals + dd_value_index(t, k) * V, K, V, ctr);
          c0 += c1;
          if (t.outcome) t.outcome[k] = static_cast<uint8_t>(c1);
        }
      }
      claimed += static_cast<uint32_t>(__shfl(static_cast<int>(c0), 0, 64));
    }
    wave_sync();
    if (!windowed) break;
    lo = hi;
    hi = lo + lcap;
  }
  if (t.nfree) {  // the arc's real claims, in the next bitmaps
    __threadfence_block();  // the replay's flag stores, seen by every lane
    for (uint64_t wi = lane; wi < nwords; wi += 64) {
      const uint64_t w = (w0 + wi) % words;
      unsigned long long bits = P[w] & ~load_word(&Q[w]);
      unsigned long long taken = 0;
      while (bits) {
        const int bt = __ffsll(bits) - 1;
        bits &= bits - 1;
        const uint64_t p = (w << 6) + bt;
        if (((p + t.cap - s) & t.mask) >= arc) continue;
        if ((t.flags[p] & 0xFu) == kOccupied) taken |= 1ull << bt;
      }
      if (taken) {
        atomicAnd(&t.nfree[w], ~taken);
        atomicAnd(&t.nempty[w], ~taken);
        atomicAnd(&t.nq[w], ~taken);
      }
    }
  }
  return claimed;
}

// One single-wave workgroup per queued closure.  With host_out (counted
// inserts whose claims go to the next bitmaps) every workgroup takes a
// completion ticket on ctr->count once it is done, and the last one hands
// the counters to the host (hand_to_host).
template <int KS, int VS>
__global__ void __launch_bounds__(64)
k_ins_big(Table t, const uint8_t* __restrict__ in_keys, const uint8_t* __restrict__ in_vals, uint64_t n,
          const unsigned long long* __restrict__ Q, const uint32_t* __restrict__ big, Counters* ctr,
          Counters* host_out, uint32_t seq) {
  // static sizes: the member list, then the LDS replays' arrays
  constexpr uint32_t kPool = (KS > 0 && VS > 0 && KS % 4 == 0 && VS % 4 == 0) ? kBigPoolWords : 0u;
  __shared__ uint32_t lst[kPool ? kPool : kBigMax];
  __shared__ uint32_t carry[2 * kMaxGenericSize / 4];
  __shared__ uint32_t cnt;
  __shared__ uint32_t cur;
  const uint32_t nbig = __hip_atomic_load(&ctr->nbig, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // Closures are handed out one at a time (ctr->nsmall is the ticket; their
  // replay times differ by orders of magnitude).  Workgroups beyond the
  // queue's length take none (the others take every ticket below it).
  const uint32_t parts = min(nbig, gridDim.x);  // the workgroups that take a closure
  for (bool more = blockIdx.x < parts; more;) {
    if (threadIdx.x == 0) cur = atomicAdd(&ctr->nsmall, 1u);
    __syncthreads();
    const uint32_t c = cur;
    __syncthreads();
    if (c >= nbig) break;
    const uint32_t claimed = replay_arc<KS, VS>(t, in_keys, in_vals, n, const_cast<unsigned long long*>(Q), big[c],
                                                lst, kBigMax, carry, &cnt, ctr, kPool);
    if (threadIdx.x == 0 && claimed) atomicAdd(&ctr->count, static_cast<unsigned long long>(claimed));
  }
  // Tickets from the workgroups that took closures only (tickets on one
  // word serialise); a workgroup's count adds precede its ticket on that
  // word.  With no closure queued, workgroup 0 hands over at once.
  if (host_out && threadIdx.x == 0) {
    if (parts == 0) {
      if (blockIdx.x == 0)
        hand_to_host(ctr, host_out, seq, __hip_atomic_load(&ctr->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    } else if (blockIdx.x < parts) {
      const unsigned long long old = atomicAdd(&ctr->count, kTicket);
      if ((old >> kTicketShift) == parts - 1u) hand_to_host(ctr, host_out, seq, old & (kTicket - 1u));
    }
  }
}

// The hand-over of a settle pass that found closures queued for k_ins_big
// (not yet run): the host copy's status carries kStatusBigPending, the
// device counters stay as they are (the count so far, the queue length)
// except the completion tickets, which the settle after k_ins_big reuses.
__device__ inline void hand_pending(Counters* ctr, Counters* host_out, uint32_t seq) {
  volatile Counters* h = host_out;
  h->status = ld_agent(&ctr->status) | kStatusBigPending;
  h->nbig = ld_agent(&ctr->nbig);
  h->count = __hip_atomic_load(&ctr->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  h->nsmall = 0;
  ctr->done = 0;
  __hip_atomic_store(&host_out->done, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Publishes the batch's real claims (a subset of its dry claims) in both
// bitmaps.  Q (optional): the dry pass's
// claim bitmap; a dry claim that ended unused (the key updated an existing
// slot) gets its Q bit back, which leaves Q equal to the new freebits, so
// the next insert needs no copy.  host_out (optional): the last workgroup
// writes the op's final counters there and zeroes them for the next op
// (hand_to_host).  defer_big: k_ins_big has not run; if the leader queued
// closures for it, their slots are not final, so the pass settles nothing
// (deduplication state included) and hands over kStatusBigPending instead.
// early: the leader has already handed the counters over (hand_early); with
// closures pending the pass does nothing (k_ins_big and a settle with a
// hand-over follow), otherwise it settles and its last workgroup zeroes the
// counters for the next op.
static __global__ void __launch_bounds__(1024)
k_ins_bits(Table t, const uint32_t* __restrict__ dslot, uint64_t n, Counters* ctr, unsigned long long* Q,
           Counters* host_out, uint32_t seq, bool defer_big, bool early) {
  const bool failed = (ctr->status & kStatusDryFail) != 0;
  const bool pending = (defer_big || early) && ld_agent(&ctr->nbig) != 0;  // the same in every workgroup
  if (early && pending) return;
  const uint64_t end = pending ? 0 : n;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < end;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (t.dd_first) {
      const uint32_t b = t.dd_bucket[i] & ~kDupKey;
      t.dd_first[b] = kNoOwner;
      t.dd_last[b] = 0;
    }
    if (failed) continue;
    const uint64_t s = dslot[i];
    if (s == kNoSlot) continue;
    const unsigned long long bit = 1ull << (s & 63);
    if ((t.flags[s] & 0xFu) == kOccupied) {
      atomicAnd(&t.freebits[s >> 6], ~bit);
      atomicAnd(&t.emptybits[s >> 6], ~bit);
    } else if (Q) {
      atomicOr(&Q[s >> 6], bit);
    }
  }
  // No fence: this kernel changes none of the published fields and every
  // workgroup's read of ctr->status precedes its ticket.  (An agent-scope
  // fence writes back the L2 on this part, which costs more than the whole
  // kernel; only the last workgroup's hand-over releases, once.)
  if (!host_out && !early) return;
  __syncthreads();
  if (threadIdx.x == 0 && atomicAdd(&ctr->done, 1u) == gridDim.x - 1) {
    if (early) *ctr = Counters{};
    else if (pending) hand_pending(ctr, host_out, seq);
    else hand_to_host(ctr, host_out, seq, __hip_atomic_load(&ctr->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
}

// k_ins_bits for a batch that touches most bitmap words (the caller uses it
// when n * 16 >= cap and no deduplication is on): every word is recomputed
// from the flags, one wave per 64 slots, instead of one atomic per key.
// After the replay the flags are the truth, so freebits (not OCCUPIED),
// emptybits (EMPTY) and Q (= freebits) come out exactly as k_ins_bits
// leaves them.  A failed dry pass left
// the flags untouched, so the words come out unchanged.  The last
// workgroup hands the counters to host memory as k_ins_bits does.
static __global__ void __launch_bounds__(1024)
k_ins_bits_rebuild(Table t, Counters* ctr, unsigned long long* Q, Counters* host_out, uint32_t seq) {
  const uint64_t words = (t.cap + 63) >> 6;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t w = wave; w < words; w += waves) {  // wave-uniform
    const uint64_t i = (w << 6) + lane;
    const uint32_t st = i < t.cap ? (t.flags[i] & 0xFu) : kOccupied;
    const unsigned long long fb = __ballot(st != kOccupied);
    const unsigned long long eb = __ballot(st == kEmpty);
    if (lane == 0) {
      t.freebits[w] = fb;
      t.emptybits[w] = eb;
      if (Q) Q[w] = fb;
    }
  }
  if (!host_out) return;
  __syncthreads();
  if (threadIdx.x == 0 && atomicAdd(&ctr->done, 1u) == gridDim.x - 1)
    hand_to_host(ctr, host_out, seq, __hip_atomic_load(&ctr->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// ---------------------------------------------------------------------------
// bulk_remove, exact (see header).
// ---------------------------------------------------------------------------

// Pass 1: first event of each key on the pre-batch table; matches register
// their index in rowner[slot] with atomicMin (the earliest remover wins).
template <int KS, int G>
__global__ void __launch_bounds__(1024)
k_rem_probe(Table t, const uint8_t* __restrict__ qkeys, uint64_t n, unsigned long long* __restrict__ ev) {
  const uint32_t K = KS ? KS : t.K;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t gl = lane & (G - 1);
  const uint32_t gshift = lane & ~(G - 1u);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x / G;
  for (uint64_t qi = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; qi < n; qi += stride) {
    Blob<KS> q;
    q.load(qkeys + qi * K, K);
    const uint64_t home = q.hash(K, t.mode) & t.mask;
    uint32_t kind;
    uint64_t slot;
    group_probe<KS, G>(t, q, K, home, gl, gshift, &kind, &slot);
    if (gl == 0) {
      ev[qi] = (static_cast<unsigned long long>(kind) << 32) | slot;
      if (kind == kEvMatch) atomicMin(&t.rowner[slot], static_cast<uint32_t>(qi));
    }
  }
}

// Pass 2: conflict detection.  Key i keeps its pre-batch event unless an
// earlier key tombstones that slot: a duplicate match (owner != i) or an
// occupied terminator removed by an earlier key (owner < i).  Removes
// neither cross nor create EMPTY slots, so keys interact only inside the
// run that ends at the first EMPTY slot at/after their home (the first
// EMPTY at/after their event slot: the probe stops at EMPTY).  A conflict
// marks its run dirty (bit r of `dirty`, which is all zero between ops);
// the key that sets the bit lists the run (druns, counted in ctr->nbig) and
// empties its key list (head owner[r], an insert scratch entry that inserts
// only read after writing it).  A conflict with no EMPTY slot anywhere makes
// the whole batch one run (kStatusConflictAll).
static __global__ void __launch_bounds__(1024)
k_rem_check(Table t, const unsigned long long* __restrict__ ev, uint64_t n, uint8_t* __restrict__ action,
            unsigned long long* __restrict__ dirty, uint32_t* __restrict__ druns, Counters* ctr) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t kind = static_cast<uint32_t>(ev[i] >> 32);
    const uint64_t slot = ev[i] & 0xFFFFFFFFull;
    uint8_t a = 0;
    bool conflict = false;
    if (kind == kEvMatch) {
      const uint32_t o = t.rowner[slot];
      a = (o == static_cast<uint32_t>(i)) ? 1 : 0;
      conflict = !a;
    } else if (kind == kEvTermOcc) {
      conflict = t.rowner[slot] < static_cast<uint32_t>(i);
    }
    action[i] = a;
    if (conflict) {
      const uint64_t r = bm_next_set(t.emptybits, slot, t.cap);
      if (r >= t.cap) {
        atomicOr(&ctr->status, kStatusConflict | kStatusConflictAll);
      } else {
        atomicOr(&ctr->status, kStatusConflict);
        const unsigned long long bit = 1ull << (r & 63);
        if (!(atomicOr(&dirty[r >> 6], bit) & bit)) {
          t.owner[r] = kNoOwner;
          druns[atomicAdd(&ctr->nbig, 1u)] = static_cast<uint32_t>(r);
        }
      }
    }
  }
}

// Pass 3: apply, and reset rowner[].  With no conflict every match that
// won its slot is tombstoned.  With conflicts, the keys of dirty runs are
// linked into their run's list (nxt[i] = previous head) and left to
// k_rem_replay, the others applied here (a clean run's keys interact with
// no key of another run); with kStatusConflictAll nothing is applied.
// host_out (optional): the last workgroup hands the counters, status and
// dirty-run count included, to host memory (hand_to_host).  Each workgroup
// adds its count and its completion ticket in one atomic on ctr->count
// (ticket above bit kRemTicketShift), so the last one holds every other's
// count without a fence.  Without host_out the counters stay on the device
// for a gated k_rem_replay.
constexpr int kRemTicketShift = 40;  // counts < 2^40 keys, tickets < 2^24 workgroups
static __global__ void __launch_bounds__(1024)
k_rem_apply(Table t, const unsigned long long* __restrict__ ev, const uint8_t* __restrict__ action, uint64_t n,
            const unsigned long long* __restrict__ dirty, uint32_t* __restrict__ nxt, Counters* ctr,
            Counters* host_out, uint32_t seq) {
  const uint32_t status = ctr->status;
  const bool conflict = (status & kStatusConflict) != 0;
  const bool all = (status & kStatusConflictAll) != 0;
  unsigned long long total = 0;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x;; base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (base >= n) break;
    const uint64_t i = base + threadIdx.x;
    if (i >= n) continue;
    const uint32_t kind = static_cast<uint32_t>(ev[i] >> 32);
    const uint64_t slot = ev[i] & 0xFFFFFFFFull;
    bool deferred = all;
    if (conflict && !all) {
      const uint64_t r = kind == kEvTermEmpty ? slot : bm_next_set(t.emptybits, slot, t.cap);
      if ((dirty[r >> 6] >> (r & 63)) & 1ull) {
        nxt[i] = atomicExch(&t.owner[r], static_cast<uint32_t>(i));
        deferred = true;
      }
    }
    uint32_t removed = 0;
    if (kind == kEvMatch) {
      if (!deferred && action[i]) {
        t.flags[slot] = kTombstone;  // raw 2: distance bits cleared
        atomicOr(&t.freebits[slot >> 6], 1ull << (slot & 63));
        if (t.freecopy) atomicOr(&t.freecopy[slot >> 6], 1ull << (slot & 63));
        removed = 1;
      }
      t.rowner[slot] = kNoOwner;
    }
    if (t.outcome && !deferred) t.outcome[i] = static_cast<uint8_t>(removed);
    total += removed;
  }
  if (!host_out) {
    block_add(&ctr->count, total);
    return;
  }
  const unsigned long long sum = block_sum(total);
  if (threadIdx.x == 0) {
    constexpr unsigned long long kMask = (1ull << kRemTicketShift) - 1;
    const unsigned long long old = atomicAdd(&ctr->count, (1ull << kRemTicketShift) + sum);
    if ((old >> kRemTicketShift) == gridDim.x - 1) hand_to_host(ctr, host_out, seq, (old & kMask) + sum);
  }
}

// Sequential remove of one key (rayon_cpu.rs:276-321 applied in order).
template <int KS>
__device__ inline uint32_t remove_one(const Table& t, const uint8_t* __restrict__ kin, uint32_t K) {
  Blob<KS> q;
  q.load(kin, K);
  const uint64_t home = q.hash(K, t.mode) & t.mask;
  for (uint64_t p = 0; p <= t.mask; ++p) {
    const uint64_t idx = (home + p) & t.mask;
    const uint32_t f = t.flags[idx];
    const uint32_t st = f & 0xFu;
    const uint64_t d = f >> 4;
    if (st == kEmpty) return 0;
    if (st == kOccupied) {
      if (p > d) return 0;
      if (d == p && q.eq(t.keys + idx * K, K)) {
        t.flags[idx] = kTombstone;
        atomicOr(&t.freebits[idx >> 6], 1ull << (idx & 63));
        if (t.freecopy) atomicOr(&t.freecopy[idx >> 6], 1ull << (idx & 63));
        return 1;
      }
    }
  }
  return 0;
}

// Pass 4 (runs only after a conflict): one lane per dirty run replays the
// run's keys in batch order from the run's pre-batch state (their list,
// kRunKeys keys at a time insertion-sorted in LDS), clears the run's dirty
// bit, and the whole batch in order under kStatusConflictAll.  host_out
// (optional): the last workgroup hands the counters over as k_rem_apply
// does.
constexpr uint32_t kRunKeys = 64;
constexpr uint32_t kRemReplayBlocks = 64;
// gated: status and dirty-run count from ctr (left there by k_rem_apply
// without host_out), the kernel exiting at once without a conflict;
// otherwise they come as arguments (read back from the hand-over).
template <int KS>
__global__ void __launch_bounds__(64)
k_rem_replay(Table t, const uint8_t* __restrict__ qkeys, uint64_t n, unsigned long long* __restrict__ dirty,
             const uint32_t* __restrict__ druns, const uint32_t* __restrict__ nxt, Counters* ctr, bool gated,
             uint32_t status_arg, uint32_t nd_arg, Counters* host_out, uint32_t seq) {
  __shared__ uint32_t col[kRunKeys * 64];  // column per lane: col[j * 64 + lane]
  const uint32_t status = gated ? ctr->status : status_arg;
  if (!(status & kStatusConflict)) return;  // grid-uniform: nothing to replay
  const uint32_t K = KS ? KS : t.K;
  const uint32_t lane = threadIdx.x;
  unsigned long long removed = 0;
  if (status & kStatusConflictAll) {
    if (blockIdx.x == 0 && lane == 0) {
      for (uint64_t i = 0; i < n; ++i) {
        const uint32_t r = remove_one<KS>(t, qkeys + i * K, K);
        removed += r;
        if (t.outcome) t.outcome[i] = static_cast<uint8_t>(r);
      }
    }
  } else {
    const uint32_t nd = gated ? ctr->nbig : nd_arg;
    auto one = [&](uint32_t k) {
      const uint32_t x = remove_one<KS>(t, qkeys + static_cast<uint64_t>(k) * K, K);
      removed += x;
      if (t.outcome) t.outcome[k] = static_cast<uint8_t>(x);
    };
    for (uint32_t d = blockIdx.x * 64 + lane; d < nd; d += gridDim.x * 64) {
      const uint32_t r = druns[d];
      const uint32_t head = t.owner[r];
      // kRunKeys keys at a time, in batch order: each pass keeps the
      // smallest indices above the last one replayed (one pass for most runs)
      uint32_t after = 0;
      bool first = true;
      for (;;) {
        uint32_t m = 0;
        for (uint32_t k = head; k != kNoOwner; k = nxt[k]) {
          if (!first && k <= after) continue;
          if (m == kRunKeys && k > col[(kRunKeys - 1) * 64 + lane]) continue;
          uint32_t c = m < kRunKeys ? m++ : kRunKeys - 1;  // insertion sort, the largest dropped when full
          while (c > 0 && col[(c - 1) * 64 + lane] > k) {
            col[c * 64 + lane] = col[(c - 1) * 64 + lane];
            --c;
          }
          col[c * 64 + lane] = k;
        }
        for (uint32_t j = 0; j < m; ++j) one(col[j * 64 + lane]);
        if (m < kRunKeys) break;
        after = col[(kRunKeys - 1) * 64 + lane];
        first = false;
      }
      atomicAnd(&dirty[r >> 6], ~(1ull << (r & 63)));
    }
  }
  if (!host_out) {
    block_add(&ctr->count, removed);
    return;
  }
  const unsigned long long sum = block_sum(removed);
  if (threadIdx.x == 0) {
    constexpr unsigned long long kMask = (1ull << kRemTicketShift) - 1;
    const unsigned long long old = atomicAdd(&ctr->count, (1ull << kRemTicketShift) + sum);
    if ((old >> kRemTicketShift) == gridDim.x - 1) hand_to_host(ctr, host_out, seq, (old & kMask) + sum);
  }
}

// ---------------------------------------------------------------------------
// Table maintenance.
// ---------------------------------------------------------------------------
static __global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

// clear (kernels.cu:319-326): flags EMPTY; both bitmaps all-free.
static __global__ void k_clear(Table t, unsigned long long* q) {  // q (optional): a claim bitmap to keep equal to freebits
  const uint64_t words = (t.cap + 63) >> 6;
  const unsigned long long full = t.cap >= 64 ? ~0ull : ((1ull << t.cap) - 1ull);
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < t.cap;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    t.flags[i] = kEmpty;
    if (i < words) {
      t.freebits[i] = full;
      t.emptybits[i] = full;
      if (q) q[i] = full;
    }
  }
}

// Rebuilds both bitmaps from flags (after an import): one wave per 64 slots.
static __global__ void k_rebuild_bits(Table t) {
  const uint64_t words = (t.cap + 63) >> 6;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t w = wave; w < words; w += waves) {
    const uint64_t i = (w << 6) + lane;
    const uint32_t st = i < t.cap ? (t.flags[i] & 0xFu) : kOccupied;
    const unsigned long long fb = __ballot(st != kOccupied);
    const unsigned long long eb = __ballot(st == kEmpty);
    if (lane == 0) {
      t.freebits[w] = fb;
      t.emptybits[w] = eb;
    }
  }
}

// recount (kernels.cu:328-337): OCCUPIED slots, one atomic per wave.
static __global__ void k_count(Table t, Counters* ctr) {
  unsigned long long occ = 0;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x;; base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (base >= t.cap) break;
    const uint64_t i = base + threadIdx.x;
    occ += (i < t.cap && (t.flags[i] & 0xFu) == kOccupied) ? 1u : 0u;
  }
  block_add(&ctr->count, occ);
}

}  // namespace dev
}  // namespace fm
