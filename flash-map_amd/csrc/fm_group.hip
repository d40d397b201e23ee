// fm_group.hip -- many independent FlashMaps driven by fused launches.
//
// Every op runs over all tables of the group at once: the host cuts each
// table's batch into chunks of `block` keys (one table per chunk, so every
// workgroup works on one table) and persistent kernels walk the chunk list.
// The per-table semantics are exactly the single-table ones (fm_kernels.h):
// dry parking pass + led closure replay for inserts, conflict-checked
// removes, group-cooperative gets.  A table whose insert batch may fill it
// (len + n == cap) is replayed serially by one lane; a table whose batch would
// overflow it is skipped with TableFull (rayon_cpu.rs:134-142); a remove
// batch with a conflict is replayed serially for that table only.
// fm_group_step captures insert -> get -> remove into one hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fm_internal.h"
#include "fm_kernels.h"

namespace fm {
namespace dev {

constexpr uint32_t kModeLeader = 0;  // closures led without sorting
constexpr uint32_t kModeSkip = 1;    // no keys, or TableFull
constexpr uint32_t kModeSerial = 2;  // batch may fill the table: one lane, in order

struct TableOp {
  unsigned long long begin;  // global index of the table's first key
  unsigned long long count;  // keys for this table
  uint32_t mode;
  uint32_t pad;
};

struct Chunk {
  uint32_t table;
  uint32_t count;            // keys in the chunk (<= block)
  unsigned long long begin;  // global index of the chunk's first key
  unsigned long long local;  // index of that key within the table's batch
};

struct OpDesc {  // device-resident description of one op over the group
  const TableOp* tops;
  const Chunk* chunks;
  const uint32_t* n_chunks;
};

// nbig[0]: closures queued for the big replays, nbig[1]: kg_ins_big's work
// ticket, nbig[2 + b]: queued closures in size class b,
// nbig[2 + kBigClasses + b]: the order pass's fill cursor of class b,
// nbig[kSmallTicket]: kg_ins_big_small's work ticket, nbig[kOverflow]: the
// closures it passed on to kg_ins_big.
constexpr uint32_t kBigClasses = 32;
constexpr uint32_t kSmallTicket = 2 + 2 * kBigClasses;
constexpr uint32_t kOverflow = kSmallTicket + 1;
constexpr uint32_t kNbigWords = kOverflow + 1;

__global__ void kg_zero(Counters* a, Counters* b, uint32_t n, uint32_t* nbig) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    a[i] = Counters{0, 0, 0};
    b[i] = Counters{0, 0, 0};
  }
  if (blockIdx.x == 0 && threadIdx.x < kNbigWords) nbig[threadIdx.x] = 0;  // blockDim >= kNbigWords
}

// Longest replays first.  The big queue fills in whatever order the leader
// lanes reach it, so a closure of a thousand members can start after most
// of the others have finished and alone keep kg_ins_big running.  These two
// passes sort the queue by member count in powers of two (largest class
// first; any order inside a class).  Closures own disjoint arcs, so the
// order of their replays does not change the result.
// Adds, for every active thread of the block, 1 to ctr[c] for its class c,
// with one global atomic per class present in the block (the classes are
// few, so per-thread or per-wave atomics would queue on a handful of
// addresses).  Returns the counter's old value plus the thread's rank among
// the block's threads of its class.  Every thread of the block calls it.
__device__ inline uint32_t block_class_add(uint32_t* ctr, uint32_t c, bool active, uint32_t* hist, uint32_t* gbase) {
  if (threadIdx.x < kBigClasses) hist[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t rank = active ? atomicAdd(&hist[c], 1u) : 0u;
  __syncthreads();
  if (threadIdx.x < kBigClasses && hist[threadIdx.x]) gbase[threadIdx.x] = atomicAdd(&ctr[threadIdx.x], hist[threadIdx.x]);
  __syncthreads();
  const uint32_t r = active ? gbase[c] + rank : 0u;
  __syncthreads();  // hist and gbase are reused by the next call
  return r;
}

// Pass 1: a closure's members are the dry-claimed slots (P & ~Q) of its
// arc, which runs from its first member s up to the next slot still free in
// Q.  Only the first kClassWords bitmap words from s are read (all at once):
// a closure whose arc runs past them is put in the top class.
constexpr uint32_t kClassWords = 16;

__global__ void __launch_bounds__(256)
kg_big_class(const Table* __restrict__ tables, const unsigned long long* __restrict__ Q,
             const unsigned long long* __restrict__ qoff, const unsigned long long* __restrict__ big, uint32_t* nbig,
             uint8_t* __restrict__ cls) {
  __shared__ uint32_t hist[kBigClasses], gbase[kBigClasses];
  const uint32_t nb = nbig[0];
  for (uint32_t base = blockIdx.x * blockDim.x; base < nb; base += gridDim.x * blockDim.x) {  // block-uniform
    const uint32_t c = base + threadIdx.x;
    const bool active = c < nb;
    uint32_t b = 0;
    if (active) {
      const uint32_t table = static_cast<uint32_t>(big[c] >> 32);
      const uint64_t s = big[c] & 0xFFFFFFFFull;
      const Table t = tables[table];
      const unsigned long long* Qt = Q + qoff[table];
      const uint64_t words = (t.cap + 63) >> 6;
      unsigned long long qw[kClassWords], mw[kClassWords];
#pragma unroll
      for (uint32_t j = 0; j < kClassWords; ++j) {
        const uint64_t w = ((s >> 6) + j) % words;
        qw[j] = Qt[w];
        mw[j] = t.freebits[w] & ~qw[j];
      }
      // Bits at or after s in the first word; the barrier is the first Q bit.
      // Tables of fewer than kClassWords words repeat words here, but the
      // barrier comes within the first lap, so nothing is counted twice.
      uint64_t members = 0;
      bool found = false;
#pragma unroll
      for (uint32_t j = 0; j < kClassWords; ++j) {  // no early exit: constant indices keep qw/mw in registers
        const unsigned long long from = j == 0 ? (~0ull << (s & 63)) : ~0ull;
        const unsigned long long q = found ? 0ull : (qw[j] & from);
        const unsigned long long upto = q ? ((q & (0ull - q)) - 1ull) : ~0ull;  // bits below the barrier
        members += found ? 0u : static_cast<uint32_t>(__popcll(mw[j] & from & upto));
        found = found || q != 0;
      }
      const uint32_t k = members ? 63u - static_cast<uint32_t>(__clzll(members)) : 0u;
      b = found ? (k < kBigClasses - 1 ? k : kBigClasses - 2) : kBigClasses - 1;
      cls[c] = static_cast<uint8_t>(b);
    }
    (void)block_class_add(nbig + 2, b, active, hist, gbase);
  }
}

// Pass 2: scatter the queue into `ordered`, largest class first.
__global__ void __launch_bounds__(256)
kg_big_order(const unsigned long long* __restrict__ big, const uint8_t* __restrict__ cls, uint32_t* nbig,
             unsigned long long* __restrict__ ordered) {
  __shared__ uint32_t hist[kBigClasses], gbase[kBigClasses];
  const uint32_t nb = nbig[0];
  for (uint32_t base = blockIdx.x * blockDim.x; base < nb; base += gridDim.x * blockDim.x) {  // block-uniform
    const uint32_t c = base + threadIdx.x;
    const bool active = c < nb;
    const uint32_t b = active ? cls[c] : 0u;
    const uint32_t at = block_class_add(nbig + 2 + kBigClasses, b, active, hist, gbase);
    if (active) {
      uint32_t before = 0;
      for (uint32_t h = b + 1; h < kBigClasses; ++h) before += nbig[2 + h];
      ordered[before + at] = big[c];
    }
  }
}

__global__ void kg_copy_words(unsigned long long* __restrict__ dst, const unsigned long long* __restrict__ src,
                              uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

template <int KS>
__global__ void __launch_bounds__(512)
kg_ins_dry(const Table* __restrict__ tables, OpDesc op, const uint8_t* __restrict__ keys,
           unsigned long long* __restrict__ Q, const unsigned long long* __restrict__ qoff,
           uint32_t* __restrict__ dslot, Counters* ctr) {
  const uint32_t nch = *op.n_chunks;
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const Chunk ch = op.chunks[c];
    const TableOp to = op.tops[ch.table];
    if (to.mode != kModeLeader || threadIdx.x >= ch.count) continue;
    const Table t = tables[ch.table];
    const uint32_t K = KS ? KS : t.K;
    const uint64_t g = ch.begin + threadIdx.x;
    Blob<KS> k;
    k.load(keys + g * K, K);
    const uint64_t h = k.hash(K, t.mode) & t.mask;
    uint64_t s;
    if (!bm_claim(Q + qoff[ch.table], h, t.cap, ((t.cap + 63) >> 6) + to.count + 2, &s)) {
      atomicOr(&ctr[ch.table].status, kStatusDryFail);
      dslot[g] = kNoOwner;
      continue;
    }
    dslot[g] = static_cast<uint32_t>(s);
    t.owner[s] = static_cast<uint32_t>(ch.local + threadIdx.x);
  }
}

// LDS of one wave of kg_ins_leader: with the plan (static sizes), the
// leader step's arrays (leader_wave) for closures of up to
// kGroupLeaderMax members; without it, the member lists only.
constexpr uint32_t kGroupLeaderMax = 8;
template <int KS, int VS, bool PLAN>
__host__ __device__ constexpr uint32_t group_leader_wave_bytes() {
  return PLAN ? 2 * 64 * kPlanMoves * 8 + 64 * 8 + 64 * 16 * (VS % 16 == 0 ? VS / 16 : 0) + kGroupLeaderMax * 64 * 4 + 64
              : kLeaderMax * 64 * 4;
}

// The leader pass over every chunk of a leader-mode table: each wave runs
// the single map's leader step (leader_wave) on its 64 keys of the chunk,
// with the table's claim bitmap; closures past its reach are queued for
// the big replays with their table.
template <int KS, int VS, bool PLAN>
__global__ void __launch_bounds__(512)
kg_ins_leader(const Table* __restrict__ tables, OpDesc op, const uint8_t* __restrict__ keys,
              const uint8_t* __restrict__ vals, const unsigned long long* __restrict__ Q,
              const unsigned long long* __restrict__ qoff, const uint32_t* __restrict__ dslot,
              unsigned long long* __restrict__ big, uint32_t* nbig, Counters* ctr, uint32_t coop_arc) {
  extern __shared__ u32x4 gl_lds[];  // per wave: group_leader_wave_bytes<KS, VS, PLAN>() bytes
  constexpr bool kPlan = PLAN && KS > 0 && KS % 4 == 0 && VS > 0 && VS % 4 == 0;
  constexpr uint32_t kWaveBytes = group_leader_wave_bytes<KS, VS, kPlan>();
  unsigned char* w = reinterpret_cast<unsigned char*>(gl_lds) + (threadIdx.x >> 6) * kWaveBytes;
  uint64_t* mv_src = reinterpret_cast<uint64_t*>(w);
  uint64_t* mv_dst = mv_src + (kPlan ? 64 * kPlanMoves : 0);
  uint64_t* arc_at = mv_dst + (kPlan ? 64 * kPlanMoves : 0);
  u32x4* mv_stage = reinterpret_cast<u32x4*>(arc_at + (kPlan ? 64 : 0));
  uint32_t* lst = reinterpret_cast<uint32_t*>(mv_stage + (kPlan ? 64 * (VS % 16 == 0 ? VS / 16 : 0) : 0));
  uint8_t* arc_src = reinterpret_cast<uint8_t*>(lst + 64 * (kPlan ? kGroupLeaderMax : kLeaderMax));
  const uint32_t leader_max = kPlan ? kGroupLeaderMax : kLeaderMax;
  const uint32_t nch = *op.n_chunks;
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const Chunk ch = op.chunks[c];
    const TableOp to = op.tops[ch.table];
    if (to.mode != kModeLeader || (ctr[ch.table].status & kStatusDryFail)) continue;  // block-uniform
    const Table t = tables[ch.table];
    const uint32_t K = KS ? KS : t.K;
    const uint32_t V = VS ? VS : t.V;
    const uint32_t sd = threadIdx.x < ch.count ? dslot[ch.begin + threadIdx.x] : kNoSlot;
    const unsigned long long claimed = leader_wave<KS, VS, kPlan>(
        t, keys + to.begin * K, vals + to.begin * V, ch.local + threadIdx.x, sd, Q + qoff[ch.table], &ctr[ch.table],
        coop_arc, leader_max, lst, mv_src, mv_dst, mv_stage, arc_at, arc_src, [&](uint64_t s) {
          big[atomicAdd(nbig, 1u)] = (static_cast<unsigned long long>(ch.table) << 32) | s;
        });
    block_add(&ctr[ch.table].count, claimed);  // chunk-uniform table
  }
}

// The closures of fewer than 2^kSmallClasses members whose arcs fit
// kSmallArcWords of LDS: one wave each, many workgroups per CU (the big
// replay's LDS leaves room for one).  They come from the back of the
// ordered queue (the small classes); a closure that does not fit is passed
// on to kg_ins_big through `overflow`.  No workgroup waits for another.
constexpr uint32_t kSmallClasses = 9;
constexpr uint32_t kSmallLst = 1u << kSmallClasses;
constexpr uint32_t kSmallArcWords = 4096;  // 16 KiB

__device__ inline uint32_t large_closures(const uint32_t* nbig) {
  uint32_t n = 0;
  for (uint32_t b = kSmallClasses; b < kBigClasses; ++b) n += nbig[2 + b];
  return n;
}

template <int KS, int VS>
__global__ void __launch_bounds__(64)
kg_ins_big_small(const Table* __restrict__ tables, OpDesc op, const uint8_t* __restrict__ keys,
                 const uint8_t* __restrict__ vals, const unsigned long long* __restrict__ Q,
                 const unsigned long long* __restrict__ qoff, const unsigned long long* __restrict__ ordered,
                 unsigned long long* __restrict__ overflow, uint32_t* nbig_p, Counters* ctr) {
  static_assert(KS > 0 && VS > 0 && KS % 4 == 0 && VS % 4 == 0, "static sizes");
  __shared__ uint32_t lst[kSmallLst];
  __shared__ uint32_t arc_lds[kSmallArcWords];
  __shared__ uint32_t cnt, cur;
  __shared__ unsigned long long barrier_at, barrier_before;
  const uint32_t nbig = nbig_p[0];
  const uint32_t n_large = large_closures(nbig_p);
  if (n_large + blockIdx.x >= nbig) return;
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(&nbig_p[kSmallTicket], 1u);
    __syncthreads();
    const uint32_t c = n_large + cur;
    __syncthreads();
    if (c >= nbig) break;
    const unsigned long long e = ordered[c];
    const uint32_t table = static_cast<uint32_t>(e >> 32);
    const uint64_t s = e & 0xFFFFFFFFull;
    const Table t = tables[table];
    const TableOp to = op.tops[table];
    const unsigned long long* Qt = Q + qoff[table];
    const unsigned long long* P = t.freebits;
    const uint64_t words = (t.cap + 63) >> 6;
    if (threadIdx.x == 0) {
      barrier_at = bm_next_set(Qt, s, t.cap);
      barrier_before = bm_prev_set(P, s, t.cap);
      cnt = 0;
    }
    __syncthreads();
    const uint64_t arc = (barrier_at + t.cap - s) & t.mask;
    const uint64_t b = barrier_before;
    const uint64_t L = (barrier_at - b - 1) & t.mask;
    bool fits = b != s && (2 + KS / 4) * L <= kSmallArcWords;
    if (fits) {
      const uint64_t w0 = s >> 6;
      const uint64_t nwords = min(((s & 63) + arc + 63) >> 6, words);
      for (uint64_t wi = threadIdx.x; wi < nwords; wi += blockDim.x) {
        const uint64_t w = (w0 + wi) % words;
        unsigned long long bits = P[w] & ~Qt[w];
        while (bits) {
          const int bt = __ffsll(bits) - 1;
          bits &= bits - 1;
          const uint64_t p = (w << 6) + bt;
          if (((p + t.cap - s) & t.mask) >= arc) continue;
          const uint32_t k = atomicAdd(&cnt, 1u);
          if (k < kSmallLst) lst[k] = t.owner[p];
        }
      }
      __syncthreads();
      fits = cnt <= kSmallLst;
    }
    if (!fits) {
      if (threadIdx.x == 0) overflow[atomicAdd(&nbig_p[kOverflow], 1u)] = e;
      __syncthreads();
      continue;
    }
    const uint32_t m = cnt;
    uint32_t p2 = 1;
    while (p2 < m) p2 <<= 1;
    for (uint32_t k = m + threadIdx.x; k < p2; k += blockDim.x) lst[k] = 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t size = 2; size <= p2; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t k = threadIdx.x; k < p2; k += blockDim.x) {
          const uint32_t j = k ^ stride;
          if (j > k) {
            const uint32_t a = lst[k], b2 = lst[j];
            if ((a > b2) == ((k & size) == 0)) {
              lst[k] = b2;
              lst[j] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    const uint32_t claimed = replay_arc_lds<KS, VS>(t, keys + to.begin * KS, vals + to.begin * VS, (b + 1) & t.mask,
                                                    static_cast<uint32_t>(L), lst, m, arc_lds, &ctr[table]);
    if (threadIdx.x == 0 && claimed) atomicAdd(&ctr[table].count, static_cast<unsigned long long>(claimed));
    __syncthreads();
  }
}

template <int KS, int VS>
__global__ void __launch_bounds__(64)
kg_ins_big(const Table* __restrict__ tables, OpDesc op, const uint8_t* __restrict__ keys,
           const uint8_t* __restrict__ vals, const unsigned long long* __restrict__ Q,
           const unsigned long long* __restrict__ qoff, const unsigned long long* __restrict__ ordered,
           const unsigned long long* __restrict__ overflow, uint32_t* nbig_p, Counters* ctr) {
  // static sizes: the member list, then the LDS replays' arrays
  constexpr uint32_t kPool = (KS > 0 && VS > 0 && KS % 4 == 0 && VS % 4 == 0) ? kBigPoolWords : 0u;
  __shared__ uint32_t lst[kPool ? kPool : kBigMax];
  __shared__ uint32_t carry[2 * kMaxGenericSize / 4];
  __shared__ uint32_t cnt;
  __shared__ unsigned long long barrier_at, barrier_before;
  __shared__ uint32_t cur;
  // The closures of the large classes (the front of the ordered queue),
  // then those kg_ins_big_small passed on (its launch has ended).
  const uint32_t n_large = kPool ? large_closures(nbig_p) : nbig_p[0];  // (no small replay for runtime sizes)
  const uint32_t nbig = n_large + nbig_p[kOverflow];
  // Closures are handed out one at a time (their replay times differ by
  // orders of magnitude, so a fixed assignment would leave workgroups idle).
  // Workgroups beyond the queue's length leave at once.
  if (blockIdx.x >= nbig) return;
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(nbig_p + 1, 1u);
    __syncthreads();
    const uint32_t c = cur;
    __syncthreads();
    if (c >= nbig) break;
    const unsigned long long e = c < n_large ? ordered[c] : overflow[c - n_large];
    const uint32_t table = static_cast<uint32_t>(e >> 32);
    const uint64_t s = e & 0xFFFFFFFFull;
    const Table t = tables[table];
    const TableOp to = op.tops[table];
    const uint32_t K = KS ? KS : t.K;
    const uint32_t V = VS ? VS : t.V;
    const unsigned long long* Qt = Q + qoff[table];
    const unsigned long long* P = t.freebits;
    const uint64_t words = (t.cap + 63) >> 6;
    if (threadIdx.x == 0) {
      barrier_at = bm_next_set(Qt, s, t.cap);
      barrier_before = kPool ? bm_prev_set(P, s, t.cap) : s;
    }
    __syncthreads();
    const uint64_t arc = (barrier_at + t.cap - s) & t.mask;
    const uint64_t w0 = s >> 6;
    const uint64_t nwords = min(((s & 63) + arc + 63) >> 6, words);
    // the whole arc, from the barrier before s; a canonical replay takes
    // all of its members at once
    const uint64_t b = barrier_before;
    const uint64_t L = (barrier_at - b - 1) & t.mask;
    const bool canon = kPool && t.canon_min && b != s && L <= kCanonMaxL && 2 * L <= kPool;
    const uint32_t lcap = canon ? static_cast<uint32_t>(L) : kBigMax;
    uint64_t lo = 0, hi = ~0ull;
    bool windowed = false;
    while (lo < to.count) {
      if (threadIdx.x == 0) cnt = 0;
      __syncthreads();
      for (uint64_t wi = threadIdx.x; wi < nwords; wi += blockDim.x) {
        const uint64_t w = (w0 + wi) % words;
        unsigned long long bits = P[w] & ~Qt[w];
        while (bits) {
          const int b = __ffsll(bits) - 1;
          bits &= bits - 1;
          const uint64_t p = (w << 6) + b;
          if (((p + t.cap - s) & t.mask) >= arc) continue;
          const uint32_t idx = t.owner[p];
          if (idx < lo || idx >= hi) continue;
          const uint32_t k = atomicAdd(&cnt, 1u);
          if (k < lcap) lst[k] = idx;
        }
      }
      __syncthreads();
      const uint32_t m = cnt;
      if (m > lcap) {
        windowed = true;
        hi = lo + lcap;
        __syncthreads();
        continue;
      }
      uint32_t p2 = 1;
      while (p2 < m) p2 <<= 1;
      if (kPool && m > 64 && 2 * ((m + 3) & ~3u) + p2 <= kPool) {
        sort_by_buckets(lst, m, to.count, lst + ((m + 3) & ~3u), lst + 2 * ((m + 3) & ~3u), p2);
        p2 = 0;  // (sorted)
      }
      for (uint32_t k = m + threadIdx.x; k < p2; k += blockDim.x) lst[k] = 0xFFFFFFFFu;
      __syncthreads();
      for (uint32_t size = 2; size <= p2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
          for (uint32_t k = threadIdx.x; k < p2; k += blockDim.x) {
            const uint32_t j = k ^ stride;
            if (j > k) {
              const uint32_t a = lst[k], b2 = lst[j];
              if ((a > b2) == ((k & size) == 0)) {
                lst[k] = b2;
                lst[j] = a;
              }
            }
          }
          __syncthreads();
        }
      }
      if constexpr (KS > 0 && VS > 0 && KS % 4 == 0 && VS % 4 == 0) {
        // A canonical arc from its groups; else the whole arc (from the
        // barrier before s to the one after) in LDS when it fits; otherwise
        // one wave walks each member in HBM.
        uint32_t claimed = 0;
        bool done = false;
        if (canon && !windowed && m >= t.canon_min && canon_words(m, static_cast<uint32_t>(L)) <= kPool) {
          const uint32_t c = replay_arc_canon<KS, VS>(t, keys + to.begin * K, vals + to.begin * V, (b + 1) & t.mask,
                                                      static_cast<uint32_t>(L), lst, m, lst);
          // owner[] is the removes' conflict scratch here: every entry of the
          // arc back to "none"
          __threadfence_block();
          for (uint64_t q = threadIdx.x; q < L; q += blockDim.x) t.owner[(b + 1 + q) & t.mask] = kNoOwner;
          __syncthreads();
          if (c != kNotCanon) {
            claimed = c;
            done = true;
          }
        }
        const uint32_t arc_off = max(kBigMax, (m + 3u) & ~3u);
        if (done) {
        } else if (!windowed && b != s && L <= arc_lds_cap<KS>() && arc_off + kArcLdsWords <= kPool) {
          claimed = replay_arc_lds<KS, VS>(t, keys + to.begin * K, vals + to.begin * V, (b + 1) & t.mask,
                                           static_cast<uint32_t>(L), lst, m, lst + arc_off, &ctr[table]);
        } else {
          for (uint32_t j = 0; j < m; ++j) {  // one wave per workgroup walks each member
            const uint64_t gk = to.begin + lst[j];
            claimed += insert_one_wave<KS, VS>(t, keys + gk * K, vals + gk * V, carry, &ctr[table]);
          }
        }
        if (threadIdx.x == 0 && claimed) atomicAdd(&ctr[table].count, static_cast<unsigned long long>(claimed));
      } else if (threadIdx.x == 0) {
        uint32_t claimed = 0;
        for (uint32_t j = 0; j < m; ++j) {
          const uint64_t gk = to.begin + lst[j];
          claimed += insert_one<KS, VS, false>(t, keys + gk * K, vals + gk * V, K, V, &ctr[table]);
        }
        if (claimed) atomicAdd(&ctr[table].count, static_cast<unsigned long long>(claimed));
      }
      __syncthreads();
      if (!windowed) break;
      lo = hi;
      hi = lo + kBigMax;
    }
  }
}

// Tables whose batch may fill them: one lane applies the batch in order.
template <int KS, int VS>
__global__ void kg_ins_serial(const Table* __restrict__ tables, const TableOp* __restrict__ tops, uint32_t n_tables,
                              const uint8_t* __restrict__ keys, const uint8_t* __restrict__ vals, Counters* ctr) {
  for (uint32_t tb = blockIdx.x * blockDim.x + threadIdx.x; tb < n_tables; tb += gridDim.x * blockDim.x) {
    const TableOp to = tops[tb];
    if (to.mode != kModeSerial) continue;
    const Table t = tables[tb];
    const uint32_t K = KS ? KS : t.K;
    const uint32_t V = VS ? VS : t.V;
    unsigned long long claimed = 0;
    for (uint64_t i = 0; i < to.count; ++i) {
      const uint64_t g = to.begin + i;
      claimed += insert_one<KS, VS, true>(t, keys + g * K, vals + g * V, K, V, &ctr[tb]);
    }
    ctr[tb].count += claimed;
  }
}

__global__ void __launch_bounds__(512)
kg_ins_bits(const Table* __restrict__ tables, OpDesc op, const uint32_t* __restrict__ dslot, const Counters* ctr) {
  const uint32_t nch = *op.n_chunks;
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const Chunk ch = op.chunks[c];
    if (op.tops[ch.table].mode != kModeLeader || threadIdx.x >= ch.count) continue;
    const uint64_t s = dslot[ch.begin + threadIdx.x];
    if (s == kNoOwner) continue;
    const Table t = tables[ch.table];
    if (!(ctr[ch.table].status & kStatusDryFail) && (t.flags[s] & 0xFu) == kOccupied) {
      const unsigned long long bit = 1ull << (s & 63);
      atomicAnd(&t.freebits[s >> 6], ~bit);
      atomicAnd(&t.emptybits[s >> 6], ~bit);
    }
    t.owner[s] = kNoOwner;
  }
}

template <int KS, int VS, int G>
__global__ void __launch_bounds__(512)
kg_get(const Table* __restrict__ tables, OpDesc op, const uint8_t* __restrict__ keys, uint8_t* __restrict__ out_vals,
       uint8_t* __restrict__ out_found) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t gl = lane & (G - 1);
  const uint32_t gshift = lane & ~(G - 1u);
  const uint32_t per_pass = blockDim.x / G;
  const uint32_t nch = *op.n_chunks;
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const Chunk ch = op.chunks[c];
    const Table t = tables[ch.table];
    const uint32_t K = KS ? KS : t.K;
    const uint32_t V = VS ? VS : t.V;
    for (uint32_t k0 = 0; k0 < ch.count; k0 += per_pass) {
      const uint32_t ki = k0 + threadIdx.x / G;
      if (ki >= ch.count) continue;
      const uint64_t g = ch.begin + ki;
      Blob<KS> q;
      q.load(keys + g * K, K);
      const uint64_t home = q.hash(K, t.mode) & t.mask;
      uint32_t kind;
      uint64_t slot;
      group_probe<KS, G>(t, q, K, home, gl, gshift, &kind, &slot);
      const bool hit = kind == kEvMatch;
      group_copy<VS, G>(out_vals + g * V, hit ? t.vals + slot * V : nullptr, V, gl);
      if (out_found && gl == 0) out_found[g] = hit ? 1 : 0;
    }
  }
}

template <int KS, int G>
__global__ void __launch_bounds__(512)
kg_rem_probe(const Table* __restrict__ tables, OpDesc op, const uint8_t* __restrict__ keys,
             unsigned long long* __restrict__ ev) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t gl = lane & (G - 1);
  const uint32_t gshift = lane & ~(G - 1u);
  const uint32_t per_pass = blockDim.x / G;
  const uint32_t nch = *op.n_chunks;
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const Chunk ch = op.chunks[c];
    const Table t = tables[ch.table];
    const uint32_t K = KS ? KS : t.K;
    for (uint32_t k0 = 0; k0 < ch.count; k0 += per_pass) {
      const uint32_t ki = k0 + threadIdx.x / G;
      if (ki >= ch.count) continue;
      const uint64_t g = ch.begin + ki;
      Blob<KS> q;
      q.load(keys + g * K, K);
      const uint64_t home = q.hash(K, t.mode) & t.mask;
      uint32_t kind;
      uint64_t slot;
      group_probe<KS, G>(t, q, K, home, gl, gshift, &kind, &slot);
      if (gl == 0) {
        ev[g] = (static_cast<unsigned long long>(kind) << 32) | slot;
        if (kind == kEvMatch) atomicMin(&t.owner[slot], static_cast<uint32_t>(ch.local + ki));
      }
    }
  }
}

__global__ void __launch_bounds__(512)
kg_rem_check(const Table* __restrict__ tables, OpDesc op, const unsigned long long* __restrict__ ev,
             uint8_t* __restrict__ action, Counters* ctr) {
  const uint32_t nch = *op.n_chunks;
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const Chunk ch = op.chunks[c];
    if (threadIdx.x >= ch.count) continue;
    const Table t = tables[ch.table];
    const uint64_t g = ch.begin + threadIdx.x;
    const uint32_t local = static_cast<uint32_t>(ch.local + threadIdx.x);
    const uint32_t kind = static_cast<uint32_t>(ev[g] >> 32);
    const uint64_t slot = ev[g] & 0xFFFFFFFFull;
    uint8_t a = 0;
    bool conflict = false;
    if (kind == kEvMatch) {
      a = t.owner[slot] == local ? 1 : 0;
      conflict = !a;
    } else if (kind == kEvTermOcc) {
      conflict = t.owner[slot] < local;
    }
    action[g] = a;
    if (conflict) atomicOr(&ctr[ch.table].status, kStatusConflict);
  }
}

__global__ void __launch_bounds__(512)
kg_rem_apply(const Table* __restrict__ tables, OpDesc op, const unsigned long long* __restrict__ ev,
             const uint8_t* __restrict__ action, Counters* ctr) {
  const uint32_t nch = *op.n_chunks;
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const Chunk ch = op.chunks[c];
    const Table t = tables[ch.table];
    const bool conflict = (ctr[ch.table].status & kStatusConflict) != 0;
    uint32_t removed = 0;
    if (threadIdx.x < ch.count) {
      const uint64_t g = ch.begin + threadIdx.x;
      if (static_cast<uint32_t>(ev[g] >> 32) == kEvMatch) {
        const uint64_t slot = ev[g] & 0xFFFFFFFFull;
        if (!conflict && action[g]) {
          t.flags[slot] = kTombstone;
          atomicOr(&t.freebits[slot >> 6], 1ull << (slot & 63));
          removed = 1;
        }
        t.owner[slot] = kNoOwner;
      }
    }
    block_add(&ctr[ch.table].count, removed);  // chunk-uniform table
  }
}

// Tables whose remove batch had a conflict: one lane replays it in order
// from the untouched pre-batch state.
template <int KS>
__global__ void kg_rem_serial(const Table* __restrict__ tables, const TableOp* __restrict__ tops, uint32_t n_tables,
                              const uint8_t* __restrict__ keys, Counters* ctr) {
  for (uint32_t tb = blockIdx.x * blockDim.x + threadIdx.x; tb < n_tables; tb += gridDim.x * blockDim.x) {
    if (!(ctr[tb].status & kStatusConflict)) continue;
    const TableOp to = tops[tb];
    const Table t = tables[tb];
    const uint32_t K = KS ? KS : t.K;
    unsigned long long removed = 0;
    for (uint64_t i = 0; i < to.count; ++i) removed += remove_one<KS>(t, keys + (to.begin + i) * K, K);
    ctr[tb].count = removed;
  }
}

}  // namespace dev

namespace {

using dev::Chunk;
using dev::Counters;
using dev::OpDesc;
using dev::Table;
using dev::TableOp;

inline uint32_t gridn(uint64_t threads, uint32_t block) {
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((threads + block - 1) / block, 1u << 16)));
}

// Launch sequences for one (K,V) specialisation.
struct GroupOps {
  void (*insert)(const Table*, OpDesc, const TableOp*, uint32_t, const uint8_t*, const uint8_t*,
                 unsigned long long*, const unsigned long long*, const unsigned long long*, uint64_t, uint32_t*,
                 unsigned long long*, uint64_t, uint32_t*, Counters*, uint32_t, uint32_t, uint32_t, hipStream_t);
  void (*get)(const Table*, OpDesc, const uint8_t*, uint8_t*, uint8_t*, uint32_t, uint32_t, uint32_t, hipStream_t);
  void (*remove)(const Table*, OpDesc, const TableOp*, uint32_t, const uint8_t*, unsigned long long*, uint8_t*,
                 Counters*, uint32_t, uint32_t, uint32_t, hipStream_t);
};

template <int KS, int VS>
void g_insert(const Table* tables, OpDesc op, const TableOp* tops, uint32_t n_tables, const uint8_t* keys,
              const uint8_t* vals, unsigned long long* Q, const unsigned long long* freebits_all,
              const unsigned long long* qoff, uint64_t total_words, uint32_t* dslot, unsigned long long* big,
              uint64_t big_cap, uint32_t* nbig, Counters* ctr, uint32_t block, uint32_t grid, uint32_t coop_arc,
              hipStream_t s) {
  // big: [queue | ordered queue] of big_cap entries each, then big_cap class bytes
  unsigned long long* ordered = big + big_cap;
  uint8_t* cls = reinterpret_cast<uint8_t*>(big + 2 * big_cap);
  hipLaunchKernelGGL(dev::kg_copy_words, dim3(gridn(total_words, 256)), dim3(256), 0, s, Q, freebits_all, total_words);
  hipLaunchKernelGGL((dev::kg_ins_dry<KS>), dim3(grid), dim3(block), 0, s, tables, op, keys, Q, qoff, dslot, ctr);
  // the planned leader step when its LDS fits 64 KiB per workgroup
  constexpr uint32_t kPlanBytes = dev::group_leader_wave_bytes<KS, VS, true>();
  constexpr uint32_t kListBytes = dev::group_leader_wave_bytes<KS, VS, false>();
  const uint32_t plan_lds = (block / 64) * kPlanBytes, list_lds = (block / 64) * kListBytes;
  if (KS > 0 && VS > 0 && KS % 4 == 0 && VS % 4 == 0 && plan_lds <= 65536)
    hipLaunchKernelGGL((dev::kg_ins_leader<KS, VS, true>), dim3(grid), dim3(block), plan_lds, s, tables, op, keys, vals,
                       Q, qoff, dslot, big, nbig, ctr, coop_arc);
  else
    hipLaunchKernelGGL((dev::kg_ins_leader<KS, VS, false>), dim3(grid), dim3(block), list_lds, s, tables, op, keys,
                       vals, Q, qoff, dslot, big, nbig, ctr, coop_arc);
  hipLaunchKernelGGL(dev::kg_big_class, dim3(256), dim3(256), 0, s, tables, Q, qoff, big, nbig, cls);
  hipLaunchKernelGGL(dev::kg_big_order, dim3(256), dim3(256), 0, s, big, cls, nbig, ordered);
  if constexpr (KS > 0 && VS > 0 && KS % 4 == 0 && VS % 4 == 0)
    hipLaunchKernelGGL((dev::kg_ins_big_small<KS, VS>), dim3(2048), dim3(64), 0, s, tables, op, keys, vals, Q, qoff,
                       ordered, big, nbig, ctr);
  hipLaunchKernelGGL((dev::kg_ins_big<KS, VS>), dim3(1024), dim3(64), 0, s, tables, op, keys, vals, Q, qoff, ordered,
                     big, nbig, ctr);
  hipLaunchKernelGGL((dev::kg_ins_serial<KS, VS>), dim3(gridn(n_tables, 64)), dim3(64), 0, s, tables, tops, n_tables,
                     keys, vals, ctr);
  hipLaunchKernelGGL(dev::kg_ins_bits, dim3(grid), dim3(block), 0, s, tables, op, dslot, ctr);
}

template <int KS, int VS, int G>
void g_get_g(const Table* tables, OpDesc op, const uint8_t* keys, uint8_t* ov, uint8_t* of, uint32_t block,
             uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((dev::kg_get<KS, VS, G>), dim3(grid), dim3(block), 0, s, tables, op, keys, ov, of);
}

template <int KS, int VS>
void g_get(const Table* tables, OpDesc op, const uint8_t* keys, uint8_t* ov, uint8_t* of, uint32_t G, uint32_t block,
           uint32_t grid, hipStream_t s) {
  switch (G) {
    case 1: return g_get_g<KS, VS, 1>(tables, op, keys, ov, of, block, grid, s);
    case 2: return g_get_g<KS, VS, 2>(tables, op, keys, ov, of, block, grid, s);
    case 4: return g_get_g<KS, VS, 4>(tables, op, keys, ov, of, block, grid, s);
    case 8: return g_get_g<KS, VS, 8>(tables, op, keys, ov, of, block, grid, s);
    default: return g_get_g<KS, VS, 16>(tables, op, keys, ov, of, block, grid, s);
  }
}

template <int KS, int G>
void g_rem_probe(const Table* tables, OpDesc op, const uint8_t* keys, unsigned long long* ev, uint32_t block,
                 uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((dev::kg_rem_probe<KS, G>), dim3(grid), dim3(block), 0, s, tables, op, keys, ev);
}

template <int KS, int VS>
void g_remove(const Table* tables, OpDesc op, const TableOp* tops, uint32_t n_tables, const uint8_t* keys,
              unsigned long long* ev, uint8_t* action, Counters* ctr, uint32_t G, uint32_t block, uint32_t grid,
              hipStream_t s) {
  switch (G) {
    case 1: g_rem_probe<KS, 1>(tables, op, keys, ev, block, grid, s); break;
    case 2: g_rem_probe<KS, 2>(tables, op, keys, ev, block, grid, s); break;
    case 4: g_rem_probe<KS, 4>(tables, op, keys, ev, block, grid, s); break;
    case 8: g_rem_probe<KS, 8>(tables, op, keys, ev, block, grid, s); break;
    default: g_rem_probe<KS, 16>(tables, op, keys, ev, block, grid, s); break;
  }
  hipLaunchKernelGGL(dev::kg_rem_check, dim3(grid), dim3(block), 0, s, tables, op, ev, action, ctr);
  hipLaunchKernelGGL(dev::kg_rem_apply, dim3(grid), dim3(block), 0, s, tables, op, ev, action, ctr);
  hipLaunchKernelGGL((dev::kg_rem_serial<KS>), dim3(gridn(n_tables, 64)), dim3(64), 0, s, tables, tops, n_tables, keys,
                     ctr);
}

template <int KS, int VS>
GroupOps make_group_ops() {
  return GroupOps{&g_insert<KS, VS>, &g_get<KS, VS>, &g_remove<KS, VS>};
}

GroupOps select_group_ops(uint32_t K, uint32_t V) {
  if (K == 8 && V == 32) return make_group_ops<8, 32>();
  if (K == 8 && V == 8) return make_group_ops<8, 8>();
  if (K == 32 && V == 8) return make_group_ops<32, 8>();
  if (K == 32 && V == 128) return make_group_ops<32, 128>();
  return make_group_ops<0, 0>();
}

std::string hmsg(const char* what, hipError_t e) { return std::string(what) + ": " + hipGetErrorString(e); }

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Host mirror of one op's descriptors, uploaded before each launch.
struct HostOp {
  std::vector<TableOp> tops;
  std::vector<Chunk> chunks;
};

}  // namespace

class GroupImpl {
 public:
  ~GroupImpl() { release(); }

  fm_status init(uint32_t n, const uint64_t* caps, uint32_t K, uint32_t V, uint32_t mode, int dev, std::string* err);
  fm_status insert(const uint8_t* keys, const uint8_t* vals, const uint64_t* counts, uint64_t* out_new,
                   int32_t* out_status);
  fm_status get(const uint8_t* keys, const uint64_t* counts, uint8_t* ov, uint8_t* of);
  fm_status remove(const uint8_t* keys, const uint64_t* counts, uint64_t* out_removed);
  fm_status step(const uint8_t* ik, const uint8_t* iv, const uint64_t* ic, const uint8_t* gk, const uint64_t* gc,
                 uint8_t* ov, uint8_t* of, const uint8_t* rk, const uint64_t* rc, uint64_t* out_new,
                 int32_t* out_status, uint64_t* out_removed);
  fm_status export_slots(uint32_t table, uint32_t* flags, void* keys, void* values);
  fm_status copy_from(GroupImpl& src);
  fm_status set_tuning(const fm_group_tuning& t);
  fm_status autotune(GroupImpl* scratch, const uint8_t* ik, const uint8_t* iv, const uint64_t* ic, const uint8_t* gk,
                     const uint64_t* gc, uint8_t* ov, uint8_t* of, const uint8_t* rk, const uint64_t* rc,
                     const fm_group_tuning* cand, uint32_t n_cand, uint32_t reps, fm_group_tuning* best,
                     double* best_us, uint32_t* evaluated);

  uint32_t n_ = 0;
  int dev_ = -1;
  hipStream_t stream_ = nullptr;
  std::vector<Table> h_tables_;
  std::vector<uint64_t> len_;
  fm_group_tuning tune_{4, 256, 1024, 64};

 private:
  fm_status fail(fm_status c, const std::string& m) {
    set_error(c, m);
    return c;
  }
  fm_status use() {
    hipError_t e = hipSetDevice(dev_);
    return e == hipSuccess ? FM_OK : fail(FM_ERR_DEVICE_INIT, hmsg("hipSetDevice", e));
  }
  void build_op(HostOp& op, const uint64_t* counts, bool insert_modes);
  fm_status stage(int which, const HostOp& op, uint64_t total);
  fm_status ensure_keys(uint64_t total);
  fm_status enqueue_insert(const uint8_t* keys, const uint8_t* vals);
  fm_status enqueue_get(const uint8_t* keys, uint8_t* ov, uint8_t* of);
  fm_status enqueue_remove(const uint8_t* keys);
  fm_status finish(bool ins, bool rem, uint64_t* out_new, int32_t* out_status, uint64_t* out_removed,
                   const uint64_t* ins_counts);
  void release();

  uint32_t K_ = 0, V_ = 0, mode_ = 0;
  GroupOps ops_{};
  uint8_t* keys_all_ = nullptr;
  uint8_t* vals_all_ = nullptr;
  uint32_t* flags_all_ = nullptr;
  unsigned long long* free_all_ = nullptr;
  unsigned long long* empty_all_ = nullptr;
  uint32_t* owner_all_ = nullptr;
  uint64_t total_slots_ = 0, total_words_ = 0;
  Table* d_tables_ = nullptr;
  unsigned long long* d_qoff_ = nullptr;
  unsigned long long* Q_ = nullptr;

  // per-op descriptors: [0] insert, [1] get, [2] remove
  TableOp* d_tops_[3] = {};
  Chunk* d_chunks_[3] = {};
  uint32_t* d_nch_ = nullptr;   // 3 counts
  uint64_t chunk_cap_[3] = {};
  HostOp hop_[3];
  uint32_t h_nch_[3] = {};

  // per-key scratch
  uint64_t key_cap_ = 0;
  uint32_t* dslot_ = nullptr;
  unsigned long long* ev_ = nullptr;
  uint8_t* action_ = nullptr;
  unsigned long long* big_ = nullptr;
  uint32_t* nbig_ = nullptr;
  Counters* ins_ctr_ = nullptr;
  Counters* rem_ctr_ = nullptr;
  Counters* h_ins_ctr_ = nullptr;
  Counters* h_rem_ctr_ = nullptr;

  // captured step graph
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  std::vector<const void*> graph_key_;
};

void GroupImpl::release() {
  if (dev_ < 0) return;
  (void)hipSetDevice(dev_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
  dfree(keys_all_);
  dfree(vals_all_);
  dfree(flags_all_);
  dfree(free_all_);
  dfree(empty_all_);
  dfree(owner_all_);
  dfree(d_tables_);
  dfree(d_qoff_);
  dfree(Q_);
  for (int i = 0; i < 3; ++i) {
    dfree(d_tops_[i]);
    dfree(d_chunks_[i]);
  }
  dfree(d_nch_);
  dfree(dslot_);
  dfree(ev_);
  dfree(action_);
  dfree(big_);
  dfree(nbig_);
  dfree(ins_ctr_);
  dfree(rem_ctr_);
  if (h_ins_ctr_) (void)hipHostFree(h_ins_ctr_);
  if (h_rem_ctr_) (void)hipHostFree(h_rem_ctr_);
  h_ins_ctr_ = h_rem_ctr_ = nullptr;
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
  dev_ = -1;
}

fm_status GroupImpl::init(uint32_t n, const uint64_t* caps, uint32_t K, uint32_t V, uint32_t mode, int dev,
                          std::string* err) {
  if (n == 0 || !caps || K == 0 || V == 0 || K > dev::kMaxGenericSize || V > dev::kMaxGenericSize) {
    *err = "fm_group_create: need n > 0 tables and key/value sizes in 1..256";
    return FM_ERR_INVALID_ARGUMENT;
  }
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || dev < 0 || dev >= count) {
    *err = e != hipSuccess ? hmsg("hipGetDeviceCount", e) : std::string("device ordinal out of range");
    return FM_ERR_DEVICE_INIT;
  }
  dev_ = dev;
  if ((e = hipSetDevice(dev_)) != hipSuccess || (e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking)) != hipSuccess) {
    *err = hmsg("device init", e);
    return FM_ERR_DEVICE_INIT;
  }
  n_ = n;
  K_ = K;
  V_ = V;
  mode_ = mode == FM_HASH_MURMUR3 ? 1u : 0u;
  ops_ = select_group_ops(K, V);
  std::vector<uint64_t> cap(n), slot_off(n), word_off(n);
  for (uint32_t i = 0; i < n; ++i) {
    if (caps[i] == 0) {
      *err = "capacity must be positive";
      return FM_ERR_ZERO_CAPACITY;
    }
    cap[i] = round_capacity(caps[i]);
    if (cap[i] > (1ull << 31)) {
      *err = "capacity above 2^31 slots";
      return FM_ERR_INVALID_ARGUMENT;
    }
    slot_off[i] = total_slots_;
    word_off[i] = total_words_;
    total_slots_ += cap[i];
    total_words_ += (cap[i] + 63) / 64;
  }
  if ((e = hipMalloc(&keys_all_, total_slots_ * K + 16)) != hipSuccess ||
      (e = hipMalloc(&vals_all_, total_slots_ * V + 16)) != hipSuccess ||
      (e = hipMalloc(&flags_all_, total_slots_ * 4)) != hipSuccess ||
      (e = hipMalloc(&free_all_, total_words_ * 8)) != hipSuccess ||
      (e = hipMalloc(&empty_all_, total_words_ * 8)) != hipSuccess ||
      (e = hipMalloc(&owner_all_, total_slots_ * 4)) != hipSuccess ||
      (e = hipMalloc(&Q_, total_words_ * 8)) != hipSuccess || (e = hipMalloc(&d_tables_, n * sizeof(Table))) != hipSuccess ||
      (e = hipMalloc(&d_qoff_, n * 8)) != hipSuccess || (e = hipMalloc(&d_nch_, 3 * 4)) != hipSuccess ||
      (e = hipMalloc(&nbig_, dev::kNbigWords * 4)) != hipSuccess || (e = hipMalloc(&ins_ctr_, n * sizeof(Counters))) != hipSuccess ||
      (e = hipMalloc(&rem_ctr_, n * sizeof(Counters))) != hipSuccess ||
      (e = hipHostMalloc(&h_ins_ctr_, n * sizeof(Counters), hipHostMallocDefault)) != hipSuccess ||
      (e = hipHostMalloc(&h_rem_ctr_, n * sizeof(Counters), hipHostMallocDefault)) != hipSuccess) {
    *err = hmsg("hipMalloc (group)", e);
    return FM_ERR_GPU_ALLOC;
  }
  h_tables_.resize(n);
  // Arcs of at least this many members replayed from their groups when
  // canonical (0: never).
  uint32_t canon_min = 32;
  if (const char* env = std::getenv("FM_INSERT_CANON_MIN")) canon_min = static_cast<uint32_t>(std::strtoul(env, nullptr, 10));
  for (uint32_t i = 0; i < n; ++i) {
    Table& t = h_tables_[i];
    t.keys = keys_all_ + slot_off[i] * K;
    t.vals = vals_all_ + slot_off[i] * V;
    t.flags = flags_all_ + slot_off[i];
    t.freebits = free_all_ + word_off[i];
    t.emptybits = empty_all_ + word_off[i];
    t.owner = owner_all_ + slot_off[i];
    t.cap = cap[i];
    t.mask = cap[i] - 1;
    t.K = K;
    t.V = V;
    t.mode = mode_;
    t.canon_min = canon_min;
    hipLaunchKernelGGL(dev::k_clear, dim3(gridn(t.cap, 256)), dim3(256), 0, stream_, t, nullptr);
  }
  hipLaunchKernelGGL(dev::k_fill_u32, dim3(gridn(total_slots_, 256)), dim3(256), 0, stream_, owner_all_, total_slots_,
                     dev::kNoOwner);
  std::vector<unsigned long long> qoff(word_off.begin(), word_off.end());
  if ((e = hipMemcpyAsync(d_tables_, h_tables_.data(), n * sizeof(Table), hipMemcpyHostToDevice, stream_)) != hipSuccess ||
      (e = hipMemcpyAsync(d_qoff_, qoff.data(), n * 8, hipMemcpyHostToDevice, stream_)) != hipSuccess ||
      (e = hipStreamSynchronize(stream_)) != hipSuccess) {
    *err = hmsg("group init", e);
    return FM_ERR_DEVICE_INIT;
  }
  len_.assign(n, 0);
  return FM_OK;
}

fm_status GroupImpl::set_tuning(const fm_group_tuning& t) {
  const bool ok = (t.group == 1 || t.group == 2 || t.group == 4 || t.group == 8 || t.group == 16) && t.block >= 64 &&
                  t.block <= 512 && t.block % 64 == 0 && t.grid >= 64 && t.grid <= 8192;
  if (!ok) return fail(FM_ERR_INVALID_ARGUMENT, "group tuning: group 1..16 (pow2), block 64..512 (x64), grid 64..8192");
  tune_ = t;
  return FM_OK;
}

// Per-table op descriptors and the chunk list (one table per chunk).
void GroupImpl::build_op(HostOp& op, const uint64_t* counts, bool insert_modes) {
  op.tops.resize(n_);
  op.chunks.clear();
  uint64_t begin = 0;
  for (uint32_t t = 0; t < n_; ++t) {
    TableOp& to = op.tops[t];
    to.begin = begin;
    to.count = counts[t];
    to.pad = 0;
    to.mode = dev::kModeLeader;
    if (insert_modes) {
      const uint64_t cap = h_tables_[t].cap;
      if (counts[t] == 0 || len_[t] + counts[t] > cap) to.mode = dev::kModeSkip;  // rayon_cpu.rs:134-142
      else if (len_[t] + counts[t] >= cap) to.mode = dev::kModeSerial;           // may leave no barrier
    }
    if (to.mode == dev::kModeLeader || !insert_modes) {
      for (uint64_t c = 0; c < counts[t]; c += tune_.block) {
        Chunk ch;
        ch.table = t;
        ch.count = static_cast<uint32_t>(std::min<uint64_t>(tune_.block, counts[t] - c));
        ch.begin = begin + c;
        ch.local = c;
        op.chunks.push_back(ch);
      }
    }
    begin += counts[t];
  }
}

fm_status GroupImpl::ensure_keys(uint64_t total) {
  if (total <= key_cap_) return FM_OK;
  if (total > 0x7FFFFFFFull) return fail(FM_ERR_INVALID_ARGUMENT, "group batch above 2^31-1 keys");
  (void)hipStreamSynchronize(stream_);
  dfree(dslot_);
  dfree(ev_);
  dfree(action_);
  dfree(big_);
  key_cap_ = 0;
  const uint64_t want = std::max<uint64_t>(total, 4096);
  hipError_t e;
  if ((e = hipMalloc(&dslot_, want * 4)) != hipSuccess || (e = hipMalloc(&ev_, want * 8)) != hipSuccess ||
      (e = hipMalloc(&action_, want)) != hipSuccess || (e = hipMalloc(&big_, want * 17)) != hipSuccess)
    return fail(FM_ERR_GPU_ALLOC, hmsg("hipMalloc (group scratch)", e));
  key_cap_ = want;
  return FM_OK;
}

// Uploads one op's descriptors (grows the device arrays when needed).
fm_status GroupImpl::stage(int which, const HostOp& op, uint64_t total) {
  fm_status st;
  if ((st = ensure_keys(total)) != FM_OK) return st;
  hipError_t e;
  if (!d_tops_[which] && (e = hipMalloc(&d_tops_[which], n_ * sizeof(TableOp))) != hipSuccess)
    return fail(FM_ERR_GPU_ALLOC, hmsg("hipMalloc (ops)", e));
  if (op.chunks.size() > chunk_cap_[which]) {
    (void)hipStreamSynchronize(stream_);
    dfree(d_chunks_[which]);
    chunk_cap_[which] = 0;
    const uint64_t want = std::max<uint64_t>(op.chunks.size() * 2, 1024);
    if ((e = hipMalloc(&d_chunks_[which], want * sizeof(Chunk))) != hipSuccess)
      return fail(FM_ERR_GPU_ALLOC, hmsg("hipMalloc (chunks)", e));
    chunk_cap_[which] = want;
  }
  h_nch_[which] = static_cast<uint32_t>(op.chunks.size());
  if ((e = hipMemcpyAsync(d_tops_[which], op.tops.data(), n_ * sizeof(TableOp), hipMemcpyHostToDevice, stream_)) !=
          hipSuccess ||
      (op.chunks.size() && (e = hipMemcpyAsync(d_chunks_[which], op.chunks.data(), op.chunks.size() * sizeof(Chunk),
                                               hipMemcpyHostToDevice, stream_)) != hipSuccess) ||
      (e = hipMemcpyAsync(d_nch_ + which, &h_nch_[which], 4, hipMemcpyHostToDevice, stream_)) != hipSuccess)
    return fail(FM_ERR_TRANSFER, hmsg("descriptor upload", e));
  // The host vectors are read by the async copies: wait before they change.
  if ((e = hipStreamSynchronize(stream_)) != hipSuccess) return fail(FM_ERR_TRANSFER, hmsg("descriptor upload", e));
  return FM_OK;
}

fm_status GroupImpl::enqueue_insert(const uint8_t* keys, const uint8_t* vals) {
  OpDesc op{d_tops_[0], d_chunks_[0], d_nch_ + 0};
  ops_.insert(d_tables_, op, d_tops_[0], n_, keys, vals, Q_, free_all_, d_qoff_, total_words_, dslot_, big_, key_cap_,
              nbig_, ins_ctr_, tune_.block, tune_.grid, tune_.coop_arc, stream_);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? FM_OK : fail(FM_ERR_KERNEL_LAUNCH, hmsg("group insert", e));
}

fm_status GroupImpl::enqueue_get(const uint8_t* keys, uint8_t* ov, uint8_t* of) {
  OpDesc op{d_tops_[1], d_chunks_[1], d_nch_ + 1};
  ops_.get(d_tables_, op, keys, ov, of, tune_.group, tune_.block, tune_.grid, stream_);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? FM_OK : fail(FM_ERR_KERNEL_LAUNCH, hmsg("group get", e));
}

fm_status GroupImpl::enqueue_remove(const uint8_t* keys) {
  OpDesc op{d_tops_[2], d_chunks_[2], d_nch_ + 2};
  ops_.remove(d_tables_, op, d_tops_[2], n_, keys, ev_, action_, rem_ctr_, tune_.group, tune_.block, tune_.grid,
              stream_);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? FM_OK : fail(FM_ERR_KERNEL_LAUNCH, hmsg("group remove", e));
}

fm_status GroupImpl::finish(bool ins, bool rem, uint64_t* out_new, int32_t* out_status, uint64_t* out_removed,
                            const uint64_t* ins_counts) {
  hipError_t e = hipSuccess;
  if (ins) e = hipMemcpyAsync(h_ins_ctr_, ins_ctr_, n_ * sizeof(Counters), hipMemcpyDeviceToHost, stream_);
  if (e == hipSuccess && rem) e = hipMemcpyAsync(h_rem_ctr_, rem_ctr_, n_ * sizeof(Counters), hipMemcpyDeviceToHost, stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  if (e != hipSuccess) return fail(FM_ERR_TRANSFER, hmsg("group counters", e));
  fm_status worst = FM_OK;
  if (ins) {
    for (uint32_t t = 0; t < n_; ++t) {
      int32_t s = FM_OK;
      const uint32_t mode = hop_[0].tops[t].mode;
      if (mode == dev::kModeSkip && ins_counts[t] > 0) s = FM_ERR_TABLE_FULL;
      if (h_ins_ctr_[t].status & (dev::kStatusBudget | dev::kStatusDryFail)) s = FM_ERR_KERNEL_LAUNCH;
      if (s == FM_OK) len_[t] += h_ins_ctr_[t].count;
      if (out_new) out_new[t] = s == FM_OK ? h_ins_ctr_[t].count : 0;
      if (out_status) out_status[t] = s;
      if (s == FM_ERR_KERNEL_LAUNCH) worst = FM_ERR_KERNEL_LAUNCH;
    }
  }
  if (rem) {
    for (uint32_t t = 0; t < n_; ++t) {
      len_[t] -= std::min<uint64_t>(len_[t], h_rem_ctr_[t].count);
      if (out_removed) out_removed[t] = h_rem_ctr_[t].count;
    }
  }
  if (worst != FM_OK) return fail(worst, "group insert: a table's walk exceeded its budget");
  return FM_OK;
}

static uint64_t sum_counts(const uint64_t* c, uint32_t n) {
  uint64_t s = 0;
  for (uint32_t i = 0; i < n; ++i) s += c[i];
  return s;
}

fm_status GroupImpl::insert(const uint8_t* keys, const uint8_t* vals, const uint64_t* counts, uint64_t* out_new,
                            int32_t* out_status) {
  fm_status st;
  if ((st = use()) != FM_OK) return st;
  build_op(hop_[0], counts, true);
  if ((st = stage(0, hop_[0], sum_counts(counts, n_))) != FM_OK) return st;
  hipLaunchKernelGGL(dev::kg_zero, dim3(gridn(n_, 256)), dim3(256), 0, stream_, ins_ctr_, rem_ctr_, n_, nbig_);
  if ((st = enqueue_insert(keys, vals)) != FM_OK) return st;
  return finish(true, false, out_new, out_status, nullptr, counts);
}

fm_status GroupImpl::get(const uint8_t* keys, const uint64_t* counts, uint8_t* ov, uint8_t* of) {
  fm_status st;
  if ((st = use()) != FM_OK) return st;
  build_op(hop_[1], counts, false);
  if ((st = stage(1, hop_[1], sum_counts(counts, n_))) != FM_OK) return st;
  if ((st = enqueue_get(keys, ov, of)) != FM_OK) return st;
  hipError_t e = hipStreamSynchronize(stream_);
  return e == hipSuccess ? FM_OK : fail(FM_ERR_KERNEL_LAUNCH, hmsg("group get", e));
}

fm_status GroupImpl::remove(const uint8_t* keys, const uint64_t* counts, uint64_t* out_removed) {
  fm_status st;
  if ((st = use()) != FM_OK) return st;
  build_op(hop_[2], counts, false);
  if ((st = stage(2, hop_[2], sum_counts(counts, n_))) != FM_OK) return st;
  hipLaunchKernelGGL(dev::kg_zero, dim3(gridn(n_, 256)), dim3(256), 0, stream_, ins_ctr_, rem_ctr_, n_, nbig_);
  if ((st = enqueue_remove(keys)) != FM_OK) return st;
  return finish(false, true, nullptr, nullptr, out_removed, nullptr);
}

fm_status GroupImpl::step(const uint8_t* ik, const uint8_t* iv, const uint64_t* ic, const uint8_t* gk,
                          const uint64_t* gc, uint8_t* ov, uint8_t* of, const uint8_t* rk, const uint64_t* rc,
                          uint64_t* out_new, int32_t* out_status, uint64_t* out_removed) {
  fm_status st;
  if ((st = use()) != FM_OK) return st;
  build_op(hop_[0], ic, true);
  build_op(hop_[1], gc, false);
  build_op(hop_[2], rc, false);
  const uint64_t total = std::max(sum_counts(ic, n_), std::max(sum_counts(gc, n_), sum_counts(rc, n_)));
  for (int w = 0; w < 3; ++w)
    if ((st = stage(w, hop_[w], total)) != FM_OK) return st;
  // The graph bakes in every pointer and launch shape: re-capture when any
  // of them changed.
  std::vector<const void*> key = {ik, iv, gk, ov, of, rk, d_chunks_[0], d_chunks_[1], d_chunks_[2], dslot_,
                                  reinterpret_cast<const void*>(static_cast<uintptr_t>(tune_.group)),
                                  reinterpret_cast<const void*>(static_cast<uintptr_t>(tune_.block)),
                                  reinterpret_cast<const void*>(static_cast<uintptr_t>(tune_.grid)),
                                  reinterpret_cast<const void*>(static_cast<uintptr_t>(tune_.coop_arc))};
  hipError_t e;
  if (!exec_ || key != graph_key_) {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    if ((e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return fail(FM_ERR_KERNEL_LAUNCH, hmsg("hipStreamBeginCapture", e));
    hipLaunchKernelGGL(dev::kg_zero, dim3(gridn(n_, 256)), dim3(256), 0, stream_, ins_ctr_, rem_ctr_, n_, nbig_);
    fm_status a = enqueue_insert(ik, iv);
    fm_status b = enqueue_get(gk, ov, of);
    fm_status c = enqueue_remove(rk);
    e = hipStreamEndCapture(stream_, &graph_);
    if (a != FM_OK || b != FM_OK || c != FM_OK || e != hipSuccess)
      return fail(FM_ERR_KERNEL_LAUNCH, hmsg("graph capture", e));
    if ((e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0)) != hipSuccess)
      return fail(FM_ERR_KERNEL_LAUNCH, hmsg("hipGraphInstantiate", e));
    graph_key_ = key;
  }
  if ((e = hipGraphLaunch(exec_, stream_)) != hipSuccess) return fail(FM_ERR_KERNEL_LAUNCH, hmsg("hipGraphLaunch", e));
  return finish(true, true, out_new, out_status, out_removed, ic);
}

fm_status GroupImpl::export_slots(uint32_t table, uint32_t* flags, void* keys, void* values) {
  fm_status st;
  if ((st = use()) != FM_OK) return st;
  if (table >= n_) return fail(FM_ERR_INVALID_ARGUMENT, "table index out of range");
  const Table& t = h_tables_[table];
  hipError_t e;
  if ((e = hipMemcpyAsync(flags, t.flags, t.cap * 4, hipMemcpyDeviceToHost, stream_)) != hipSuccess ||
      (keys && (e = hipMemcpyAsync(keys, t.keys, t.cap * K_, hipMemcpyDeviceToHost, stream_)) != hipSuccess) ||
      (values && (e = hipMemcpyAsync(values, t.vals, t.cap * V_, hipMemcpyDeviceToHost, stream_)) != hipSuccess) ||
      (e = hipStreamSynchronize(stream_)) != hipSuccess)
    return fail(FM_ERR_TRANSFER, hmsg("group export", e));
  for (uint64_t i = 0; i < t.cap; ++i) {
    if ((flags[i] & 0xFu) == dev::kOccupied) continue;
    if (keys) std::memset(static_cast<uint8_t*>(keys) + i * K_, 0, K_);
    if (values) std::memset(static_cast<uint8_t*>(values) + i * V_, 0, V_);
  }
  return FM_OK;
}

fm_status GroupImpl::copy_from(GroupImpl& src) {
  if (src.n_ != n_ || src.dev_ != dev_ || src.K_ != K_ || src.V_ != V_ || src.mode_ != mode_ ||
      src.total_slots_ != total_slots_)
    return fail(FM_ERR_INVALID_ARGUMENT, "fm_group_copy: groups differ in shape");
  for (uint32_t t = 0; t < n_; ++t)
    if (src.h_tables_[t].cap != h_tables_[t].cap) return fail(FM_ERR_INVALID_ARGUMENT, "fm_group_copy: capacities differ");
  fm_status st;
  if ((st = use()) != FM_OK) return st;
  hipError_t e = hipStreamSynchronize(src.stream_);
  if (e == hipSuccess) e = hipMemcpyAsync(keys_all_, src.keys_all_, total_slots_ * K_, hipMemcpyDeviceToDevice, stream_);
  if (e == hipSuccess) e = hipMemcpyAsync(vals_all_, src.vals_all_, total_slots_ * V_, hipMemcpyDeviceToDevice, stream_);
  if (e == hipSuccess) e = hipMemcpyAsync(flags_all_, src.flags_all_, total_slots_ * 4, hipMemcpyDeviceToDevice, stream_);
  if (e == hipSuccess) e = hipMemcpyAsync(free_all_, src.free_all_, total_words_ * 8, hipMemcpyDeviceToDevice, stream_);
  if (e == hipSuccess) e = hipMemcpyAsync(empty_all_, src.empty_all_, total_words_ * 8, hipMemcpyDeviceToDevice, stream_);
  if (e != hipSuccess) return fail(FM_ERR_TRANSFER, hmsg("fm_group_copy", e));
  len_ = src.len_;
  return FM_OK;
}

// Launch-shape search.  Every run restores `scratch` from this group's
// tables, so each candidate steps from the same state and must give the same
// per-table counts and statuses; the wall time of the whole step (descriptor
// upload, graph launch, counter read-back) is what is compared.
fm_status GroupImpl::autotune(GroupImpl* scratch, const uint8_t* ik, const uint8_t* iv, const uint64_t* ic,
                              const uint8_t* gk, const uint64_t* gc, uint8_t* ov, uint8_t* of, const uint8_t* rk,
                              const uint64_t* rc, const fm_group_tuning* cand, uint32_t n_cand, uint32_t reps,
                              fm_group_tuning* best, double* best_us, uint32_t* evaluated) {
  fm_status st;
  if ((st = use()) != FM_OK) return st;
  if (reps == 0 || reps > 64) return fail(FM_ERR_INVALID_ARGUMENT, "fm_group_autotune: reps must be 1..64");
  std::unique_ptr<GroupImpl> own;
  if (!scratch) {
    std::vector<uint64_t> caps(n_);
    for (uint32_t t = 0; t < n_; ++t) caps[t] = h_tables_[t].cap;
    own.reset(new (std::nothrow) GroupImpl());
    if (!own) return fail(FM_ERR_GPU_ALLOC, "fm_group_autotune: host allocation failed");
    std::string err;
    if ((st = own->init(n_, caps.data(), K_, V_, mode_, dev_, &err)) != FM_OK) return fail(st, err);
    scratch = own.get();
  }
  std::vector<uint64_t> ref_new(n_), ref_rem(n_), cur_new(n_), cur_rem(n_);
  std::vector<int32_t> ref_status(n_), cur_status(n_);
  bool have_ref = false;
  uint32_t runs = 0;
  // Median wall time of `reps` restored steps (after one untimed step that
  // captures the graph for this shape), or a negative value on error.
  auto measure = [&](const fm_group_tuning& t, double* us) -> fm_status {
    fm_status s;
    if ((s = scratch->set_tuning(t)) != FM_OK) return s;
    std::vector<double> times;
    for (uint32_t r = 0; r <= reps; ++r) {
      if ((s = scratch->copy_from(*this)) != FM_OK) return s;
      hipError_t e = hipStreamSynchronize(scratch->stream_);
      if (e != hipSuccess) return fail(FM_ERR_TRANSFER, hmsg("fm_group_autotune: restore", e));
      const auto t0 = std::chrono::steady_clock::now();
      s = scratch->step(ik, iv, ic, gk, gc, ov, of, rk, rc, cur_new.data(), cur_status.data(), cur_rem.data());
      const auto t1 = std::chrono::steady_clock::now();
      if (s != FM_OK) return s;
      if (r) times.push_back(std::chrono::duration<double, std::micro>(t1 - t0).count());
    }
    ++runs;
    if (!have_ref) {
      ref_new = cur_new;
      ref_status = cur_status;
      ref_rem = cur_rem;
      have_ref = true;
    } else if (cur_new != ref_new || cur_status != ref_status || cur_rem != ref_rem) {
      return fail(FM_ERR_KERNEL_LAUNCH, "fm_group_autotune: launch shapes gave different per-table results");
    }
    std::sort(times.begin(), times.end());
    *us = times[times.size() / 2];
    return FM_OK;
  };
  fm_group_tuning top = tune_;
  double top_us = 0;
  if (n_cand) {  // exhaustive over the caller's list
    bool any = false;
    for (uint32_t i = 0; i < n_cand; ++i) {
      double us;
      if ((st = measure(cand[i], &us)) != FM_OK) return st;
      if (!any || us < top_us) top = cand[i], top_us = us, any = true;
    }
  } else {  // coordinate descent from the current tuning, one knob at a time
    static const uint32_t kGroups[] = {1, 2, 4, 8, 16};
    static const uint32_t kBlocks[] = {64, 128, 256, 512};
    static const uint32_t kGrids[] = {256, 512, 1024, 2048, 4096, 8192};
    static const uint32_t kArcs[] = {0, 16, 64, 256, 4096};
    struct Knob {
      uint32_t fm_group_tuning::*field;
      const uint32_t* vals;
      uint32_t n;
    };
    const Knob knobs[] = {{&fm_group_tuning::block, kBlocks, 4},
                          {&fm_group_tuning::grid, kGrids, 6},
                          {&fm_group_tuning::group, kGroups, 5},
                          {&fm_group_tuning::coop_arc, kArcs, 5}};
    if ((st = measure(top, &top_us)) != FM_OK) return st;
    for (int pass = 0; pass < 3; ++pass) {
      bool moved = false;
      for (const Knob& k : knobs) {
        for (uint32_t i = 0; i < k.n; ++i) {
          if (top.*k.field == k.vals[i]) continue;
          fm_group_tuning t = top;
          t.*k.field = k.vals[i];
          double us;
          if ((st = measure(t, &us)) != FM_OK) return st;
          if (us < top_us * 0.97) top = t, top_us = us, moved = true;  // 3%: above run-to-run noise
        }
      }
      if (!moved) break;
    }
  }
  tune_ = top;
  if (best) *best = top;
  if (best_us) *best_us = top_us;
  if (evaluated) *evaluated = runs;
  return FM_OK;
}

}  // namespace fm

struct fm_group {
  fm::GroupImpl impl;
  std::mutex mu;
};

extern "C" {

static fm_status group_bad(const char* m) {
  fm::set_error(FM_ERR_INVALID_ARGUMENT, m);
  return FM_ERR_INVALID_ARGUMENT;
}

fm_status fm_group_create(uint32_t n_tables, const uint64_t* capacities, uint32_t key_size, uint32_t value_size,
                          uint32_t hash_strategy, int32_t device_id, fm_group** out) {
  if (!out) return group_bad("fm_group_create: null out");
  *out = nullptr;
  std::unique_ptr<fm_group> g(new (std::nothrow) fm_group());
  if (!g) return group_bad("fm_group_create: host allocation failed");
  std::string err;
  fm_status st = g->impl.init(n_tables, capacities, key_size, value_size, hash_strategy, device_id, &err);
  if (st != FM_OK) {
    fm::set_error(st, err);
    return st;
  }
  *out = g.release();
  return FM_OK;
}

void fm_group_destroy(fm_group* g) { delete g; }
uint32_t fm_group_tables(const fm_group* g) { return g ? g->impl.n_ : 0; }
uint64_t fm_group_len(const fm_group* g, uint32_t t) { return g && t < g->impl.n_ ? g->impl.len_[t] : 0; }
uint64_t fm_group_capacity(const fm_group* g, uint32_t t) { return g && t < g->impl.n_ ? g->impl.h_tables_[t].cap : 0; }
void* fm_group_stream(fm_group* g) { return g ? g->impl.stream_ : nullptr; }

fm_status fm_group_alloc_device(fm_group* g, uint64_t bytes, void** out) {
  if (!g || !out) return group_bad("fm_group_alloc_device: null argument");
  std::lock_guard<std::mutex> l(g->mu);
  (void)hipSetDevice(g->impl.dev_);
  hipError_t e = hipMalloc(out, std::max<uint64_t>(bytes, 1));
  if (e == hipSuccess) e = hipMemset(*out, 0, std::max<uint64_t>(bytes, 1));
  if (e != hipSuccess) {
    fm::set_error(FM_ERR_GPU_ALLOC, fm::hmsg("hipMalloc", e));
    return FM_ERR_GPU_ALLOC;
  }
  return FM_OK;
}

fm_status fm_group_free_device(fm_group* g, void* p) {
  if (!g) return group_bad("null group");
  std::lock_guard<std::mutex> l(g->mu);
  (void)hipSetDevice(g->impl.dev_);
  (void)hipStreamSynchronize(g->impl.stream_);
  return hipFree(p) == hipSuccess ? FM_OK : FM_ERR_GPU_ALLOC;
}

fm_status fm_group_upload(fm_group* g, const void* host, uint64_t bytes, void** out) {
  fm_status st = fm_group_alloc_device(g, bytes, out);
  if (st != FM_OK) return st;
  std::lock_guard<std::mutex> l(g->mu);
  if (bytes && hipMemcpy(*out, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    fm::set_error(FM_ERR_TRANSFER, "fm_group_upload failed");
    return FM_ERR_TRANSFER;
  }
  return FM_OK;
}

fm_status fm_group_download(fm_group* g, const void* dptr, uint64_t bytes, void* host) {
  if (!g) return group_bad("null group");
  std::lock_guard<std::mutex> l(g->mu);
  (void)hipSetDevice(g->impl.dev_);
  (void)hipStreamSynchronize(g->impl.stream_);
  if (bytes && hipMemcpy(host, dptr, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
    fm::set_error(FM_ERR_TRANSFER, "fm_group_download failed");
    return FM_ERR_TRANSFER;
  }
  return FM_OK;
}

fm_status fm_group_bulk_insert_device(fm_group* g, const void* d_keys, const void* d_values, const uint64_t* counts,
                                      uint64_t* out_new, int32_t* out_status) {
  if (!g || !counts) return group_bad("fm_group_bulk_insert_device: null argument");
  std::lock_guard<std::mutex> l(g->mu);
  return g->impl.insert(static_cast<const uint8_t*>(d_keys), static_cast<const uint8_t*>(d_values), counts, out_new,
                        out_status);
}

fm_status fm_group_bulk_get_device(fm_group* g, const void* d_keys, const uint64_t* counts, void* d_out_values,
                                   uint8_t* d_out_found) {
  if (!g || !counts) return group_bad("fm_group_bulk_get_device: null argument");
  std::lock_guard<std::mutex> l(g->mu);
  return g->impl.get(static_cast<const uint8_t*>(d_keys), counts, static_cast<uint8_t*>(d_out_values), d_out_found);
}

fm_status fm_group_bulk_remove_device(fm_group* g, const void* d_keys, const uint64_t* counts, uint64_t* out_removed) {
  if (!g || !counts) return group_bad("fm_group_bulk_remove_device: null argument");
  std::lock_guard<std::mutex> l(g->mu);
  return g->impl.remove(static_cast<const uint8_t*>(d_keys), counts, out_removed);
}

fm_status fm_group_step_device(fm_group* g, const void* d_ins_keys, const void* d_ins_values,
                               const uint64_t* ins_counts, const void* d_get_keys, const uint64_t* get_counts,
                               void* d_out_values, uint8_t* d_out_found, const void* d_rem_keys,
                               const uint64_t* rem_counts, uint64_t* out_new, int32_t* out_status,
                               uint64_t* out_removed) {
  if (!g || !ins_counts || !get_counts || !rem_counts) return group_bad("fm_group_step_device: null argument");
  std::lock_guard<std::mutex> l(g->mu);
  return g->impl.step(static_cast<const uint8_t*>(d_ins_keys), static_cast<const uint8_t*>(d_ins_values), ins_counts,
                      static_cast<const uint8_t*>(d_get_keys), get_counts, static_cast<uint8_t*>(d_out_values),
                      d_out_found, static_cast<const uint8_t*>(d_rem_keys), rem_counts, out_new, out_status,
                      out_removed);
}

fm_status fm_group_export_slots(fm_group* g, uint32_t table, uint32_t* flags, void* keys, void* values) {
  if (!g || !flags) return group_bad("fm_group_export_slots: null argument");
  std::lock_guard<std::mutex> l(g->mu);
  return g->impl.export_slots(table, flags, keys, values);
}

fm_status fm_group_copy(fm_group* dst, fm_group* src) {
  if (!dst || !src) return group_bad("fm_group_copy: null argument");
  if (dst == src) return FM_OK;
  std::scoped_lock l(dst->mu, src->mu);
  return dst->impl.copy_from(src->impl);
}

fm_status fm_group_autotune(fm_group* g, fm_group* scratch, const void* d_ins_keys, const void* d_ins_values,
                            const uint64_t* ins_counts, const void* d_get_keys, const uint64_t* get_counts,
                            void* d_out_values, uint8_t* d_out_found, const void* d_rem_keys,
                            const uint64_t* rem_counts, const fm_group_tuning* candidates, uint32_t n_candidates,
                            uint32_t reps, fm_group_tuning* out_best, double* out_step_us, uint32_t* out_evaluated) {
  if (!g || !ins_counts || !get_counts || !rem_counts || (n_candidates && !candidates))
    return group_bad("fm_group_autotune: null argument");
  if (scratch == g) return group_bad("fm_group_autotune: scratch must be another group");
  if (scratch) {
    std::scoped_lock l(g->mu, scratch->mu);
    return g->impl.autotune(&scratch->impl, static_cast<const uint8_t*>(d_ins_keys),
                            static_cast<const uint8_t*>(d_ins_values), ins_counts,
                            static_cast<const uint8_t*>(d_get_keys), get_counts, static_cast<uint8_t*>(d_out_values),
                            d_out_found, static_cast<const uint8_t*>(d_rem_keys), rem_counts, candidates, n_candidates,
                            reps, out_best, out_step_us, out_evaluated);
  }
  std::lock_guard<std::mutex> l(g->mu);
  return g->impl.autotune(nullptr, static_cast<const uint8_t*>(d_ins_keys), static_cast<const uint8_t*>(d_ins_values),
                          ins_counts, static_cast<const uint8_t*>(d_get_keys), get_counts,
                          static_cast<uint8_t*>(d_out_values), d_out_found, static_cast<const uint8_t*>(d_rem_keys),
                          rem_counts, candidates, n_candidates, reps, out_best, out_step_us, out_evaluated);
}

fm_status fm_group_get_tuning(const fm_group* g, fm_group_tuning* out) {
  if (!g || !out) return group_bad("fm_group_get_tuning: null argument");
  *out = g->impl.tune_;
  return FM_OK;
}

fm_status fm_group_set_tuning(fm_group* g, const fm_group_tuning* t) {
  if (!g || !t) return group_bad("fm_group_set_tuning: null argument");
  std::lock_guard<std::mutex> l(g->mu);
  return g->impl.set_tuning(*t);
}

}  // extern "C"
