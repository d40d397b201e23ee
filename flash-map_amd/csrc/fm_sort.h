// fm_sort.h -- stable sort of (u32 key, u32 value) pairs for gfx950.
//
// Used to group insert keys by closure id (the sort-grouped insert of a batch
// that may fill the table) and remove keys by EMPTY-delimited run (the exact
// conflict replay of a remove batch).  Every kernel takes an optional gate:
// a device status word and a mask; the kernel exits at once unless
// (*gate & mask) != 0, so a pipeline can queue the sort behind a kernel that
// decides on the device whether it is needed (no host readback).
//
//  * n <= kSortSmallMax: one workgroup sorts the batch in LDS.  Each pair
//    becomes the 64-bit word key << 32 | position, sorted by a bitonic
//    network; the position breaks ties, so equal keys keep their input order
//    (stable), and each value is gathered from its input position.
//  * larger n: least-significant-digit radix sort, kSortBits bits per pass,
//    as many passes as end_bit needs.  A pass counts digits per tile of
//    kSortTile pairs (k_sort_hist), scans the digit-major count matrix in
//    chunks (k_sort_scan, plus k_sort_scan_top when there is more than one
//    chunk), and scatters each tile stably (k_sort_scatter): within a wave,
//    lanes with equal digits are found with one ballot per digit bit and
//    ranked by lane order; waves of a round, and rounds of a tile, are ranked
//    in order through per-digit LDS counters.  Passes alternate between the
//    output arrays and the scratch buffer so the last one ends in the output.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fm {
namespace sort {

constexpr uint32_t kSortSmallMax = 16384;       // pairs sorted by one workgroup in LDS (128 KiB)
constexpr uint32_t kSortSmallBlock = 1024;
constexpr int kSortBits = 11;                   // radix digit width
constexpr uint32_t kSortBins = 1u << kSortBits;
constexpr uint32_t kSortBlock = 256;            // four waves
constexpr uint32_t kSortRounds = 16;            // rounds of kSortBlock pairs per tile
constexpr uint32_t kSortTile = kSortBlock * kSortRounds;
constexpr uint32_t kScanBlock = 1024;
constexpr uint32_t kScanItems = 4;
constexpr uint32_t kScanChunk = kScanBlock * kScanItems;  // count-matrix entries per scan workgroup

__device__ __forceinline__ bool gate_closed(const uint32_t* gate, uint32_t mask) {
  return gate && (__hip_atomic_load(gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & mask) == 0u;
}

// ---- one-workgroup sort ----------------------------------------------------
static __global__ void __launch_bounds__(kSortSmallBlock)
k_sort_small(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
             uint32_t* __restrict__ vout, uint32_t n, const uint32_t* gate, uint32_t mask) {
  extern __shared__ unsigned long long w[];
  if (gate_closed(gate, mask)) return;
  uint32_t p2 = 1;
  while (p2 < n) p2 <<= 1;
  for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x)
    w[i] = i < n ? (static_cast<unsigned long long>(kin[i]) << 32) | i : ~0ull;
  __syncthreads();
  for (uint32_t size = 2; size <= p2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t i = threadIdx.x; i < p2 / 2; i += blockDim.x) {
        // i-th compare-exchange pair of this stage: lo has bit `stride` clear
        const uint32_t lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));
        const uint32_t hi = lo | stride;
        const unsigned long long a = w[lo], b = w[hi];
        const bool up = (lo & size) == 0;
        if ((a > b) == up) {
          w[lo] = b;
          w[hi] = a;
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const unsigned long long x = w[i];
    kout[i] = static_cast<uint32_t>(x >> 32);
    vout[i] = vin[static_cast<uint32_t>(x)];
  }
}

// ---- radix passes ----------------------------------------------------------
__device__ __forceinline__ uint32_t digit_of(uint32_t k, int shift) { return (k >> shift) & (kSortBins - 1u); }

// H[d * nblk + b] = pairs of tile b whose digit is d.
static __global__ void __launch_bounds__(kSortBlock)
k_sort_hist(const uint32_t* __restrict__ kin, uint64_t n, int shift, uint32_t* __restrict__ H, uint32_t nblk,
            const uint32_t* gate, uint32_t mask) {
  __shared__ uint32_t c[kSortBins];
  if (gate_closed(gate, mask)) return;
  for (uint32_t d = threadIdx.x; d < kSortBins; d += blockDim.x) c[d] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  for (uint32_t r = 0; r < kSortRounds; ++r) {
    const uint64_t i = base + r * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&c[digit_of(kin[i], shift)], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kSortBins; d += blockDim.x) H[static_cast<uint64_t>(d) * nblk + blockIdx.x] = c[d];
}

// Exclusive scan of one chunk of kScanChunk entries of H in place; the chunk
// total goes to S[chunk].
static __global__ void __launch_bounds__(kScanBlock)
k_sort_scan(uint32_t* __restrict__ H, uint64_t len, uint32_t* __restrict__ S, const uint32_t* gate, uint32_t mask) {
  __shared__ uint32_t part[kScanBlock / 64];
  if (gate_closed(gate, mask)) return;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanChunk + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
  for (uint32_t j = 0; j < kScanItems; ++j) {
    v[j] = base + j < len ? H[base + j] : 0u;
    sum += v[j];
  }
  // inclusive scan of per-thread sums: wave shuffles, then across waves
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t inc = sum;
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) part[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (uint32_t k = 0; k < kScanBlock / 64; ++k) {
      const uint32_t t = part[k];
      part[k] = run;
      run += t;
    }
    S[blockIdx.x] = run;
  }
  __syncthreads();
  uint32_t ex = part[wave] + inc - sum;
  for (uint32_t j = 0; j < kScanItems; ++j) {
    if (base + j < len) H[base + j] = ex;
    ex += v[j];
  }
}

// Exclusive scan of the chunk totals (one workgroup; S has at most a few
// thousand entries for any batch this backend sorts).
static __global__ void __launch_bounds__(kScanBlock)
k_sort_scan_top(uint32_t* __restrict__ S, uint32_t m, const uint32_t* gate, uint32_t mask) {
  __shared__ uint32_t part[kScanBlock / 64];
  __shared__ uint32_t carry;
  if (gate_closed(gate, mask)) return;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < m; base += kScanBlock) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < m ? S[i] : 0u;
    uint32_t inc = v;
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    if (lane == 63) part[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = carry;
      for (uint32_t k = 0; k < kScanBlock / 64; ++k) {
        const uint32_t t = part[k];
        part[k] = run;
        run += t;
      }
      carry = run;
    }
    __syncthreads();
    if (i < m) S[i] = part[wave] + inc - v;
    __syncthreads();
  }
}

// Stable scatter of tile b: output position of a pair = (pairs with a
// smaller digit) + (pairs with its digit in earlier tiles) + (pairs with its
// digit earlier in this tile).
static __global__ void __launch_bounds__(kSortBlock)
k_sort_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
               uint32_t* __restrict__ vout, uint64_t n, int shift, const uint32_t* __restrict__ H,
               const uint32_t* __restrict__ S, uint32_t nblk, const uint32_t* gate, uint32_t mask) {
  __shared__ uint32_t base[kSortBins];
  __shared__ uint32_t wcnt[kSortBlock / 64][kSortBins];
  if (gate_closed(gate, mask)) return;
  for (uint32_t d = threadIdx.x; d < kSortBins; d += blockDim.x) {
    const uint64_t h = static_cast<uint64_t>(d) * nblk + blockIdx.x;
    base[d] = H[h] + S[h / kScanChunk];
    for (uint32_t w = 0; w < kSortBlock / 64; ++w) wcnt[w][d] = 0;
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const uint64_t tile = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  for (uint32_t r = 0; r < kSortRounds; ++r) {
    const uint64_t i = tile + r * kSortBlock + threadIdx.x;
    const bool valid = i < n;
    uint32_t k = 0, v = 0, d = 0;
    if (valid) {
      k = kin[i];
      v = vin[i];
      d = digit_of(k, shift);
    }
    unsigned long long peers = __ballot(valid);
    for (int b = 0; b < kSortBits; ++b) {
      const bool set = (d >> b) & 1u;
      const unsigned long long m = __ballot(valid && set);
      peers &= set ? m : ~m;
    }
    const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt));
    const bool leader = valid && rank == 0;
    if (leader) wcnt[wave][d] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    if (valid) {
      uint32_t pos = base[d] + rank;
      for (uint32_t w = 0; w < wave; ++w) pos += wcnt[w][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    if (leader) {
      atomicAdd(&base[d], wcnt[wave][d]);
      wcnt[wave][d] = 0;
    }
    __syncthreads();
  }
}

}  // namespace sort

constexpr uint64_t sort_round256(uint64_t b) { return (b + 255) & ~uint64_t(255); }

// Scratch bytes sort_pairs needs for n pairs (0 for the one-workgroup path).
constexpr uint64_t sort_scratch_bytes(uint64_t n) {
  if (n <= sort::kSortSmallMax) return 0;
  const uint64_t nblk = (n + sort::kSortTile - 1) / sort::kSortTile;
  const uint64_t len = nblk * sort::kSortBins;
  const uint64_t chunks = (len + sort::kScanChunk - 1) / sort::kScanChunk;
  return 2 * sort_round256(4 * n) + sort_round256(4 * len) + sort_round256(4 * chunks);
}

// Stable sort of n (key, value) pairs by key bits [0, end_bit) into
// kout/vout (inputs unchanged).  tmp: sort_scratch_bytes(n) bytes, 256-byte
// aligned.  gate (optional): every kernel exits unless (*gate & mask) != 0.
inline hipError_t sort_pairs(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                             int end_bit, void* tmp, const uint32_t* gate, uint32_t mask, hipStream_t s) {
  using namespace sort;
  if (n == 0) return hipSuccess;
  if (n <= kSortSmallMax) {
    uint32_t p2 = 1;
    while (p2 < n) p2 <<= 1;
    hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(kSortSmallBlock), p2 * sizeof(unsigned long long), s, kin, vin,
                       kout, vout, static_cast<uint32_t>(n), gate, mask);
    return hipGetLastError();
  }
  const uint64_t nblk = (n + kSortTile - 1) / kSortTile;
  const uint64_t len = nblk * kSortBins;
  const uint64_t chunks = (len + kScanChunk - 1) / kScanChunk;
  uint8_t* p = static_cast<uint8_t*>(tmp);
  uint32_t* tk = reinterpret_cast<uint32_t*>(p);
  uint32_t* tv = reinterpret_cast<uint32_t*>(p + sort_round256(4 * n));
  uint32_t* H = reinterpret_cast<uint32_t*>(p + 2 * sort_round256(4 * n));
  uint32_t* S = reinterpret_cast<uint32_t*>(p + 2 * sort_round256(4 * n) + sort_round256(4 * len));
  const int passes = end_bit <= 0 ? 1 : (end_bit + kSortBits - 1) / kSortBits;
  // pass p writes the output arrays when (passes - 1 - p) is even
  const uint32_t* ck = kin;
  const uint32_t* cv = vin;
  for (int pass = 0; pass < passes; ++pass) {
    const bool to_out = ((passes - 1 - pass) & 1) == 0;
    uint32_t* nk = to_out ? kout : tk;
    uint32_t* nv = to_out ? vout : tv;
    const int shift = pass * kSortBits;
    hipLaunchKernelGGL(k_sort_hist, dim3(static_cast<uint32_t>(nblk)), dim3(kSortBlock), 0, s, ck, n, shift, H,
                       static_cast<uint32_t>(nblk), gate, mask);
    hipLaunchKernelGGL(k_sort_scan, dim3(static_cast<uint32_t>(chunks)), dim3(kScanBlock), 0, s, H, len, S, gate,
                       mask);
    if (chunks > 1)
      hipLaunchKernelGGL(k_sort_scan_top, dim3(1), dim3(kScanBlock), 0, s, S, static_cast<uint32_t>(chunks), gate,
                         mask);
    else
      hipLaunchKernelGGL(k_sort_scan_top, dim3(1), dim3(kScanBlock), 0, s, S, 1u, gate, mask);
    hipLaunchKernelGGL(k_sort_scatter, dim3(static_cast<uint32_t>(nblk)), dim3(kSortBlock), 0, s, ck, cv, nk, nv, n,
                       shift, H, S, static_cast<uint32_t>(nblk), gate, mask);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    ck = nk;
    cv = nv;
  }
  return hipSuccess;
}

}  // namespace fm
