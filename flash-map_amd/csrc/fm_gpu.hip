// fm_gpu.hip -- GPU backend (MI355X / gfx950) of the FlashMap C ABI.
//
// Replaces the reference's GpuFlashMap (src/gpu.rs:32-712): no NVRTC JIT,
// the kernels are compiled ahead of time for gfx950; device memory is sized
// once per map; per-call scratch and staging buffers grow geometrically and
// are reused, so steady-state calls allocate nothing.  All work for a map is
// ordered on one stream (default priority).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <memory>
#include <mutex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fm_copy.h"
#include "fm_internal.h"
#include "fm_kernels.h"
#include "fm_sort.h"

namespace fm {
namespace {

using dev::Counters;
using dev::Table;

constexpr uint64_t kMaxGpuCapacity = 1ull << 31;  // u32 slot ids in scratch
constexpr uint64_t kQKeepMax = 1u << 15;          // removes up to this size keep Q_ == freebits

// ---------------------------------------------------------------------------
// Scratch of the insert that allocates nothing (fm_bulk_insert_device_noalloc):
// one copy of `scratch` per GPU (32 MiB, held until the process ends),
// shared by every such call on that GPU.  In host code the array's name is
// only a host-side stand-in; each device's own address comes from
// hipGetSymbolAddress, taken once per device at map creation with that
// device current, and every kernel argument uses that address.  Nothing in
// it needs to start at zero: every kernel writes an element before any
// kernel of the same call reads it.  Layout for up to kNoAllocMaxKeys keys
// (the led path and the sort-grouped path of one call use one or the other):
//   led:  dslot [0, 4 MiB), big [4, 8 MiB)
//   sort: home [0, 4), cid [4, 8), idx [8, 12), cid_s [12, 16),
//         idx_s [16, 20), sort scratch [20, 32 MiB)
// ---------------------------------------------------------------------------
constexpr uint64_t kNoAllocMaxKeys = 1ull << 20;
constexpr uint64_t kNoAllocArray = 4 * kNoAllocMaxKeys;
constexpr uint64_t kNoAllocBytes = 32ull << 20;
static_assert(5 * kNoAllocArray + sort_scratch_bytes(kNoAllocMaxKeys) <= kNoAllocBytes, "sort-path layout");
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

// Records this device's address of `scratch` (the device is current).
hipError_t scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> l(g_scratch_mu);
  if (g_scratch_ptr[dev]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_ptr[dev] = static_cast<uint8_t*>(p);
  return e;
}

uint8_t* scratch_ptr(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> l(g_scratch_mu);
  return g_scratch_ptr[dev];
}

// ---------------------------------------------------------------------------
// Per-(K,V) launchers.  Specialised sizes get register-resident keys and
// vector loads; everything else runs the generic byte path (K,V <= 256).
// ---------------------------------------------------------------------------
struct Ops {
  hipError_t (*get)(const Table&, const uint8_t*, uint64_t, uint8_t*, uint8_t*, uint32_t, uint32_t, hipStream_t);
  hipError_t (*ins_dry)(const Table&, const uint8_t*, uint64_t, unsigned long long*, uint32_t*, uint32_t*, Counters*,
                        uint32_t, hipStream_t);
  hipError_t (*ins_leader)(const Table&, const uint8_t*, const uint8_t*, uint64_t, const unsigned long long*,
                           const uint32_t*, uint32_t*, Counters*, uint32_t, uint32_t, bool, Counters*, uint32_t,
                           hipStream_t);
  // null for the generic sizes
  hipError_t (*ins_arcs)(const Table&, const uint8_t*, const uint8_t*, const unsigned long long*, uint32_t*,
                         Counters*, hipStream_t);
  hipError_t (*ins_big)(const Table&, const uint8_t*, const uint8_t*, uint64_t, const unsigned long long*,
                        const uint32_t*, Counters*, Counters*, uint32_t, hipStream_t);
  hipError_t (*ins_replay)(const Table&, const uint8_t*, const uint8_t*, uint64_t, const uint32_t*, const uint32_t*,
                           Counters*, uint32_t, hipStream_t);
  hipError_t (*rem_probe)(const Table&, const uint8_t*, uint64_t, unsigned long long*, uint32_t, uint32_t, hipStream_t);
  hipError_t (*rem_replay)(const Table&, const uint8_t*, uint64_t, unsigned long long*, const uint32_t*,
                           const uint32_t*, Counters*, bool, uint32_t, uint32_t, Counters*, uint32_t, hipStream_t);
  hipError_t (*ins_dedup)(const Table&, const uint8_t*, uint64_t, uint32_t, hipStream_t);
};

inline uint32_t grid_for(uint64_t threads, uint32_t block) {
  const uint64_t g = (threads + block - 1) / block;
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(g, 1u << 20)));
}

template <int KS, int VS, int G>
hipError_t launch_get_g(const Table& t, const uint8_t* q, uint64_t n, uint8_t* ov, uint8_t* of, uint32_t block,
                        hipStream_t s) {
  hipLaunchKernelGGL((dev::k_get<KS, VS, G>), dim3(grid_for(n * G, block)), dim3(block), 0, s, t, q, n, ov, of);
  return hipGetLastError();
}

template <int KS, int VS>
hipError_t launch_get(const Table& t, const uint8_t* q, uint64_t n, uint8_t* ov, uint8_t* of, uint32_t G,
                      uint32_t block, hipStream_t s) {
  switch (G) {
    case 1: return launch_get_g<KS, VS, 1>(t, q, n, ov, of, block, s);
    case 2: return launch_get_g<KS, VS, 2>(t, q, n, ov, of, block, s);
    case 4: return launch_get_g<KS, VS, 4>(t, q, n, ov, of, block, s);
    case 8: return launch_get_g<KS, VS, 8>(t, q, n, ov, of, block, s);
    default: return launch_get_g<KS, VS, 16>(t, q, n, ov, of, block, s);
  }
}

template <int KS>
hipError_t launch_ins_dry(const Table& t, const uint8_t* k, uint64_t n, unsigned long long* Q, uint32_t* home,
                          uint32_t* dslot, Counters* c, uint32_t block, hipStream_t s) {
  // with next bitmaps to fill, at most 4 words copied per thread
  const uint64_t threads = t.nfree ? std::max<uint64_t>(n, ((t.cap + 63) / 64 + 3) / 4) : n;
  hipLaunchKernelGGL((dev::k_ins_dry<KS>), dim3(grid_for(threads, block)), dim3(block), 0, s, t, k, n, Q, home, dslot,
                     c);
  return hipGetLastError();
}

template <int KS>
hipError_t launch_ins_dedup(const Table& t, const uint8_t* k, uint64_t n, uint32_t block, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_ins_dedup<KS>), dim3(grid_for(n, block)), dim3(block), 0, s, t, k, n);
  return hipGetLastError();
}

template <int KS, int VS>
hipError_t launch_ins_leader(const Table& t, const uint8_t* k, const uint8_t* v, uint64_t n,
                             const unsigned long long* Q, const uint32_t* dslot, uint32_t* big, Counters* c,
                             uint32_t coop_arc, uint32_t leader_max, bool plan, Counters* host_out, uint32_t seq,
                             hipStream_t s) {
  constexpr bool kCanPlan = KS > 0 && KS % 4 == 0 && VS > 0 && VS % 4 == 0;
  if (kCanPlan && plan)
    hipLaunchKernelGGL((dev::k_ins_leader<KS, VS, kCanPlan>), dim3(grid_for(n, 64)), dim3(64), 0, s, t, k, v, n, Q,
                       dslot, big, c, coop_arc, leader_max, host_out, seq);
  else
    hipLaunchKernelGGL((dev::k_ins_leader<KS, VS, false>), dim3(grid_for(n, 64)), dim3(64), 0, s, t, k, v, n, Q,
                       dslot, big, c, coop_arc, leader_max, host_out, seq);
  return hipGetLastError();
}

// One wave per bitmap word (cap >= 256, checked by the caller).
template <int KS, int VS>
hipError_t launch_ins_arcs(const Table& t, const uint8_t* k, const uint8_t* v, const unsigned long long* Q,
                           uint32_t* big, Counters* c, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_ins_arcs<KS, VS>), dim3(grid_for(t.cap, 1024)), dim3(1024), 0, s, t, k, v, Q, big, c);
  return hipGetLastError();
}

// One wave per queued closure; enough waves to cover ~4 per SIMD.
template <int KS, int VS>
hipError_t launch_ins_big(const Table& t, const uint8_t* k, const uint8_t* v, uint64_t n, const unsigned long long* Q,
                          const uint32_t* big, Counters* c, Counters* host_out, uint32_t seq, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_ins_big<KS, VS>), dim3(1024), dim3(64), 0, s, t, k, v, n, Q, big, c, host_out, seq);
  return hipGetLastError();
}

template <int KS, int VS>
hipError_t launch_ins_replay(const Table& t, const uint8_t* k, const uint8_t* v, uint64_t n, const uint32_t* cid,
                             const uint32_t* idx, Counters* c, uint32_t block, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_ins_replay<KS, VS>), dim3(grid_for(n, block)), dim3(block), 0, s, t, k, v, n, cid, idx,
                     c);
  return hipGetLastError();
}

template <int KS, int G>
hipError_t launch_rem_probe_g(const Table& t, const uint8_t* q, uint64_t n, unsigned long long* ev, uint32_t block,
                              hipStream_t s) {
  hipLaunchKernelGGL((dev::k_rem_probe<KS, G>), dim3(grid_for(n * G, block)), dim3(block), 0, s, t, q, n, ev);
  return hipGetLastError();
}

template <int KS>
hipError_t launch_rem_probe(const Table& t, const uint8_t* q, uint64_t n, unsigned long long* ev, uint32_t G,
                            uint32_t block, hipStream_t s) {
  switch (G) {
    case 1: return launch_rem_probe_g<KS, 1>(t, q, n, ev, block, s);
    case 2: return launch_rem_probe_g<KS, 2>(t, q, n, ev, block, s);
    case 4: return launch_rem_probe_g<KS, 4>(t, q, n, ev, block, s);
    case 8: return launch_rem_probe_g<KS, 8>(t, q, n, ev, block, s);
    default: return launch_rem_probe_g<KS, 16>(t, q, n, ev, block, s);
  }
}

template <int KS>
hipError_t launch_rem_replay(const Table& t, const uint8_t* q, uint64_t n, unsigned long long* dirty,
                             const uint32_t* druns, const uint32_t* nxt, Counters* c, bool gated, uint32_t status,
                             uint32_t nd, Counters* host_out, uint32_t seq, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_rem_replay<KS>), dim3(dev::kRemReplayBlocks), dim3(64), 0, s, t, q, n, dirty, druns, nxt,
                     c, gated, status, nd, host_out, seq);
  return hipGetLastError();
}

template <int KS, int VS>
Ops make_ops() {
  Ops o;
  o.get = &launch_get<KS, VS>;
  o.ins_dry = &launch_ins_dry<KS>;
  o.ins_leader = &launch_ins_leader<KS, VS>;
  o.ins_big = &launch_ins_big<KS, VS>;
  if constexpr (KS > 0 && VS > 0 && KS % 4 == 0 && VS % 4 == 0) o.ins_arcs = &launch_ins_arcs<KS, VS>;
  else o.ins_arcs = nullptr;
  o.ins_replay = &launch_ins_replay<KS, VS>;
  o.rem_probe = &launch_rem_probe<KS>;
  o.rem_replay = &launch_rem_replay<KS>;
  o.ins_dedup = &launch_ins_dedup<KS>;
  return o;
}

// Sizes with register-resident fast paths: the five benchmark shapes
// (u64->u64, [u8;32]->[u8;128], [u8;32]->u64, u64->[u8;32]) and the
// reference test shapes ([u8;32]->[u8;64]).
Ops select_ops(uint32_t K, uint32_t V) {
#define FM_SIZE(k, v) \
  if (K == k && V == v) return make_ops<k, v>();
  FM_SIZE(8, 8)
  FM_SIZE(32, 128)
  FM_SIZE(32, 8)
  FM_SIZE(8, 32)
  FM_SIZE(32, 64)
  FM_SIZE(32, 32)
  FM_SIZE(16, 16)
  FM_SIZE(8, 16)
#undef FM_SIZE
  return make_ops<0, 0>();
}

uint32_t vec_align(uint32_t size) {
  if (size % 16 == 0) return 16;
  if (size % 8 == 0) return 8;
  if (size % 4 == 0) return 4;
  return 1;
}

std::string hip_msg(const char* what, hipError_t e) { return std::string(what) + ": " + hipGetErrorString(e); }

class GpuBackend final : public Backend {
 public:
  GpuBackend() = default;
  ~GpuBackend() override { release(); }

  fm_status init(const fm_builder& b, std::string* err);

  fm_backend_kind kind() const override { return FM_BACKEND_GPU; }
  uint64_t capacity() const override { return t_.cap; }
  uint64_t len() const override { return len_; }
  uint32_t key_size() const override { return t_.K; }
  uint32_t value_size() const override { return t_.V; }

  fm_status bulk_get(const void* keys, uint64_t n, void* out_values, uint8_t* out_found) override;
  fm_status bulk_insert(const void* keys, const void* values, uint64_t n, uint64_t* out_new) override;
  fm_status bulk_remove(const void* keys, uint64_t n, uint64_t* out_removed) override;
  fm_status clear() override;
  fm_status export_slots(uint32_t* flags, void* keys, void* values) override;
  fm_status import_slots(const uint32_t* flags, const void* keys, const void* values, uint64_t len) override;

  void* device_stream() override { return stream_; }
  int32_t device_id() const override { return dev_; }
  fm_status alloc_device(uint64_t bytes, void** out) override;
  fm_status free_device(void* p) override;
  fm_status upload(const void* host, uint64_t bytes, void** out) override;
  fm_status download(const void* dptr, uint64_t bytes, void* host) override;
  fm_status bulk_get_device(const void* d_keys, uint64_t n, void* d_out, uint8_t* d_found) override;
  fm_status bulk_insert_device(const void* d_keys, const void* d_vals, uint64_t n, uint64_t* out_new,
                               bool counted) override;
  fm_status bulk_insert_device_noalloc(const void* d_keys, const void* d_vals, uint64_t n,
                                       uint64_t* out_new) override;
  fm_status bulk_remove_device(const void* d_keys, uint64_t n, uint64_t* out_removed) override;
  fm_status recount(uint64_t* out) override;
  fm_status synchronize() override;
  fm_status copy_from(Backend& src, uint32_t parts) override;
  fm_status serve_get(const void* d_keys, uint64_t n, void* d_out, uint8_t* d_found) override {
    fm_status st;
    if ((st = use_device()) != FM_OK) return st;
    return run_get(static_cast<const uint8_t*>(d_keys), n, static_cast<uint8_t*>(d_out), d_found);
  }
  fm_status serve_insert(const void* d_keys, const void* d_vals, uint64_t n, uint8_t* d_claimed) override {
    fm_status st;
    if ((st = use_device()) != FM_OK) return st;
    if ((st = run_insert(static_cast<const uint8_t*>(d_keys), static_cast<const uint8_t*>(d_vals), n, d_claimed)) !=
        FM_OK)
      return st;
    // No status readback here: len_ is an upper bound while serving (an
    // insert adds its whole batch, a remove subtracts once its flags are
    // back), so a led insert's dry pass cannot fail and its settle pass
    // leaves Q_ equal to freebits: the next insert needs no bitmap copy.
    q_synced_ = last_insert_ == InsertPath::kLed && q_pending_;
    len_ += n;
    return FM_OK;
  }
  void serve_settle_insert(uint64_t n, uint64_t claimed) override { len_ -= std::min(len_, n - claimed); }
  void serve_mode(bool on) override { serving_ = on; }
  fm_status serve_reserve(uint64_t max_keys) override {
    fm_status st;
    if ((st = use_device()) != FM_OK || (st = ensure_scratch(max_keys)) != FM_OK) return st;
    return ensure_dedup(max_keys);
  }
  fm_status serve_remove(const void* d_keys, uint64_t n, uint8_t* d_removed) override {
    fm_status st;
    if ((st = use_device()) != FM_OK) return st;
    return run_remove(static_cast<const uint8_t*>(d_keys), n, nullptr, d_removed, true);
  }
  void serve_settle_remove(uint64_t removed) override { len_ -= std::min(len_, removed); }
  fm_status profile_enable(bool on) override;
  fm_status profile_only(const char* name) override;
  fm_status profile_read(fm_kernel_stat* out, uint32_t max, uint32_t* count) override;
  fm_status get_tuning(fm_tuning* out) const override {
    *out = tune_;
    return FM_OK;
  }
  fm_status set_tuning(const fm_tuning* t) override;

 private:
  fm_status fail(fm_status code, const std::string& msg) {
    set_error(code, msg);
    return code;
  }
  fm_status check_launch(hipError_t e, const char* what) {
    return e == hipSuccess ? FM_OK : fail(FM_ERR_KERNEL_LAUNCH, hip_msg(what, e));
  }
  fm_status use_device() {
    hipError_t e = hipSetDevice(dev_);
    return e == hipSuccess ? FM_OK : fail(FM_ERR_DEVICE_INIT, hip_msg("hipSetDevice", e));
  }
  fm_status ensure_scratch(uint64_t n);
  fm_status ensure_staging(uint64_t bytes_keys, uint64_t bytes_vals, uint64_t bytes_out, uint64_t bytes_found);
  fm_status read_counters(Counters* out);
  bool host_seq_arrived(uint32_t seq) const;
  uint32_t next_ctr_seq() {  // never 0, the value a copied-back op leaves in h_ctr_->done
    ctr_seq_ = ctr_seq_ == UINT32_MAX ? 1 : ctr_seq_ + 1;
    return ctr_seq_;
  }
  fm_status zero_counters();
  fm_status check_device_ptr(const void* p, uint32_t align, const char* what);
  fm_status check_insert_status(const Counters& c) {
    if (c.status & dev::kStatusBigPending)  // the host owes a k_ins_big and a settle: counts are partial
      return fail(FM_ERR_KERNEL_LAUNCH, "insert: closures still queued for the big-closure replay");
    if (c.status & dev::kStatusBudget)
      return fail(FM_ERR_KERNEL_LAUNCH, "insert walk found no free slot in a full table (uncounted inserts overfilled it)");
    if ((c.status & dev::kStatusDryFail) && last_insert_ != InsertPath::kSort) {
      // Unreachable while len never undercounts occupied slots; the leader
      // kernels skipped the batch, so only owner[] (and Q) need restoring.
      q_synced_ = false;
      if (last_insert_ == InsertPath::kLed) {
        hipLaunchKernelGGL(dev::k_fill_u32, dim3(grid_for(t_.cap, 256)), dim3(256), 0, stream_, t_.owner, t_.cap,
                           dev::kNoOwner);
        (void)hipStreamSynchronize(stream_);
      }
      return fail(FM_ERR_KERNEL_LAUNCH, "insert: free-slot accounting disagrees with the table; batch not applied");
    }
    return FM_OK;
  }

  // Device pipelines (inputs on the device, ordered on stream_).
  fm_status run_get(const uint8_t* d_keys, uint64_t n, uint8_t* d_out, uint8_t* d_found);
  // leaves count in ctr.  defer_big (counted calls, which then read the
  // counters with read_insert_counters): the led path launches no k_ins_big;
  // its settle pass reports queued closures and the host runs them then.
  // Per-call buffers of an insert: the map's growable scratch, or the
  // per-GPU `scratch` array for the insert that allocates nothing.
  struct InsertScratch {
    uint32_t* home;
    uint32_t* dslot;
    uint32_t* big;
    uint32_t* cid;
    uint32_t* idx;
    uint32_t* cid_s;
    uint32_t* idx_s;
    void* sort_tmp;
  };
  // fixed: use these buffers (no allocation, no deduplication).
  fm_status run_insert(const uint8_t* d_keys, const uint8_t* d_vals, uint64_t n, uint8_t* outcome = nullptr,
                       bool defer_big = false, const InsertScratch* fixed = nullptr);
  // early_seq != 0: the leader hands the counters over with that sequence
  // number, and the settle pass only settles (k_ins_bits `early`).
  fm_status launch_bits(const Table& t, uint64_t n, bool dedup, bool defer_big, const uint32_t* dslot,
                        uint32_t early_seq = 0);
  fm_status read_insert_counters(Counters* out);
  // async: no status readback and no count (serving reads the outcome
  // flags); the conflict replay is gated on the device.
  fm_status run_remove(const uint8_t* d_keys, uint64_t n, uint64_t* removed, uint8_t* outcome = nullptr,
                       bool async = false);

  void release();

  int dev_ = -1;
  hipStream_t stream_ = nullptr;
  Table t_{};
  uint64_t len_ = 0;
  Ops ops_{};
  fm_tuning tune_{4, 256, 4, 256, 64, 0, 0};
  bool serving_ = false;  // inside fm_serve_log: inserts deduplicate as with insert_dedup
  // insert deduplication buckets (allocated on first use)
  uint64_t dd_s_ = 0;
  uint32_t* dd_first_ = nullptr;
  uint32_t* dd_last_ = nullptr;
  uint32_t* dd_bucket_ = nullptr;
  uint64_t dd_n_ = 0;
  fm_status ensure_dedup(uint64_t n);
  bool tune_sort_only_ = false;  // FM_INSERT_SORT_PATH=1 forces the sort-grouped insert
  uint32_t leader_max_ = dev::kLeaderMax;  // FM_LEADER_MAX: members a leader lane replays (more: k_ins_big)
  bool arcs_on_ = true;                    // FM_INSERT_ARCS=0: dense batches led by k_ins_leader instead
  bool next_bits_on_ = true;               // FM_INSERT_NEXT_BITS=0: sparse batches settled by k_ins_bits
  enum class InsertPath { kSort, kLed };
  InsertPath last_insert_ = InsertPath::kSort;
  // Stream-ordered facts about device state, kept by the host:
  bool q_synced_ = false;      // Q_ == freebits
  bool q_pending_ = false;     // the last led insert restored Q_; valid once its status is clean
  bool ctr_zero_ = false;      // ctr_ holds zeros
  bool ctr_in_host_ = false;   // the last op left its final counters in h_ctr_
  uint32_t ctr_seq_ = 0;       // sequence number of the last counted insert's hand-over
  uint32_t poll_seq_ = 0;      // nonzero: h_ctr_->done takes this value when the counters are there
  struct PendingBig {  // the last deferred led insert, for its k_ins_big and settle
    Table t;
    const uint8_t* keys;
    const uint8_t* vals;
    uint64_t n;
    bool dedup;
    const uint32_t* dslot;
    uint32_t* big;
    unsigned long long* Q;  // the batch's claim bitmap
    bool next_bits;         // claims went to the next bitmaps (Table::nfree): nothing left to settle
  };
  PendingBig pend_{};
  // The spare bitmap set of inserts that publish their claims into the next
  // bitmaps (Table::nfree, nempty, nq); swapped with freebits, emptybits and
  // Q_ after such a batch.  Allocated on first use.
  unsigned long long* nfree_ = nullptr;
  unsigned long long* nempty_ = nullptr;
  unsigned long long* nq_ = nullptr;
  // Each set is one block of three words-long arrays (free, empty, claim),
  // so a kernel reaches the other two from the first (Table::nfree).
  unsigned long long* bits_a_ = nullptr;  // the set allocated with the map
  unsigned long long* bits_b_ = nullptr;  // the spare set
  bool ensure_next_bits();

  // scratch (grown geometrically, reused)
  uint64_t scratch_n_ = 0;
  uint32_t* home_ = nullptr;
  uint32_t* dslot_ = nullptr;
  uint32_t* big_ = nullptr;
  uint32_t* cid_ = nullptr;
  uint32_t* idx_ = nullptr;
  uint32_t* cid_s_ = nullptr;
  uint32_t* idx_s_ = nullptr;
  unsigned long long* ev_ = nullptr;
  uint8_t* action_ = nullptr;
  void* sort_tmp_ = nullptr;
  size_t sort_tmp_bytes_ = 0;
  unsigned long long* Q_ = nullptr;
  unsigned long long* dirty_ = nullptr;  // remove conflicts: runs to replay, all zero between ops
  Counters* ctr_ = nullptr;
  Counters* h_ctr_ = nullptr;  // pinned

  // per-kernel event timing
  enum {
    kKidGet, kKidInsCopy, kKidInsDry, kKidInsClosure, kKidInsSort, kKidInsReplay, kKidRemProbe, kKidRemCheck,
    kKidRemApply, kKidRemRuns, kKidRemSort, kKidRemReplay, kKidCopyTable, kKidInsLeader, kKidInsBig, kKidInsBits,
    kKidInsDedup, kKidInsArcs, kNumKid
  };
  struct PendingTiming {
    int kid;
    hipEvent_t a, b;
  };
  struct Prof {  // brackets the launches of one scope when profiling is on
    GpuBackend* g;
    int kid;
    Prof(GpuBackend* g_, int k) : g(g_), kid(k) {
      if (g->timed(kid)) g->prof_mark(kid, true);
    }
    ~Prof() {
      if (g->timed(kid)) g->prof_mark(kid, false);
    }
  };
  static const char* kid_name(int k);
  hipEvent_t take_event();
  void prof_mark(int kid, bool begin);
  void harvest();
  bool prof_ = false;
  int prof_only_ = -1;    // >= 0: time this kernel id only (fm_profile_only)
  bool timed(int kid) const { return prof_ && !noalloc_ && (prof_only_ < 0 || prof_only_ == kid); }
  bool noalloc_ = false;  // inside the insert that allocates nothing: no timing events either
  hipError_t scratch_err_ = hipSuccess;  // hipGetSymbolAddress of `scratch` at map creation
  hipEvent_t prof_open_ = nullptr;
  std::vector<PendingTiming> pending_;
  std::vector<hipEvent_t> ev_pool_;
  double prof_ms_[kNumKid] = {};
  uint64_t prof_n_[kNumKid] = {};
  int sort_end_bit_ = 1;

  // host-API staging
  uint8_t* st_keys_ = nullptr;
  uint64_t st_keys_bytes_ = 0;
  uint8_t* st_vals_ = nullptr;
  uint64_t st_vals_bytes_ = 0;
  uint8_t* st_out_ = nullptr;
  uint64_t st_out_bytes_ = 0;
  uint8_t* st_found_ = nullptr;
  uint64_t st_found_bytes_ = 0;
  // Host buffers of large calls: pageable memory goes through pinned
  // buffers with the copy pool (buffers from fm_host_alloc are moved by DMA
  // directly); small calls use the runtime's pageable copies.
  static constexpr uint64_t kPinnedMin = 1ull << 20;   // bytes per call, in + out
  static constexpr uint64_t kPinnedMax = 1ull << 30;   // largest pinned buffer kept
  static constexpr unsigned kCopyWorkers = 8;
  std::unique_ptr<CopyPool> copies_;
  uint8_t* h_in_ = nullptr;
  uint64_t h_in_bytes_ = 0;
  uint8_t* h_out_ = nullptr;
  uint64_t h_out_bytes_ = 0;
  // Grows the pinned buffers; false when a call should keep the pageable path.
  bool ensure_host_staging(uint64_t in, uint64_t out);
  // Staged transfers move in pieces, so the copy pool fills (or empties) one
  // piece of the pinned buffer while the DMA engine moves the one before it:
  // a call costs the DMA time plus one piece of host copying instead of the
  // two one after the other.
  struct HostXfer {
    uint8_t* dev;
    uint8_t* host;  // source of stage_in, destination of stage_out
    uint64_t bytes;
  };
  static constexpr unsigned kStageXfers = 2;  // transfers per staged call at most
  static constexpr unsigned kStagePieces = 16;
  static constexpr uint64_t kStagePieceMin = 4ull << 20;
  hipEvent_t stage_ev_[kStagePieces + kStageXfers] = {};  // one per piece of stage_out
  static uint64_t stage_piece(const HostXfer* xs, unsigned nx);
  // Host -> h_in_ -> device on stream_; returns once the last piece is queued.
  fm_status stage_in(const HostXfer* xs, unsigned nx);
  // Device -> h_out_ -> host after the work queued on stream_; returns when
  // every byte is in place.
  fm_status stage_out(const HostXfer* xs, unsigned nx);
};

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void GpuBackend::release() {
  if (dev_ < 0) return;
  (void)hipSetDevice(dev_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  dfree(t_.keys);
  dfree(t_.vals);
  dfree(t_.flags);
  dfree(bits_a_);
  t_.freebits = t_.emptybits = nullptr;
  dfree(t_.owner);
  dfree(t_.rowner);
  dfree(home_);
  dfree(dslot_);
  dfree(big_);
  dfree(cid_);
  dfree(idx_);
  dfree(cid_s_);
  dfree(idx_s_);
  dfree(ev_);
  dfree(action_);
  if (sort_tmp_) (void)hipFree(sort_tmp_);
  sort_tmp_ = nullptr;
  dfree(bits_b_);
  nfree_ = nempty_ = nq_ = Q_ = nullptr;
  dfree(dirty_);
  dfree(ctr_);
  dfree(dd_first_);
  dfree(dd_last_);
  dfree(dd_bucket_);
  if (h_ctr_) (void)hipHostFree(h_ctr_);
  h_ctr_ = nullptr;
  dfree(st_keys_);
  dfree(st_vals_);
  dfree(st_out_);
  dfree(st_found_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (h_in_) (void)hipHostFree(h_in_);
  if (h_out_) (void)hipHostFree(h_out_);
  h_in_ = h_out_ = nullptr;
  h_in_bytes_ = h_out_bytes_ = 0;
  for (hipEvent_t& e : stage_ev_) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  copies_.reset();
  pending_.clear();
  for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
  ev_pool_.clear();
  if (prof_open_) (void)hipEventDestroy(prof_open_);
  prof_open_ = nullptr;
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
  dev_ = -1;
}

fm_status GpuBackend::init(const fm_builder& b, std::string* err) {
  // GpuFlashMap::new (gpu.rs:45-105)
  if (b.capacity == 0) {
    *err = "capacity must be positive";
    return FM_ERR_ZERO_CAPACITY;
  }
  if (b.key_size == 0 || b.key_size > dev::kMaxGenericSize || b.value_size == 0 ||
      b.value_size > dev::kMaxGenericSize) {
    *err = "key/value size outside 1..256 bytes";
    return FM_ERR_INVALID_ARGUMENT;
  }
  const uint64_t cap = round_capacity(b.capacity);
  if (cap > kMaxGpuCapacity) {
    *err = "capacity above 2^31 slots";
    return FM_ERR_INVALID_ARGUMENT;
  }
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    *err = e != hipSuccess ? hip_msg("hipGetDeviceCount", e) : std::string("no HIP device");
    return FM_ERR_DEVICE_INIT;
  }
  if (b.device_id < 0 || b.device_id >= count) {
    *err = "device ordinal " + std::to_string(b.device_id) + " out of range";
    return FM_ERR_DEVICE_INIT;
  }
  dev_ = b.device_id;
  if ((e = hipSetDevice(dev_)) != hipSuccess) {
    *err = hip_msg("hipSetDevice", e);
    return FM_ERR_DEVICE_INIT;
  }
  if ((e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking)) != hipSuccess) {
    *err = hip_msg("hipStreamCreate", e);
    return FM_ERR_DEVICE_INIT;
  }
  t_.cap = cap;
  t_.mask = cap - 1;
  t_.K = b.key_size;
  t_.V = b.value_size;
  t_.mode = b.hash_strategy == FM_HASH_MURMUR3 ? 1u : 0u;
  const uint64_t words = (cap + 63) / 64;
  // Align per-slot arrays for 16-byte vector access.
  if ((e = hipMalloc(&t_.keys, cap * t_.K + 16)) != hipSuccess ||
      (e = hipMalloc(&t_.vals, cap * t_.V + 16)) != hipSuccess ||
      (e = hipMalloc(&t_.flags, cap * sizeof(uint32_t))) != hipSuccess ||
      (e = hipMalloc(&bits_a_, 3 * words * 8)) != hipSuccess ||
      (e = hipMalloc(&t_.owner, cap * sizeof(uint32_t))) != hipSuccess ||
      (e = hipMalloc(&t_.rowner, cap * sizeof(uint32_t))) != hipSuccess ||
      (e = hipMalloc(&dirty_, words * 8)) != hipSuccess ||
      (e = hipMemset(dirty_, 0, words * 8)) != hipSuccess ||
      (e = hipMalloc(&ctr_, sizeof(Counters))) != hipSuccess ||
      (e = hipHostMalloc(&h_ctr_, sizeof(Counters), hipHostMallocDefault)) != hipSuccess) {
    *err = hip_msg("hipMalloc (table)", e);
    return FM_ERR_GPU_ALLOC;
  }
  t_.freebits = bits_a_;
  t_.emptybits = bits_a_ + words;
  Q_ = bits_a_ + 2 * words;
  std::memset(h_ctr_, 0, sizeof(Counters));  // done == 0 matches no sequence number
  scratch_err_ = scratch_setup(dev_);  // checked by the insert that allocates nothing
  ops_ = select_ops(t_.K, t_.V);
  if (const char* env = std::getenv("FM_INSERT_SORT_PATH")) tune_sort_only_ = env[0] == '1';
  if (const char* env = std::getenv("FM_INSERT_ARCS")) arcs_on_ = std::strcmp(env, "0") != 0;
  if (const char* env = std::getenv("FM_INSERT_NEXT_BITS")) next_bits_on_ = std::strcmp(env, "0") != 0;
  // Arcs of at least this many members replayed from their groups when
  // canonical (0: never).
  t_.canon_min = 32;
  if (const char* env = std::getenv("FM_INSERT_CANON_MIN"))
    t_.canon_min = static_cast<uint32_t>(std::strtoul(env, nullptr, 10));
  if (const char* env = std::getenv("FM_LEADER_MAX"))
    leader_max_ = static_cast<uint32_t>(std::max(1ul, std::min<unsigned long>(std::strtoul(env, nullptr, 10), dev::kLeaderMax)));
  sort_end_bit_ = 1;
  while ((1ull << sort_end_bit_) <= cap) ++sort_end_bit_;  // ids in [0, cap]
  hipLaunchKernelGGL(dev::k_clear, dim3(grid_for(cap, 256)), dim3(256), 0, stream_, t_, Q_);
  q_synced_ = true;  // k_clear filled Q_ like freebits
  hipLaunchKernelGGL(dev::k_fill_u32, dim3(grid_for(cap, 256)), dim3(256), 0, stream_, t_.owner, cap, dev::kNoOwner);
  hipLaunchKernelGGL(dev::k_fill_u32, dim3(grid_for(cap, 256)), dim3(256), 0, stream_, t_.rowner, cap, dev::kNoOwner);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(stream_)) != hipSuccess) {
    *err = hip_msg("table init", e);
    return FM_ERR_DEVICE_INIT;
  }
  len_ = 0;
  return FM_OK;
}

fm_status GpuBackend::set_tuning(const fm_tuning* t) {
  auto pow2 = [](uint32_t g) { return g >= 1 && g <= 16 && (g & (g - 1)) == 0; };
  auto blk = [](uint32_t b) { return b >= 64 && b <= 1024 && b % 64 == 0; };
  if (!t || !pow2(t->get_group) || !pow2(t->remove_group) || !blk(t->get_block) || !blk(t->insert_block) ||
      t->insert_fused != 0 || t->insert_dedup > 1)
    return fail(FM_ERR_INVALID_ARGUMENT,
                "tuning: group must be 1,2,4,8,16; block a multiple of 64 in 64..1024; insert_fused 0 (the "
                "cooperative insert was removed); insert_dedup 0 or 1");
  tune_ = *t;
  return FM_OK;
}

fm_status GpuBackend::ensure_scratch(uint64_t n) {
  if (n <= scratch_n_) return FM_OK;
  uint64_t want = std::max<uint64_t>(n, std::max<uint64_t>(1024, scratch_n_ * 2));
  if (want > 0x7FFFFFFFull) want = std::max<uint64_t>(n, 0x7FFFFFFFull);
  if (n > 0x7FFFFFFFull) return fail(FM_ERR_INVALID_ARGUMENT, "batch above 2^31-1 keys");
  (void)hipStreamSynchronize(stream_);
  dfree(home_);
  dfree(dslot_);
  dfree(big_);
  dfree(cid_);
  dfree(idx_);
  dfree(cid_s_);
  dfree(idx_s_);
  dfree(ev_);
  dfree(action_);
  if (sort_tmp_) (void)hipFree(sort_tmp_);
  sort_tmp_ = nullptr;
  scratch_n_ = 0;
  hipError_t e;
  if ((e = hipMalloc(&home_, want * 4)) != hipSuccess || (e = hipMalloc(&cid_, want * 4)) != hipSuccess ||
      (e = hipMalloc(&dslot_, want * 4)) != hipSuccess || (e = hipMalloc(&big_, want * 4)) != hipSuccess ||
      (e = hipMalloc(&idx_, want * 4)) != hipSuccess || (e = hipMalloc(&cid_s_, want * 4)) != hipSuccess ||
      (e = hipMalloc(&idx_s_, want * 4)) != hipSuccess || (e = hipMalloc(&ev_, want * 8)) != hipSuccess ||
      (e = hipMalloc(&action_, want)) != hipSuccess)
    return fail(FM_ERR_GPU_ALLOC, hip_msg("hipMalloc (scratch)", e));
  const uint64_t bytes = sort_scratch_bytes(want);
  if ((e = hipMalloc(&sort_tmp_, bytes + 256)) != hipSuccess)
    return fail(FM_ERR_GPU_ALLOC, hip_msg("hipMalloc (sort scratch)", e));
  sort_tmp_bytes_ = bytes + 256;
  scratch_n_ = want;
  return FM_OK;
}

fm_status GpuBackend::ensure_staging(uint64_t bk, uint64_t bv, uint64_t bo, uint64_t bf) {
  auto grow = [&](uint8_t*& p, uint64_t& have, uint64_t need) -> hipError_t {
    if (need <= have) return hipSuccess;
    (void)hipStreamSynchronize(stream_);
    dfree(p);
    have = 0;
    const uint64_t want = std::max<uint64_t>(need, 4096);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) have = want;
    return e;
  };
  hipError_t e;
  if ((e = grow(st_keys_, st_keys_bytes_, bk)) != hipSuccess || (e = grow(st_vals_, st_vals_bytes_, bv)) != hipSuccess ||
      (e = grow(st_out_, st_out_bytes_, bo)) != hipSuccess || (e = grow(st_found_, st_found_bytes_, bf)) != hipSuccess)
    return fail(FM_ERR_GPU_ALLOC, hip_msg("hipMalloc (staging)", e));
  return FM_OK;
}

bool GpuBackend::ensure_host_staging(uint64_t in, uint64_t out) {
  if (in > kPinnedMax || out > kPinnedMax) return false;
  auto grow = [&](uint8_t*& p, uint64_t& have, uint64_t need) -> bool {
    if (need <= have) return true;
    (void)hipStreamSynchronize(stream_);
    if (p) (void)hipHostFree(p);
    p = nullptr;
    have = 0;
    const uint64_t want = std::min<uint64_t>(kPinnedMax, std::max<uint64_t>(need, 4u << 20));
    if (hipHostMalloc(reinterpret_cast<void**>(&p), want, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return false;
    }
    have = want;
    return true;
  };
  if (!grow(h_in_, h_in_bytes_, in) || !grow(h_out_, h_out_bytes_, out)) return false;
  for (hipEvent_t& e : stage_ev_)
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      e = nullptr;
      return false;
    }
  if (!copies_) copies_ = std::make_unique<CopyPool>(kCopyWorkers);
  return true;
}

uint64_t GpuBackend::stage_piece(const HostXfer* xs, unsigned nx) {
  uint64_t total = 0;
  for (unsigned i = 0; i < nx; ++i) total += xs[i].bytes;
  const uint64_t even = (total + kStagePieces - 1) / kStagePieces;
  return (std::max(kStagePieceMin, even) + 0xFFFF) & ~0xFFFFull;  // nx transfers: at most kStagePieces + nx pieces
}

fm_status GpuBackend::stage_in(const HostXfer* xs, unsigned nx) {
  const uint64_t piece = stage_piece(xs, nx);
  uint64_t off = 0;
  for (unsigned i = 0; i < nx; ++i)
    for (uint64_t o = 0; o < xs[i].bytes; o += piece) {
      const uint64_t len = std::min(piece, xs[i].bytes - o);
      copies_->run({{h_in_ + off, xs[i].host + o, len}});
      hipError_t e = hipMemcpyAsync(xs[i].dev + o, h_in_ + off, len, hipMemcpyHostToDevice, stream_);
      if (e != hipSuccess) return fail(FM_ERR_TRANSFER, hip_msg("H2D (staged)", e));
      off += len;
    }
  return FM_OK;
}

fm_status GpuBackend::stage_out(const HostXfer* xs, unsigned nx) {
  const uint64_t piece = stage_piece(xs, nx);
  hipError_t e = hipSuccess;
  unsigned k = 0;
  uint64_t off = 0;
  for (unsigned i = 0; i < nx && e == hipSuccess; ++i)
    for (uint64_t o = 0; o < xs[i].bytes && e == hipSuccess; o += piece, ++k) {
      const uint64_t len = std::min(piece, xs[i].bytes - o);
      if ((e = hipMemcpyAsync(h_out_ + off, xs[i].dev + o, len, hipMemcpyDeviceToHost, stream_)) == hipSuccess)
        e = hipEventRecord(stage_ev_[k], stream_);
      off += len;
    }
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(stream_);  // nothing queued here may still write h_out_ after the call
    return fail(FM_ERR_TRANSFER, hip_msg("D2H (staged)", e));
  }
  k = 0;
  off = 0;
  for (unsigned i = 0; i < nx; ++i)
    for (uint64_t o = 0; o < xs[i].bytes; o += piece, ++k) {
      const uint64_t len = std::min(piece, xs[i].bytes - o);
      if ((e = hipEventSynchronize(stage_ev_[k])) != hipSuccess) {
        (void)hipStreamSynchronize(stream_);
        return fail(FM_ERR_TRANSFER, hip_msg("D2H (staged)", e));
      }
      copies_->run({{xs[i].host + o, h_out_ + off, len}});
      off += len;
    }
  return FM_OK;
}

bool GpuBackend::ensure_next_bits() {
  if (nfree_) return true;
  const uint64_t words = (t_.cap + 63) / 64;
  if (hipMalloc(&bits_b_, 3 * words * 8) != hipSuccess) {
    (void)hipGetLastError();  // no spare set: the settle pass publishes instead
    bits_b_ = nullptr;
    return false;
  }
  nfree_ = bits_b_;
  nempty_ = bits_b_ + words;
  nq_ = bits_b_ + 2 * words;
  return true;
}

// Zeroes the counters for an op that reads them back with read_counters.
fm_status GpuBackend::zero_counters() {
  hipError_t e = ctr_zero_ ? hipSuccess : hipMemsetAsync(ctr_, 0, sizeof(Counters), stream_);
  ctr_zero_ = false;  // the op about to run accumulates into them
  ctr_in_host_ = false;
  poll_seq_ = 0;
  return e == hipSuccess ? FM_OK : fail(FM_ERR_KERNEL_LAUNCH, hip_msg("hipMemsetAsync", e));
}

// The settle pass of a led insert: publishes the real claims, restores Q_
// and hands the counters to h_ctr_ (see k_ins_bits).
fm_status GpuBackend::launch_bits(const Table& t, uint64_t n, bool dedup, bool defer_big, const uint32_t* dslot,
                                  uint32_t early_seq) {
  Prof p(this, kKidInsBits);
  const bool early = early_seq != 0;
  const uint32_t seq = early ? early_seq : next_ctr_seq();
  if (!dedup && n * 16 >= t_.cap) {
    // Most bitmap words change: recompute them all from the flags (one
    // coalesced pass) instead of one atomic per key; at most 1024
    // workgroups, so at most 1024 completion tickets.
    const uint32_t g = std::min<uint32_t>(grid_for(t_.cap, 1024), 1024);
    // (never deferred: run_insert defers only batches settled per key)
    hipLaunchKernelGGL(dev::k_ins_bits_rebuild, dim3(g), dim3(1024), 0, stream_, t, ctr_, Q_, h_ctr_, seq);
  } else {
    // 1024-thread workgroups: one completion ticket per 1024 keys
    hipLaunchKernelGGL(dev::k_ins_bits, dim3(grid_for(n, 1024)), dim3(1024), 0, stream_, t, dslot, n, ctr_, Q_,
                       early ? nullptr : h_ctr_, seq, defer_big, early);
  }
  fm_status st;
  if ((st = check_launch(hipGetLastError(), "k_ins_bits")) != FM_OK) return st;
  poll_seq_ = seq;
  q_pending_ = true;     // Q_ == freebits once the status shows no failed dry pass
  // A settle that hands over zeroes the device counters; a deferred one
  // that reports pending closures keeps them (count, queue) for the settle
  // after k_ins_big, so they count as zero only once read_insert_counters
  // has seen a final hand-over.
  ctr_zero_ = !defer_big;
  ctr_in_host_ = true;   // the counters (or the pending report) go to h_ctr_
  return FM_OK;
}

// read_counters for a counted insert launched with defer_big: when the
// settle pass reports closures queued for k_ins_big, runs k_ins_big and the
// settle now (the device counters still hold the count and the queue) and
// reads the final counters.
fm_status GpuBackend::read_insert_counters(Counters* out) {
  fm_status st;
  auto failed = [&](fm_status e) {  // the device counters' state is unknown: clear them next time
    ctr_zero_ = q_synced_ = false;
    return e;
  };
  const bool handed = ctr_in_host_;  // the settle pass hands the counters over (the sort path does not)
  if ((st = read_counters(out)) != FM_OK) return failed(st);
  if (!(out->status & dev::kStatusBigPending)) {
    if (handed) ctr_zero_ = true;  // the hand-over zeroed them
    return FM_OK;
  }
  {
    Prof p(this, kKidInsBig);
    // claims in the next bitmaps: k_ins_big hands the counters over itself
    const uint32_t seq = pend_.next_bits ? next_ctr_seq() : 0;
    if ((st = check_launch(ops_.ins_big(pend_.t, pend_.keys, pend_.vals, pend_.n, pend_.Q, pend_.big, ctr_,
                                        pend_.next_bits ? h_ctr_ : nullptr, seq, stream_),
                           "k_ins_big")) != FM_OK)
      return failed(st);
    if (pend_.next_bits) {
      poll_seq_ = seq;
      ctr_in_host_ = true;
    }
  }
  // otherwise the settle pass and its hand-over
  if (!pend_.next_bits && (st = launch_bits(pend_.t, pend_.n, pend_.dedup, false, pend_.dslot)) != FM_OK)
    return failed(st);
  if ((st = read_counters(out)) != FM_OK) return failed(st);
  ctr_zero_ = !(out->status & dev::kStatusBigPending);
  return FM_OK;
}

// Waits for h_ctr_->done == seq by polling host memory, for at most
// kPollLimit; false (the caller then synchronises the stream) if it did not
// arrive.  A sleeping stream wait wakes the host well after the kernel's
// last store; a poll sees the hand-over as it lands.
bool GpuBackend::host_seq_arrived(uint32_t seq) const {
  // Up to 2 ms of one host core per counted op (an op longer than that is
  // waited for by the runtime's sleeping wait instead, whose late wake-up is
  // small next to it).
  constexpr auto kPollLimit = std::chrono::milliseconds(2);
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t i = 0;; ++i) {
    if (__atomic_load_n(&h_ctr_->done, __ATOMIC_ACQUIRE) == seq) return true;
    if ((i & 255u) == 255u && std::chrono::steady_clock::now() - t0 > kPollLimit) return false;
#if defined(__x86_64__)
    __builtin_ia32_pause();
#endif
  }
}

fm_status GpuBackend::read_counters(Counters* out) {
  if (ctr_in_host_ && poll_seq_ && host_seq_arrived(poll_seq_)) {
    // The last kernel handed the counters over: every table write of the op
    // precedes that release, and later ops are ordered behind it on stream_.
    ctr_in_host_ = false;
    poll_seq_ = 0;
    const volatile Counters* h = h_ctr_;
    out->status = h->status;
    out->nbig = h->nbig;
    out->count = h->count;
    out->done = out->nsmall = 0;
    return FM_OK;
  }
  hipError_t e = hipSuccess;
  if (!ctr_in_host_) e = hipMemcpyAsync(h_ctr_, ctr_, sizeof(Counters), hipMemcpyDeviceToHost, stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  ctr_in_host_ = false;
  poll_seq_ = 0;
  if (e != hipSuccess) return fail(FM_ERR_TRANSFER, hip_msg("counter readback", e));
  *out = *h_ctr_;
  return FM_OK;
}

fm_status GpuBackend::check_device_ptr(const void* p, uint32_t align, const char* what) {
  if (!p) return fail(FM_ERR_INVALID_ARGUMENT, std::string(what) + ": null device pointer");
  if (reinterpret_cast<uintptr_t>(p) % align != 0)
    return fail(FM_ERR_INVALID_ARGUMENT, std::string(what) + ": device buffer not " + std::to_string(align) +
                                             "-byte aligned");
  return FM_OK;
}

fm_status GpuBackend::run_get(const uint8_t* d_keys, uint64_t n, uint8_t* d_out, uint8_t* d_found) {
  Prof p(this, kKidGet);
  return check_launch(ops_.get(t_, d_keys, n, d_out, d_found, tune_.get_group, tune_.get_block, stream_), "k_get");
}

// Buckets for insert deduplication: 2^k >= 2n entries, first = kNoOwner and
// last = 0 between ops (k_ins_bits / k_dedup_reset restore the used ones).
fm_status GpuBackend::ensure_dedup(uint64_t n) {
  uint64_t want = 1024;
  while (want < 2 * n) want <<= 1;
  hipError_t e = hipSuccess;
  if (want > dd_s_) {
    (void)hipStreamSynchronize(stream_);
    dfree(dd_first_);
    dfree(dd_last_);
    dd_s_ = 0;
    if ((e = hipMalloc(&dd_first_, want * 4)) != hipSuccess || (e = hipMalloc(&dd_last_, want * 4)) != hipSuccess)
      return fail(FM_ERR_GPU_ALLOC, hip_msg("hipMalloc (dedup)", e));
    hipLaunchKernelGGL(dev::k_fill_u32, dim3(grid_for(want, 256)), dim3(256), 0, stream_, dd_first_, want,
                       dev::kNoOwner);
    hipLaunchKernelGGL(dev::k_fill_u32, dim3(grid_for(want, 256)), dim3(256), 0, stream_, dd_last_, want, 0u);
    if ((e = hipGetLastError()) != hipSuccess) return fail(FM_ERR_KERNEL_LAUNCH, hip_msg("dedup init", e));
    dd_s_ = want;
  }
  if (n > dd_n_) {
    (void)hipStreamSynchronize(stream_);
    dfree(dd_bucket_);
    dd_n_ = 0;
    const uint64_t m = std::max<uint64_t>(n, 1024);
    if ((e = hipMalloc(&dd_bucket_, m * 4)) != hipSuccess)
      return fail(FM_ERR_GPU_ALLOC, hip_msg("hipMalloc (dedup)", e));
    dd_n_ = m;
  }
  return FM_OK;
}

fm_status GpuBackend::run_insert(const uint8_t* d_keys, const uint8_t* d_vals, uint64_t n, uint8_t* outcome,
                                 bool defer_big, const InsertScratch* fixed) {
  fm_status st;
  Table t = t_;
  t.outcome = outcome;
  InsertScratch sc;
  if (fixed) {
    sc = *fixed;
  } else {
    if ((st = ensure_scratch(n)) != FM_OK) return st;
    sc = InsertScratch{home_, dslot_, big_, cid_, idx_, cid_s_, idx_s_, sort_tmp_};
  }
  const uint64_t words = (t_.cap + 63) / 64;
  const uint32_t blk = tune_.insert_block;
  // (deduplication changes no result; its buckets are allocated per size)
  const bool dedup = !fixed && (tune_.insert_dedup || serving_) && n > 1;
  if (dedup) {
    if ((st = ensure_dedup(n)) != FM_OK) return st;
    t.dd_first = dd_first_;
    t.dd_last = dd_last_;
    t.dd_bucket = dd_bucket_;
    t.dd_smask = dd_s_ - 1;
  }
  // With len + n < cap the dry pass provably leaves a free slot (len never
  // undercounts occupied slots), so closures are bounded by barriers and are
  // led without sorting.  Otherwise (the batch may fill the table) group by
  // closure id with a stable radix sort, which also covers "no barrier".
  const bool led = len_ + n < t_.cap && !tune_sort_only_;
  if ((st = zero_counters()) != FM_OK) return st;
  // Q_ is this op's claim bitmap: a copy of freebits, unless the last led
  // insert (or a remove after it) left it equal to freebits already.
  if (!(q_synced_ && led)) {
    Prof p(this, kKidInsCopy);
    hipError_t e = hipMemcpyAsync(Q_, t_.freebits, words * 8, hipMemcpyDeviceToDevice, stream_);
    if (e != hipSuccess) return fail(FM_ERR_TRANSFER, hip_msg("bitmap copy", e));
  }
  q_synced_ = false;
  q_pending_ = false;
  if (dedup) {
    Prof p(this, kKidInsDedup);
    if ((st = check_launch(ops_.ins_dedup(t, d_keys, n, blk, stream_), "k_ins_dedup")) != FM_OK) return st;
  }
  last_insert_ = led ? InsertPath::kLed : InsertPath::kSort;
  if (led) {
    // Sparse batches: leaders per key, one-member closures planned.  Dense
    // ones (n * 16 >= cap, most bitmap words hold claims): one wave per
    // bitmap word replays the arcs of up to 64 slots that start in it.
    const bool plan = n * 16 < t_.cap;
    // Only sparse batches (settled per key) defer: a batch that covers a
    // large share of the table (n * 16 >= cap) nearly always queues big
    // closures, and deferring would only add a round trip.
    const bool defer = defer_big && (dedup || n * 16 < t_.cap);
    const bool arcs = !plan && arcs_on_ && ops_.ins_arcs && t_.cap >= 256;
    // A deferred batch led by k_ins_leader hands its count over from the
    // leader's last workgroup.  Its claims then either go to the next
    // bitmaps as the replay makes them (sparse batches without
    // deduplication, where copying the bitmaps costs less than a settle
    // pass and a kernel boundary: at most 24 MiB or 200 B per key of copy
    // traffic), or are settled by k_ins_bits while the host returns (the
    // settle reads no batch input).
    const bool early = defer && !arcs;
    const uint32_t seq = early ? next_ctr_seq() : 0;
    const bool next_bits = (plan || arcs) && !dedup && !fixed && next_bits_on_ &&
                           words * 40 <= std::max<uint64_t>(24ull << 20, n * 200) && ensure_next_bits();
    if (next_bits) {
      t.nfree = nfree_;
      t.nempty = nempty_;
      t.nq = nq_;
    }
    {
      Prof p(this, kKidInsDry);
      if ((st = check_launch(ops_.ins_dry(t, d_keys, n, Q_, nullptr, sc.dslot, ctr_, blk, stream_), "k_ins_dry")) !=
          FM_OK)
        return st;
    }
    if (arcs) {
      Prof p(this, kKidInsArcs);
      if ((st = check_launch(ops_.ins_arcs(t, d_keys, d_vals, Q_, sc.big, ctr_, stream_), "k_ins_arcs")) != FM_OK)
        return st;
    } else {
      Prof p(this, kKidInsLeader);
      if ((st = check_launch(ops_.ins_leader(t, d_keys, d_vals, n, Q_, sc.dslot, sc.big, ctr_, tune_.insert_coop_arc,
                                             leader_max_, plan, early ? h_ctr_ : nullptr, seq, stream_),
                             "k_ins_leader")) != FM_OK)
        return st;
    }
    unsigned long long* const Qb = Q_;  // this batch's claim bitmap
    if (next_bits) {  // later ops see the next bitmaps (k_ins_big below still reads the batch's)
      std::swap(t_.freebits, nfree_);
      std::swap(t_.emptybits, nempty_);
      std::swap(Q_, nq_);
    }
    // A counted dense batch with its claims in the next bitmaps: k_ins_big
    // (launched even with no closure queued) hands the counters over.
    const bool big_hands = next_bits && defer_big && !defer;
    const uint32_t big_seq = big_hands ? next_ctr_seq() : 0;
    if (defer) {
      pend_ = PendingBig{t, d_keys, d_vals, n, dedup, sc.dslot, sc.big, Qb, next_bits};
    } else {
      Prof p(this, kKidInsBig);
      if ((st = check_launch(ops_.ins_big(t, d_keys, d_vals, n, Qb, sc.big, ctr_, big_hands ? h_ctr_ : nullptr,
                                          big_seq, stream_),
                             "k_ins_big")) != FM_OK)
        return st;
    }
    if (!next_bits) return launch_bits(t, n, dedup, defer, sc.dslot, seq);
    q_pending_ = true;  // Q_ (the next claim bitmap) == freebits once the status shows no failed dry pass
    ctr_zero_ = false;  // an uncounted batch leaves its count; a counted one's hand-over may leave the queue
    ctr_in_host_ = early || big_hands;
    poll_seq_ = early ? seq : big_seq;
    return FM_OK;
  }
  {
    Prof p(this, kKidInsDry);
    if ((st = check_launch(ops_.ins_dry(t, d_keys, n, Q_, sc.home, nullptr, ctr_, blk, stream_), "k_ins_dry")) != FM_OK)
      return st;
  }
  {
    Prof p(this, kKidInsClosure);
    hipLaunchKernelGGL(dev::k_ins_closure, dim3(grid_for(n, blk)), dim3(blk), 0, stream_, t_.cap, Q_, sc.home, n,
                       sc.cid, sc.idx, ctr_);
    if ((st = check_launch(hipGetLastError(), "k_ins_closure")) != FM_OK) return st;
  }
  {
    Prof p(this, kKidInsSort);
    hipError_t e = sort_pairs(sc.cid, sc.cid_s, sc.idx, sc.idx_s, n, sort_end_bit_, sc.sort_tmp, nullptr, 0, stream_);
    if (e != hipSuccess) return fail(FM_ERR_KERNEL_LAUNCH, hip_msg("radix sort", e));
  }
  {
    Prof p(this, kKidInsReplay);
    if ((st = check_launch(ops_.ins_replay(t, d_keys, d_vals, n, sc.cid_s, sc.idx_s, ctr_, blk, stream_),
                           "k_ins_replay")) != FM_OK)
      return st;
  }
  if (!dedup) return FM_OK;
  hipLaunchKernelGGL(dev::k_dedup_reset, dim3(grid_for(n, blk)), dim3(blk), 0, stream_, t, n);
  return check_launch(hipGetLastError(), "k_dedup_reset");
}

fm_status GpuBackend::run_remove(const uint8_t* d_keys, uint64_t n, uint64_t* removed, uint8_t* outcome,
                                 bool async) {
  fm_status st;
  Table t = t_;
  t.outcome = outcome;
  // Keeping Q_ equal to freebits costs one atomic per removed key; above
  // kQKeepMax keys the next insert's copy of freebits is cheaper.
  t.freecopy = (q_synced_ && n <= kQKeepMax) ? Q_ : nullptr;
  if (!t.freecopy) q_synced_ = false;
  if ((st = ensure_scratch(n)) != FM_OK) return st;
  if ((st = zero_counters()) != FM_OK) return st;
  const uint32_t blk = tune_.insert_block;
  {
    Prof p(this, kKidRemProbe);
    if ((st = check_launch(ops_.rem_probe(t_, d_keys, n, ev_, tune_.remove_group, tune_.get_block, stream_),
                           "k_rem_probe")) != FM_OK)
      return st;
  }
  {
    Prof p(this, kKidRemCheck);
    hipLaunchKernelGGL(dev::k_rem_check, dim3(grid_for(n, blk)), dim3(blk), 0, stream_, t_, ev_, n, action_, dirty_,
                       cid_s_, ctr_);
    if ((st = check_launch(hipGetLastError(), "k_rem_check")) != FM_OK) return st;
  }
  // A counted remove reads the counters next: k_rem_apply hands them to
  // h_ctr_ (zeroing ctr_) with the status and the dirty-run count, and the
  // host polls for the hand-over; after a conflict it launches the replay
  // of the dirty runs and reads its count the same way.  An asynchronous
  // remove launches the replay unconditionally, gated on the device (it
  // exits at once unless k_rem_check flagged a conflict), so it reads
  // nothing back; its counts come from the outcome flags.
  const uint32_t seq = async ? 0 : next_ctr_seq();
  Counters* host = async ? nullptr : h_ctr_;
  {
    Prof p(this, kKidRemApply);
    hipLaunchKernelGGL(dev::k_rem_apply, dim3(grid_for(n, blk)), dim3(blk), 0, stream_, t, ev_, action_, n, dirty_,
                       idx_, ctr_, host, seq);
    if ((st = check_launch(hipGetLastError(), "k_rem_apply")) != FM_OK) return st;
  }
  if (async) {
    Prof p(this, kKidRemReplay);
    return check_launch(ops_.rem_replay(t, d_keys, n, dirty_, cid_s_, idx_, ctr_, true, 0, 0, nullptr, 0, stream_),
                        "k_rem_replay");
  }
  ctr_zero_ = ctr_in_host_ = true;
  poll_seq_ = seq;
  Counters c;
  if ((st = read_counters(&c)) != FM_OK) return st;
  uint64_t count = c.count;
  if (c.status & dev::kStatusConflict) {
    const uint32_t seq2 = next_ctr_seq();
    {
      Prof p(this, kKidRemReplay);
      if ((st = check_launch(ops_.rem_replay(t, d_keys, n, dirty_, cid_s_, idx_, ctr_, false, c.status, c.nbig, h_ctr_,
                                             seq2, stream_),
                             "k_rem_replay")) != FM_OK)
        return st;
    }
    ctr_zero_ = ctr_in_host_ = true;
    poll_seq_ = seq2;
    Counters c2;
    if ((st = read_counters(&c2)) != FM_OK) return st;
    count += c2.count;
  }
  *removed = count;
  return FM_OK;
}

// ---- per-kernel timing -----------------------------------------------------

const char* GpuBackend::kid_name(int k) {
  static const char* names[kNumKid] = {"k_get",        "ins_bitmap_copy", "k_ins_dry",    "k_ins_closure",
                                       "ins_radix_sort", "k_ins_replay",  "k_rem_probe",  "k_rem_check",
                                       "k_rem_apply",  "k_rem_runs",      "rem_radix_sort", "k_rem_replay",
                                       "copy_table",   "k_ins_leader",    "k_ins_big",    "k_ins_bits",
                                       "k_ins_dedup",  "k_ins_arcs"};
  return names[k];
}

hipEvent_t GpuBackend::take_event() {
  if (!ev_pool_.empty()) {
    hipEvent_t e = ev_pool_.back();
    ev_pool_.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void GpuBackend::prof_mark(int kid, bool begin) {
  if (begin) {
    prof_open_ = take_event();
    if (prof_open_) (void)hipEventRecord(prof_open_, stream_);
    return;
  }
  hipEvent_t b = take_event();
  if (!prof_open_ || !b) return;
  (void)hipEventRecord(b, stream_);
  pending_.push_back(PendingTiming{kid, prof_open_, b});
  prof_open_ = nullptr;
  if (pending_.size() > 4096) harvest();
}

void GpuBackend::harvest() {
  if (pending_.empty()) return;
  (void)hipStreamSynchronize(stream_);
  for (const PendingTiming& r : pending_) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      prof_ms_[r.kid] += ms;
      prof_n_[r.kid] += 1;
    }
    ev_pool_.push_back(r.a);
    ev_pool_.push_back(r.b);
  }
  pending_.clear();
}

fm_status GpuBackend::profile_only(const char* name) {
  if (!name[0]) {
    prof_only_ = -1;
    return FM_OK;
  }
  for (int k = 0; k < kNumKid; ++k)
    if (std::strcmp(kid_name(k), name) == 0) {
      prof_only_ = k;
      return FM_OK;
    }
  return fail(FM_ERR_INVALID_ARGUMENT, std::string("fm_profile_only: no kernel named ") + name);
}

fm_status GpuBackend::profile_enable(bool on) {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  harvest();
  prof_ = on;
  for (int k = 0; k < kNumKid; ++k) {
    prof_ms_[k] = 0.0;
    prof_n_[k] = 0;
  }
  return FM_OK;
}

fm_status GpuBackend::profile_read(fm_kernel_stat* out, uint32_t max, uint32_t* count) {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  harvest();
  uint32_t c = 0;
  for (int k = 0; k < kNumKid && c < max; ++k) {
    if (!prof_n_[k]) continue;
    std::memset(out[c].name, 0, sizeof out[c].name);
    std::strncpy(out[c].name, kid_name(k), sizeof out[c].name - 1);
    out[c].launches = prof_n_[k];
    out[c].total_ms = prof_ms_[k];
    ++c;
  }
  *count = c;
  return FM_OK;
}

// ---- host-facing API -------------------------------------------------------

fm_status GpuBackend::bulk_get(const void* keys, uint64_t n, void* out_values, uint8_t* out_found) {
  if (n == 0) return FM_OK;  // gpu.rs:112-114
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  if ((st = ensure_staging(n * t_.K, 0, n * t_.V, n)) != FM_OK) return st;
  const uint64_t kb = n * t_.K, vb = n * t_.V;
  const bool large = kb + vb + n >= kPinnedMin;
  const bool staged_in = large && !host_pinned(keys, kb) && ensure_host_staging(kb, vb + n);
  const bool staged_out =
      large && !(host_pinned(out_values, vb) && host_pinned(out_found, n)) && ensure_host_staging(kb, vb + n);
  hipError_t e;
  if (staged_in) {
    const HostXfer x{st_keys_, const_cast<uint8_t*>(static_cast<const uint8_t*>(keys)), kb};
    if ((st = stage_in(&x, 1)) != FM_OK) return st;
  } else if ((e = hipMemcpyAsync(st_keys_, keys, kb, hipMemcpyHostToDevice, stream_)) != hipSuccess) {
    return fail(FM_ERR_TRANSFER, hip_msg("H2D keys", e));
  }
  if ((st = run_get(st_keys_, n, st_out_, st_found_)) != FM_OK) return st;
  if (staged_out) {
    const HostXfer xs[2] = {{st_out_, static_cast<uint8_t*>(out_values), vb}, {st_found_, out_found, n}};
    return stage_out(xs, 2);
  }
  if ((e = hipMemcpyAsync(out_values, st_out_, vb, hipMemcpyDeviceToHost, stream_)) != hipSuccess ||
      (e = hipMemcpyAsync(out_found, st_found_, n, hipMemcpyDeviceToHost, stream_)) != hipSuccess ||
      (e = hipStreamSynchronize(stream_)) != hipSuccess)
    return fail(FM_ERR_TRANSFER, hip_msg("D2H results", e));
  return FM_OK;
}

fm_status GpuBackend::bulk_insert(const void* keys, const void* values, uint64_t n, uint64_t* out_new) {
  if (n == 0) {  // gpu.rs:194-196
    *out_new = 0;
    return FM_OK;
  }
  if (len_ + n > t_.cap) {  // gpu.rs:198-206, MAX_LOAD_FACTOR = 1.0
    set_table_full(len_, t_.cap);
    return FM_ERR_TABLE_FULL;
  }
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  if ((st = ensure_staging(n * t_.K, n * t_.V, 0, 0)) != FM_OK) return st;
  const uint64_t kb = n * t_.K, vb = n * t_.V;
  hipError_t e;
  if (kb + vb >= kPinnedMin && !(host_pinned(keys, kb) && host_pinned(values, vb)) && ensure_host_staging(kb + vb, 0)) {
    const HostXfer xs[2] = {{st_keys_, const_cast<uint8_t*>(static_cast<const uint8_t*>(keys)), kb},
                            {st_vals_, const_cast<uint8_t*>(static_cast<const uint8_t*>(values)), vb}};
    if ((st = stage_in(xs, 2)) != FM_OK) return st;
  } else if ((e = hipMemcpyAsync(st_keys_, keys, kb, hipMemcpyHostToDevice, stream_)) != hipSuccess ||
             (e = hipMemcpyAsync(st_vals_, values, vb, hipMemcpyHostToDevice, stream_)) != hipSuccess) {
    return fail(FM_ERR_TRANSFER, hip_msg("H2D pairs", e));
  }
  if ((st = run_insert(st_keys_, st_vals_, n, nullptr, true)) != FM_OK) return st;
  Counters c;
  if ((st = read_insert_counters(&c)) != FM_OK) return st;
  if ((st = check_insert_status(c)) != FM_OK) return st;
  q_synced_ = q_pending_;  // a clean led insert left Q_ equal to freebits
  len_ += c.count;
  *out_new = c.count;
  return FM_OK;
}

fm_status GpuBackend::bulk_remove(const void* keys, uint64_t n, uint64_t* out_removed) {
  if (n == 0) {  // gpu.rs:276-278
    *out_removed = 0;
    return FM_OK;
  }
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  if ((st = ensure_staging(n * t_.K, 0, 0, 0)) != FM_OK) return st;
  const uint64_t kb = n * t_.K;
  hipError_t e;
  if (kb >= kPinnedMin && !host_pinned(keys, kb) && ensure_host_staging(kb, 0)) {
    const HostXfer x{st_keys_, const_cast<uint8_t*>(static_cast<const uint8_t*>(keys)), kb};
    if ((st = stage_in(&x, 1)) != FM_OK) return st;
  } else if ((e = hipMemcpyAsync(st_keys_, keys, kb, hipMemcpyHostToDevice, stream_)) != hipSuccess) {
    return fail(FM_ERR_TRANSFER, hip_msg("H2D keys", e));
  }
  uint64_t removed = 0;
  if ((st = run_remove(st_keys_, n, &removed)) != FM_OK) return st;
  len_ -= std::min(len_, removed);
  *out_removed = removed;
  return FM_OK;
}

fm_status GpuBackend::clear() {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  q_synced_ = false;
  hipLaunchKernelGGL(dev::k_clear, dim3(grid_for(t_.cap, 256)), dim3(256), 0, stream_, t_, Q_);
  if ((st = check_launch(hipGetLastError(), "k_clear")) != FM_OK) return st;
  q_synced_ = true;  // Q_ filled like freebits
  hipError_t e = hipStreamSynchronize(stream_);
  if (e != hipSuccess) return fail(FM_ERR_KERNEL_LAUNCH, hip_msg("clear", e));
  len_ = 0;
  return FM_OK;
}

fm_status GpuBackend::export_slots(uint32_t* flags, void* keys, void* values) {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  hipError_t e;
  if ((e = hipMemcpyAsync(flags, t_.flags, t_.cap * 4, hipMemcpyDeviceToHost, stream_)) != hipSuccess ||
      (keys && (e = hipMemcpyAsync(keys, t_.keys, t_.cap * t_.K, hipMemcpyDeviceToHost, stream_)) != hipSuccess) ||
      (values && (e = hipMemcpyAsync(values, t_.vals, t_.cap * t_.V, hipMemcpyDeviceToHost, stream_)) != hipSuccess) ||
      (e = hipStreamSynchronize(stream_)) != hipSuccess)
    return fail(FM_ERR_TRANSFER, hip_msg("export", e));
  // Stale bytes of non-OCCUPIED slots are unobservable; zero them so that
  // exports compare bytewise.
  for (uint64_t i = 0; i < t_.cap; ++i) {
    if ((flags[i] & 0xFu) == dev::kOccupied) continue;
    if (keys) std::memset(static_cast<uint8_t*>(keys) + i * t_.K, 0, t_.K);
    if (values) std::memset(static_cast<uint8_t*>(values) + i * t_.V, 0, t_.V);
  }
  return FM_OK;
}

fm_status GpuBackend::import_slots(const uint32_t* flags, const void* keys, const void* values, uint64_t len) {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  hipError_t e;
  if ((e = hipMemcpyAsync(t_.flags, flags, t_.cap * 4, hipMemcpyHostToDevice, stream_)) != hipSuccess ||
      (e = hipMemcpyAsync(t_.keys, keys, t_.cap * t_.K, hipMemcpyHostToDevice, stream_)) != hipSuccess ||
      (e = hipMemcpyAsync(t_.vals, values, t_.cap * t_.V, hipMemcpyHostToDevice, stream_)) != hipSuccess)
    return fail(FM_ERR_TRANSFER, hip_msg("import", e));
  q_synced_ = false;
  hipLaunchKernelGGL(dev::k_rebuild_bits, dim3(grid_for(t_.cap, 256)), dim3(256), 0, stream_, t_);
  if ((st = check_launch(hipGetLastError(), "k_rebuild_bits")) != FM_OK) return st;
  if ((e = hipStreamSynchronize(stream_)) != hipSuccess) return fail(FM_ERR_TRANSFER, hip_msg("import", e));
  len_ = len;
  return FM_OK;
}

// ---- device-resident API ---------------------------------------------------

fm_status GpuBackend::alloc_device(uint64_t bytes, void** out) {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<uint64_t>(bytes, 1));
  if (e != hipSuccess) return fail(FM_ERR_GPU_ALLOC, hip_msg("hipMalloc", e));
  if ((e = hipMemsetAsync(p, 0, std::max<uint64_t>(bytes, 1), stream_)) != hipSuccess ||
      (e = hipStreamSynchronize(stream_)) != hipSuccess) {
    (void)hipFree(p);
    return fail(FM_ERR_GPU_ALLOC, hip_msg("zero fill", e));
  }
  *out = p;
  return FM_OK;
}

fm_status GpuBackend::free_device(void* p) {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  (void)hipStreamSynchronize(stream_);
  hipError_t e = hipFree(p);
  return e == hipSuccess ? FM_OK : fail(FM_ERR_GPU_ALLOC, hip_msg("hipFree", e));
}

fm_status GpuBackend::upload(const void* host, uint64_t bytes, void** out) {
  fm_status st;
  if ((st = alloc_device(bytes, out)) != FM_OK) return st;
  hipError_t e = hipMemcpyAsync(*out, host, bytes, hipMemcpyHostToDevice, stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  if (e != hipSuccess) return fail(FM_ERR_TRANSFER, hip_msg("upload", e));
  return FM_OK;
}

fm_status GpuBackend::download(const void* dptr, uint64_t bytes, void* host) {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  hipError_t e = hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  return e == hipSuccess ? FM_OK : fail(FM_ERR_TRANSFER, hip_msg("download", e));
}

fm_status GpuBackend::bulk_get_device(const void* d_keys, uint64_t n, void* d_out, uint8_t* d_found) {
  if (n == 0) return FM_OK;
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  if ((st = check_device_ptr(d_keys, vec_align(t_.K), "keys")) != FM_OK ||
      (st = check_device_ptr(d_out, vec_align(t_.V), "out_values")) != FM_OK)
    return st;
  return run_get(static_cast<const uint8_t*>(d_keys), n, static_cast<uint8_t*>(d_out), d_found);
}

fm_status GpuBackend::bulk_insert_device(const void* d_keys, const void* d_vals, uint64_t n, uint64_t* out_new,
                                         bool counted) {
  if (n == 0) {
    if (out_new) *out_new = 0;
    return FM_OK;
  }
  if (counted && len_ + n > t_.cap) {  // gpu.rs:551-559
    set_table_full(len_, t_.cap);
    return FM_ERR_TABLE_FULL;
  }
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  if ((st = check_device_ptr(d_keys, vec_align(t_.K), "keys")) != FM_OK ||
      (st = check_device_ptr(d_vals, vec_align(t_.V), "values")) != FM_OK)
    return st;
  if ((st = run_insert(static_cast<const uint8_t*>(d_keys), static_cast<const uint8_t*>(d_vals), n, nullptr,
                       counted)) != FM_OK)
    return st;
  if (!counted) {  // gpu.rs:670: assumes every key new
    len_ += n;
    return FM_OK;
  }
  Counters c;
  if ((st = read_insert_counters(&c)) != FM_OK) return st;
  if ((st = check_insert_status(c)) != FM_OK) return st;
  q_synced_ = q_pending_;  // a clean led insert left Q_ equal to freebits
  len_ += c.count;
  if (out_new) *out_new = c.count;
  return FM_OK;
}

// fm_bulk_insert_device with every per-call buffer in this GPU's `scratch`
// array: the same kernels and results, no allocation, no deduplication and
// no timing events (neither changes a result).
fm_status GpuBackend::bulk_insert_device_noalloc(const void* d_keys, const void* d_vals, uint64_t n,
                                                 uint64_t* out_new) {
  if (n == 0) {
    if (out_new) *out_new = 0;
    return FM_OK;
  }
  if (n > kNoAllocMaxKeys)
    return fail(FM_ERR_INVALID_ARGUMENT, "fm_bulk_insert_device_noalloc: " + std::to_string(n) +
                                             " keys; at most 1048576 per call without allocation");
  if (len_ + n > t_.cap) {  // gpu.rs:551-559
    set_table_full(len_, t_.cap);
    return FM_ERR_TABLE_FULL;
  }
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  uint8_t* base = scratch_ptr(dev_);
  if (scratch_err_ != hipSuccess || !base)
    return fail(FM_ERR_DEVICE_INIT, hip_msg("hipGetSymbolAddress (insert scratch)", scratch_err_));
  if ((st = check_device_ptr(d_keys, vec_align(t_.K), "keys")) != FM_OK ||
      (st = check_device_ptr(d_vals, vec_align(t_.V), "values")) != FM_OK)
    return st;
  auto arr = [&](uint64_t i) { return reinterpret_cast<uint32_t*>(base + i * kNoAllocArray); };
  const bool led = len_ + n < t_.cap && !tune_sort_only_;
  const InsertScratch sc = led ? InsertScratch{nullptr, arr(0), arr(1), nullptr, nullptr, nullptr, nullptr, nullptr}
                               : InsertScratch{arr(0), nullptr, nullptr, arr(1), arr(2), arr(3), arr(4), arr(5)};
  noalloc_ = true;
  st = run_insert(static_cast<const uint8_t*>(d_keys), static_cast<const uint8_t*>(d_vals), n, nullptr, true, &sc);
  Counters c{};
  if (st == FM_OK) st = read_insert_counters(&c);
  noalloc_ = false;
  if (st != FM_OK) return st;
  if ((st = check_insert_status(c)) != FM_OK) return st;
  q_synced_ = q_pending_;  // a clean led insert left Q_ equal to freebits
  len_ += c.count;
  if (out_new) *out_new = c.count;
  return FM_OK;
}

fm_status GpuBackend::bulk_remove_device(const void* d_keys, uint64_t n, uint64_t* out_removed) {
  if (n == 0) {
    if (out_removed) *out_removed = 0;
    return FM_OK;
  }
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  if ((st = check_device_ptr(d_keys, vec_align(t_.K), "keys")) != FM_OK) return st;
  uint64_t removed = 0;
  if ((st = run_remove(static_cast<const uint8_t*>(d_keys), n, &removed)) != FM_OK) return st;
  len_ -= std::min(len_, removed);
  if (out_removed) *out_removed = removed;
  return FM_OK;
}

fm_status GpuBackend::recount(uint64_t* out) {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  if ((st = zero_counters()) != FM_OK) return st;
  hipLaunchKernelGGL(dev::k_count, dim3(grid_for(t_.cap, 256)), dim3(256), 0, stream_, t_, ctr_);
  if ((st = check_launch(hipGetLastError(), "k_count")) != FM_OK) return st;
  Counters c;
  if ((st = read_counters(&c)) != FM_OK) return st;
  len_ = c.count;
  *out = c.count;
  return FM_OK;
}

fm_status GpuBackend::synchronize() {
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  hipError_t e = hipStreamSynchronize(stream_);
  return e == hipSuccess ? FM_OK : fail(FM_ERR_KERNEL_LAUNCH, hip_msg("synchronize", e));
}

fm_status GpuBackend::copy_from(Backend& other, uint32_t parts) {
  GpuBackend* src = dynamic_cast<GpuBackend*>(&other);
  if (!src || src->dev_ != dev_ || src->t_.cap != t_.cap || src->t_.K != t_.K || src->t_.V != t_.V ||
      src->t_.mode != t_.mode)
    return fail(FM_ERR_INVALID_ARGUMENT, "fm_copy_table: maps differ in device, capacity, sizes or hash");
  if (src == this) return FM_OK;
  fm_status st;
  if ((st = use_device()) != FM_OK) return st;
  hipError_t e = hipStreamSynchronize(src->stream_);
  const uint64_t words = (t_.cap + 63) / 64;
  Prof p(this, kKidCopyTable);
  if (e == hipSuccess) e = hipMemcpyAsync(t_.flags, src->t_.flags, t_.cap * 4, hipMemcpyDeviceToDevice, stream_);
  if (e == hipSuccess) e = hipMemcpyAsync(t_.keys, src->t_.keys, t_.cap * t_.K, hipMemcpyDeviceToDevice, stream_);
  if (e == hipSuccess && parts != FM_COPY_STRUCTURE)
    e = hipMemcpyAsync(t_.vals, src->t_.vals, t_.cap * t_.V, hipMemcpyDeviceToDevice, stream_);
  if (e == hipSuccess) e = hipMemcpyAsync(t_.freebits, src->t_.freebits, words * 8, hipMemcpyDeviceToDevice, stream_);
  if (e == hipSuccess) e = hipMemcpyAsync(t_.emptybits, src->t_.emptybits, words * 8, hipMemcpyDeviceToDevice, stream_);
  if (e == hipSuccess && Q_) e = hipMemcpyAsync(Q_, src->t_.freebits, words * 8, hipMemcpyDeviceToDevice, stream_);
  q_synced_ = e == hipSuccess && Q_ != nullptr;
  if (e != hipSuccess) return fail(FM_ERR_TRANSFER, hip_msg("fm_copy_table", e));
  len_ = src->len_;
  return FM_OK;
}

}  // namespace

std::unique_ptr<Backend> make_gpu_backend(const fm_builder& b, fm_status* st, std::string* err) {
  std::unique_ptr<GpuBackend> g(new GpuBackend());
  *st = g->init(b, err);
  if (*st != FM_OK) return nullptr;
  return g;
}

}  // namespace fm
