#!/usr/bin/env python3
"""FlashMap bulk-op benchmark on one MI355X.

Workload (the account-state configuration): FlashMap<[u8;32],[u8;128]>
with_capacity(10,000,000) (2^24 slots), preloaded with 8,000,000 random
pairs.  One step applies, in order, 20 bulk_insert batches of 100,000 pairs
(50% overwrites of existing keys, 50% new keys, no duplicates inside a batch),
one bulk_get of 1,000,000 queries at a 90% hit rate and one bulk_remove of
200,000 existing keys: 3,200,000 key-ops.  Every step starts from the
preloaded state: two working tables alternate and are restored from a device
snapshot between steps, outside the timed brackets.  Inputs are resident in
HBM before the timed region; results stay on the device.

The first step is checked bit for bit against the sequential CPU oracle
(counts, every get value and found flag, final length).  The CPU baseline
is the multi-threaded restatement of the reference's Rayon backend, timed
on this host on one step of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import hashlib
import json
import gc
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "flash-map_amd")]

import numpy as np  # noqa: E402

import flashmap_amd as fm  # noqa: E402
from flashmap_amd.workloads import config2  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
K_BYTES, V_BYTES, F_BYTES = 32, 128, 4
HIT = 0.9
# Algorithmic bytes per key (get: query + flag + hit*(key+value) + value out
# + found byte; insert: pair in + slot written + flag read/write; remove:
# key in + key compare + flag read/write).
B_GET = K_BYTES + F_BYTES + HIT * (K_BYTES + V_BYTES) + V_BYTES + 1
B_INS = 2 * (K_BYTES + V_BYTES) + 2 * F_BYTES
B_REM = 2 * K_BYTES + 2 * F_BYTES


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cpu():
    """The host the CPU baselines run on: lscpu's model, sockets, cores per
    socket and threads per core, and the logical CPUs this process may use
    (what nproc prints)."""
    info = {"model": host_cpu_model(), "nproc": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
            elif k == "CPU(s)":
                info["logical_cpus"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def host_text():
    h = host_cpu()
    parts = [f"host '{h['model']}'", f"nproc {h['nproc']}"]
    if "sockets" in h and "cores_per_socket" in h:
        parts.append(f"{h['sockets']} socket(s) x {h['cores_per_socket']} cores x "
                     f"{h.get('threads_per_core', 1)} thread(s)")
    return ", ".join(parts)


def cpu_threads(args):
    """Rayon's global pool: one worker per logical CPU of the host."""
    return args.cpu_threads or len(os.sched_getaffinity(0))


class Step:
    """Device-resident inputs of one step and the op sequence."""

    def __init__(self, owner, cfg):
        self.batches = [(owner.upload(k), owner.upload(v), len(k)) for k, v in cfg.batches]
        self.get_keys = owner.upload(cfg.get_keys)
        self.n_get = len(cfg.get_keys)
        self.rem_keys = owner.upload(cfg.remove_keys)
        self.n_rem = len(cfg.remove_keys)
        self.out_vals = owner.alloc_device(self.n_get * V_BYTES)
        self.out_found = owner.alloc_device(self.n_get)
        self.keys_per_step = sum(b[2] for b in self.batches) + self.n_get + self.n_rem

    def run(self, m):
        new = [m.bulk_insert_device(dk, dv, n) for dk, dv, n in self.batches]
        m.bulk_get_device(self.get_keys, self.n_get, self.out_vals, self.out_found)
        removed = m.bulk_remove_device(self.rem_keys, self.n_rem)
        return new, removed


def cpu_baseline(cfg, threads, target_s=10.0, blocks=5):
    """Rayon-port timing: one untimed warm-up block, then `blocks` timed blocks
    of the step repeated from the preloaded state (restored untimed before
    each repetition), each block about target_s / blocks seconds long.
    Returns the median block rate (Mkeys/s), the timed seconds, the
    repetitions per block and the rates of all blocks."""
    from oracle.binding import RayonPortMap

    snap = RayonPortMap(cfg.capacity, K_BYTES, V_BYTES, 0, threads=threads)
    snap.bulk_insert(cfg.preload_keys, cfg.preload_vals)
    p = RayonPortMap(cfg.capacity, K_BYTES, V_BYTES, 0, threads=threads)
    keys = sum(len(k) for k, _ in cfg.batches) + len(cfg.get_keys) + len(cfg.remove_keys)

    def one():
        p.copy_from(snap, structure_only=True)
        t0 = time.perf_counter()
        for k, v in cfg.batches:
            p.bulk_insert(k, v)
        p.bulk_get(cfg.get_keys)
        p.bulk_remove(cfg.remove_keys)
        return time.perf_counter() - t0

    first = one()  # the warm-up block's length sets the block size
    per_block = max(1, int(round(target_s / blocks / max(first, 1e-3))))
    for _ in range(per_block - 1):
        one()
    rates, timed = [], 0.0
    for _ in range(blocks):
        t = sum(one() for _ in range(per_block))
        timed += t
        rates.append(keys * per_block / t / 1e6)
    p.close()
    snap.close()
    return float(np.median(rates)), timed, per_block, rates


def oracle_step(cfg):
    """Sequential oracle: expected results of one step (checker only)."""
    from oracle.binding import OracleMap

    o = OracleMap(cfg.capacity, K_BYTES, V_BYTES, 0)
    o.bulk_insert(cfg.preload_keys, cfg.preload_vals)
    t0 = time.perf_counter()
    new = [o.bulk_insert(k, v) for k, v in cfg.batches]
    vals, found = o.bulk_get(cfg.get_keys)
    removed = o.bulk_remove(cfg.remove_keys)
    dt = time.perf_counter() - t0
    res = dict(new=new, get_sha=hashlib.sha256(vals.tobytes() + found.tobytes()).hexdigest(),
               hits=int(found.sum()), removed=removed, len=len(o), seconds=dt)
    o.close()
    return res


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every logical CPU of the host (nproc)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="timed length of the CPU-baseline sample")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the workload (testing only)")
    ap.add_argument("--config", choices=["all", "c1", "c2", "c3", "c4", "c5", "host"], default="all",
                    help="all (default): the account-state metric line (c2) with sections for every other "
                         "configuration and the host-buffer calls; c1: plain u64 map, per-op roofline; "
                         "c2: account-state mix alone; c3: join-probe index (40M load, 100M probes); "
                         "c4: 512-table fused vs plain; c5: serving loop at several offered loads; "
                         "host: host-buffer calls on 1M [u8;32]->[u8;128] pairs")
    ap.add_argument("--tables", type=int, default=512, help="c4: number of tables")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--get-shape", default="",
                    help="G,B: get launch shape (group width, block size) instead of the autotuner's "
                         "(c1, c2, c3; counter runs use it to profile the shape the timed runs choose)")
    ap.add_argument("--requests", type=int, default=1000, help="c5: requests per offered load")
    ap.add_argument("--rates", default="2000,10000,50000", help="c5: offered loads, requests/s")
    ap.add_argument("--detail", default="", help="all: also write every configuration's full line to this file")
    return ap.parse_args(argv)


def main():
    args = parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the table would shard by "
              "slot range; see the sharding design). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    single = {"c1": run_c1, "c2": run_c2, "c3": run_c3, "c4": run_c4, "c5": run_c5, "host": run_host_api}
    if args.config in single:
        line, ok = single[args.config](args)
        print(json.dumps(line), flush=True)
        sys.exit(0 if ok else 1)
    run_all(args)


def run_c2(args):
    s = args.scale
    log("generating workload")
    cfg = config2(n_preload=int(8_000_000 * s), batch=int(100_000 * s), n_get=int(1_000_000 * s),
                  n_remove=int(200_000 * s), capacity=int(10_000_000 * s))

    def build():
        m = fm.FlashMap.builder(cfg.capacity, K_BYTES, V_BYTES).require_gpu().build()
        assert m.backend() == fm.Backend.GPU
        return m

    snap, A, B = build(), build(), build()
    log(f"preloading {len(cfg.preload_keys)} pairs into {snap.capacity()} slots")
    dk, dv = snap.upload(cfg.preload_keys), snap.upload(cfg.preload_vals)
    t0 = time.perf_counter()
    assert snap.bulk_insert_device(dk, dv, len(cfg.preload_keys)) == len(cfg.preload_keys)
    preload_s = time.perf_counter() - t0
    dk.free()
    dv.free()
    step = Step(A, cfg)

    check = None
    seq_mkeys_s = None
    if not args.no_check:
        log("checking one step against the oracle")
        A.copy_from(snap)
        new, removed = step.run(A)
        vals = A.download(step.out_vals)
        found = A.download(step.out_found)
        exp = oracle_step(cfg)
        got_sha = hashlib.sha256(vals.tobytes() + found.tobytes()).hexdigest()
        ok = new == exp["new"] and removed == exp["removed"] and got_sha == exp["get_sha"] and len(A) == exp["len"]
        seq_mkeys_s = (sum(len(k) for k, _ in cfg.batches) + len(cfg.get_keys) + len(cfg.remove_keys)) / \
            exp["seconds"] / 1e6
        check = {"bit_exact": bool(ok), "hits": exp["hits"], "new_per_batch": exp["new"][:3],
                 "removed": removed}
        if not ok:
            log(f"PARITY FAILURE: gpu new={new[:3]} removed={removed} sha={got_sha[:12]} len={len(A)}; "
                f"oracle new={exp['new'][:3]} removed={exp['removed']} sha={exp['get_sha'][:12]} len={exp['len']}")

    # Between timed steps the working tables get back the preloaded structure
    # (flags, keys, bitmaps).  Value bytes are not reset: no kernel branches
    # on or indexes by a value, so each step does the same work; the checked
    # step above ran on a fully restored table.
    maps = [A, B]
    get_tune = pinned_get_shape(args)
    if get_tune is None and not args.no_autotune:
        # get launch shape (group width x block) for this table and query mix
        get_tune = snap.autotune_get(step.get_keys, step.n_get, reps=5)
        log(f"get autotune: {get_tune}")
    if get_tune is not None:
        for m in maps:
            m.set_tuning(get_group=get_tune["get_group"], get_block=get_tune["get_block"])
    A.copy_from(snap, structure_only=True)
    B.copy_from(snap, structure_only=True)
    for i in range(args.warmup):
        m = maps[i % 2]
        step.run(m)
        m.copy_from(snap, structure_only=True)

    def profiled_steps(count, only, first):
        """Runs `count` steps with event timing (`only`: that kernel's
        launches alone); returns the per-kernel stats and the elapsed time."""
        for m in maps:
            m.synchronize()
            m.profile_only(only)
            m.profile(True)
        el = 0.0
        for i in range(count):
            m = maps[(first + i) % 2]
            m.synchronize()
            t0 = time.perf_counter()
            step.run(m)
            m.synchronize()
            el += time.perf_counter() - t0
            m.copy_from(snap, structure_only=True)
        for m in maps:
            m.synchronize()
        st = {}
        for m in maps:
            for name, (n, ms) in m.profile_read().items():
                a = st.setdefault(name, [0, 0.0])
                a[0] += n
                a[1] += ms
            m.profile(False)
            m.profile_only("")
        return st, el

    # Per-kernel breakdown (every launch bracketed by events) from a few
    # untimed steps; it names the dominant kernel: the largest summed device
    # time among the kernels that move each op's algorithmic bytes (get: the
    # probe; insert: the closure replay, which reads the pairs and writes the
    # slots; remove: the probe).
    kstats, _ = profiled_steps(3, "", args.warmup)
    batch_n = len(cfg.batches[0][0])
    per_launch = {"k_get": B_GET * step.n_get, "k_ins_leader": B_INS * batch_n, "k_ins_replay": B_INS * batch_n,
                  "k_rem_probe": B_REM * step.n_rem}
    ops_only = {k: v for k, v in kstats.items() if k in per_launch}
    dom = max(ops_only, key=lambda k: ops_only[k][1])
    # Each timed step is bracketed by a device synchronisation on both sides;
    # the restore of the working table between steps (harness state reset, not
    # part of the workload) runs outside the timed brackets.  Events bracket
    # the dominant kernel's launches only (its roofline figure), so the other
    # launches carry no event records.
    log(f"timing {args.steps} steps (events around {dom})")
    tstats, elapsed = profiled_steps(args.steps, dom, args.warmup + 3)

    keys = step.keys_per_step * args.steps
    value = keys / elapsed / 1e6
    ms_per_step = elapsed * 1e3 / args.steps

    dn, dms = tstats[dom]
    avg_s = dms / dn / 1e3
    achieved = per_launch[dom] / avg_s / 1e9
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", f"{dom}_account_state_10M_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    def op_rate(names, keys_per_launch, launches_key):
        n = kstats.get(launches_key, [0, 0])[0]
        ms = sum(kstats.get(k, [0, 0.0])[1] for k in names)
        return (keys_per_launch * n / (ms / 1e3) / 1e6) if ms else None

    ins_names = ["ins_bitmap_copy", "k_ins_dry", "k_ins_closure", "ins_radix_sort", "k_ins_replay", "k_ins_leader",
                 "k_ins_arcs", "k_ins_big", "k_ins_bits"]
    rem_names = ["k_rem_probe", "k_rem_check", "k_rem_apply", "k_rem_replay"]
    get_rate = op_rate(["k_get"], step.n_get, "k_get")
    ins_rate = op_rate(ins_names, batch_n, "k_ins_dry")
    rem_rate = op_rate(rem_names, step.n_rem, "k_rem_probe")

    cpu = None
    if not args.no_cpu_baseline:
        threads = cpu_threads(args)
        log(f"CPU baseline (Rayon restatement, {threads} threads)")
        rate, dt, reps, rates = cpu_baseline(cfg, threads, target_s=args.cpu_seconds)
        cpu = {"value": round(rate, 3), "unit": "Mkeys/s", "cores": threads, "kind": "port",
               "sample": f"median of 5 timed blocks (after 1 warm-up block) of {reps} steps of "
                         f"{step.keys_per_step} key-ops, each step from the 8M preloaded state (restored untimed); "
                         f"{dt:.2f} s timed; {host_text()}",
               "block_mkeys_s": [round(r, 2) for r in rates]}
        if seq_mkeys_s is not None:  # the checker's one step of the sequential oracle, 1 thread
            cpu["sequential_1_thread_mkeys_s"] = round(seq_mkeys_s, 3)

    out = {
        "metric": "bulk_get/insert Mkeys/s + % HBM peak; batched vs plain; serving p99 latency",
        "value": round(value, 3),
        "unit": "Mkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (SplitMix64 seed 2)",
        "config": {"workload": "account-state: FlashMap<[u8;32],[u8;128]> with_capacity(10M)=2^24 slots, "
                               "preload 8M; step = 20x bulk_insert 100k (50% overwrite) + bulk_get 1M "
                               "(90% hits) + bulk_remove 200k, each step from the preloaded state",
                   "keys_per_step": step.keys_per_step, "parallelism": "single GPU",
                   "timing": "each step between device synchronisations; the table restore between steps "
                             "is outside the timed brackets; in the timed steps HIP events bracket the "
                             "dominant kernel's launches only (kernels_ms and ops: 3 untimed steps with every "
                             "launch bracketed)"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "avg_launch_us": round(avg_s * 1e6, 3), "alg_bytes_per_launch": per_launch[dom]},
        "cpu_baseline": cpu,
        "ops": {"get_mkeys_s": get_rate and round(get_rate, 2), "insert_mkeys_s": ins_rate and round(ins_rate, 2),
                "remove_mkeys_s": rem_rate and round(rem_rate, 2),
                "get_hbm_frac": get_rate and round(get_rate * 1e6 * B_GET / 1e9 / HBM_PEAK_GBS, 5),
                "insert_hbm_frac": ins_rate and round(ins_rate * 1e6 * B_INS / 1e9 / HBM_PEAK_GBS, 5)},
        "kernels_ms": {k: {"launches": v[0], "avg_us": round(v[1] / v[0] * 1e3, 3)} for k, v in sorted(kstats.items())},
        "preload_mkeys_s": round(len(cfg.preload_keys) / preload_s / 1e6, 2),
        "get_autotune": get_tune,
        "check": check,
    }
    return out, check is None or check["bit_exact"]


# ---------------------------------------------------------------------------
# Shared by configs 1 and 3: a measured stream-copy peak, per-op device time
# from the map's kernel events, and the figures derived from them.
# ---------------------------------------------------------------------------
def stream_copy_peak(m, nbytes=1 << 30, reps=5):
    """Device-to-device copy of nbytes on the map's stream (the runtime's
    copy kernel), timed with HIP events: best and median GB/s of read+write."""
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    vp = ctypes.c_void_p
    hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    hip.hipEventDestroy.argtypes = [vp]
    hip.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
    src, dst = m.alloc_device(nbytes), m.alloc_device(nbytes)
    stream = vp(m.device_stream())
    e0, e1 = vp(), vp()
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    rates = []
    for i in range(reps + 1):
        assert hip.hipEventRecord(e0, stream) == 0
        assert hip.hipMemcpyAsync(vp(dst.ptr), vp(src.ptr), nbytes, 3, stream) == 0  # device to device
        assert hip.hipEventRecord(e1, stream) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        if i:  # the first copy warms up
            rates.append(2 * nbytes / (ms.value / 1e3) / 1e9)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    src.free()
    dst.free()
    return {"bytes": nbytes, "best_gbs": round(max(rates), 1), "median_gbs": round(float(np.median(rates)), 1),
            "note": "hipMemcpyAsync device-to-device on the map's stream, read+write bytes per HIP-event time"}


def timed_op(m, fn):
    """Runs fn() on map m with kernel events on; returns (result, device ms
    summed over the op's kernels, wall ms around the call, {kernel: ms})."""
    m.synchronize()
    m.profile(True)
    t0 = time.perf_counter()
    r = fn()
    m.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    ks = m.profile_read()
    m.profile(False)
    return r, sum(ms for _, ms in ks.values()), wall, {k: ms for k, (_, ms) in ks.items()}


def op_figures(n, b_key, line_b_key, dev_ms, wall_ms, peak_gbs, copy_gbs):
    """Median device/wall time of one op over the timed steps and the rates
    derived from them."""
    d, w = float(np.median(dev_ms)), float(np.median(wall_ms))
    ach = n * b_key / (d / 1e3) / 1e9
    return {"keys": n, "device_ms": round(d, 4), "wall_ms": round(w, 4),
            "mkeys_s": round(n / (d / 1e3) / 1e6, 2), "wall_mkeys_s": round(n / (w / 1e3) / 1e6, 2),
            "alg_bytes_per_key": round(b_key, 2), "achieved_gbs": round(ach, 1),
            "frac": round(ach / peak_gbs, 5), "frac_of_copy": round(ach / copy_gbs, 5) if copy_gbs else None,
            "line_bytes_per_key": round(line_b_key, 2),
            "line_frac": round(n * line_b_key / (d / 1e3) / 1e9 / peak_gbs, 5)}


def b_get(K, V, F, h):
    return K + F + h * (K + V) + V + 1


def b_ins(K, V, F):
    return 2 * (K + V) + 2 * F


def b_rem(K, F):
    return 2 * K + 2 * F


# Line-granular companions: 64 B per random field touch (get: flag, key, and
# the value on a hit; insert: flag, key, value; remove: flag, key) plus the
# streamed query/pair input and result output.
def line_get(K, V, h):
    return 64 * (2 + h) + K + V + 1


def line_ins(K, V):
    return 64 * 3 + K + V


def line_rem(K):
    return 64 * 2 + K


def pinned_get_shape(args):
    """The --get-shape G,B launch shape as an autotune-style record, or None."""
    if not args.get_shape:
        return None
    g, b = (int(x) for x in args.get_shape.split(","))
    return {"get_group": g, "get_block": b, "pinned": True}


def load_pmc(name):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


# ---------------------------------------------------------------------------
# Config 1: FlashMap<u64,u64> with_capacity(2^20).  One step from the empty
# table: bulk_insert 500k, bulk_get 500k (50% hits), bulk_remove 100k,
# bulk_get of the same 500k queries again.  Each op is timed on its own.
# ---------------------------------------------------------------------------
def run_c1(args):
    from flashmap_amd.workloads import config1

    c = config1()
    K = V = 8
    F = 4
    m = fm.FlashMap.builder(c.capacity, K, V).require_gpu().build()
    assert m.backend() == fm.Backend.GPU
    copy = stream_copy_peak(m)
    log(f"c1: stream copy {copy}")
    ik, iv = m.upload(c.insert_keys.view(np.uint8)), m.upload(c.insert_vals.view(np.uint8))
    q, rk = m.upload(c.queries.view(np.uint8)), m.upload(c.remove_keys.view(np.uint8))
    nq, ni, nr = len(c.queries), len(c.insert_keys), len(c.remove_keys)
    out, found = m.alloc_device(nq * V), m.alloc_device(nq)
    names = ["insert", "get", "remove", "get2"]
    fns = {"insert": lambda: m.bulk_insert_device(ik, iv, ni),
           "get": lambda: m.bulk_get_device(q, nq, out, found),
           "remove": lambda: m.bulk_remove_device(rk, nr),
           "get2": lambda: m.bulk_get_device(q, nq, out, found)}
    get_tune = pinned_get_shape(args)
    if get_tune is None and not args.no_autotune:
        # get launch shape for this table and query mix, tuned on the
        # populated table (gets change nothing); every step then clears it
        m.clear()
        m.bulk_insert_device(ik, iv, ni)
        get_tune = m.autotune_get(q, nq, reps=5)
        log(f"c1: get autotune {get_tune}")
    if get_tune is not None:
        m.set_tuning(get_group=get_tune["get_group"], get_block=get_tune["get_block"])
    dev = {k: [] for k in names}
    wall = {k: [] for k in names}
    kern = {k: {} for k in names}
    results = []
    for i in range(args.warmup + args.steps):
        m.clear()
        step = {}
        for k in names:
            r, d, w, ks = timed_op(m, fns[k])
            if k.startswith("get"):
                r = (hashlib.sha256(m.download(out).tobytes() + m.download(found).tobytes()).hexdigest(),
                     int(m.download(found).astype(bool).sum())) if i == 0 else None
            step[k] = r
            if i >= args.warmup:
                dev[k].append(d)
                wall[k].append(w)
                for kn, ms in ks.items():
                    kern[k].setdefault(kn, []).append(ms)
        results.append(step)
    first = results[0]
    hits = {"get": first["get"][1], "get2": first["get2"][1]}

    check = None
    if not args.no_check:
        from oracle.binding import OracleMap
        o = OracleMap(c.capacity, K, V, 0)
        exp = {"insert": o.bulk_insert(c.insert_keys.view(np.uint8), c.insert_vals.view(np.uint8))}
        for k in ("get", "remove", "get2"):
            if k == "remove":
                exp[k] = o.bulk_remove(c.remove_keys.view(np.uint8))
            else:
                v, f = o.bulk_get(c.queries.view(np.uint8))
                exp[k] = (hashlib.sha256(v.tobytes() + f.astype(np.uint8).tobytes()).hexdigest(), int(f.sum()))
        ok = all(first[k] == exp[k] for k in names) and len(m) == len(o)
        check = {"bit_exact": bool(ok), "inserted": first["insert"], "removed": first["remove"],
                 "hits": hits}
        o.close()
        if not ok:
            log(f"PARITY FAILURE c1: gpu {first} oracle {exp}")

    # Host-buffer calls (PCIe-inclusive, the reference's gpu.rs path): wall time.
    host = {k: [] for k in names}
    hk, hv, hq, hr = c.insert_keys, c.insert_vals, c.queries, c.remove_keys
    host_fns = {"insert": lambda: m.bulk_insert(hk, hv), "get": lambda: m.bulk_get(hq),
                "remove": lambda: m.bulk_remove(hr), "get2": lambda: m.bulk_get(hq)}
    for i in range(1 + min(args.steps, 5)):
        m.clear()
        for k in names:
            t0 = time.perf_counter()
            host_fns[k]()
            if i:
                host[k].append((time.perf_counter() - t0) * 1e3)

    n = {"insert": ni, "get": nq, "remove": nr, "get2": nq}
    h = {"get": hits["get"] / nq, "get2": hits["get2"] / nq}
    bkey = {"insert": b_ins(K, V, F), "get": b_get(K, V, F, h["get"]), "remove": b_rem(K, F),
            "get2": b_get(K, V, F, h["get2"])}
    lkey = {"insert": line_ins(K, V), "get": line_get(K, V, h["get"]), "remove": line_rem(K),
            "get2": line_get(K, V, h["get2"])}
    ops = {}
    pmc = load_pmc("plain_u64_1M_pmc.json") or {}
    for k in names:
        f = op_figures(n[k], bkey[k], lkey[k], dev[k], wall[k], HBM_PEAK_GBS, copy["best_gbs"])
        f["host_api_ms"] = round(float(np.median(host[k])), 4)
        f["kernels_us"] = {kn: round(float(np.median(v)) * 1e3, 2) for kn, v in sorted(kern[k].items())}
        tr = pmc.get(k)
        if tr:
            f["hbm_measured_gbs"] = round(tr["hbm_bytes"] / (f["device_ms"] / 1e3) / 1e9, 1)
            f["traffic_bytes"] = tr["hbm_bytes"]
            f["traffic_x_alg"] = round(tr["hbm_bytes"] / (n[k] * bkey[k]), 2)
        ops[k] = f

    cpu = None
    if not args.no_cpu_baseline:
        from oracle.binding import OracleMap, RayonPortMap
        threads = cpu_threads(args)
        cpu = {}
        for kind, mk in (("port", lambda: RayonPortMap(c.capacity, K, V, 0, threads=threads)),
                         ("sequential", lambda: OracleMap(c.capacity, K, V, 0))):
            times = {k: [] for k in names}
            for i in range(6):  # one warm-up, then the median of 5
                p = mk()
                cf = {"insert": lambda: p.bulk_insert(c.insert_keys.view(np.uint8), c.insert_vals.view(np.uint8)),
                      "get": lambda: p.bulk_get(c.queries.view(np.uint8)),
                      "remove": lambda: p.bulk_remove(c.remove_keys.view(np.uint8)),
                      "get2": lambda: p.bulk_get(c.queries.view(np.uint8))}
                for k in names:
                    t0 = time.perf_counter()
                    cf[k]()
                    if i:
                        times[k].append(time.perf_counter() - t0)
                p.close()
            cpu[kind] = {"cores": threads if kind == "port" else 1,
                         "mkeys_s": {k: round(n[k] / float(np.median(times[k])) / 1e6, 2) for k in names},
                         "sequence_mkeys_s": round(sum(n.values()) / sum(float(np.median(times[k])) for k in names)
                                                   / 1e6, 2)}
    total_dev = sum(ops[k]["device_ms"] for k in names)
    value = sum(n.values()) / (total_dev / 1e3) / 1e6
    dom = max(names, key=lambda k: ops[k]["device_ms"])
    out_line = {
        "metric": "bulk_get/insert Mkeys/s + % HBM peak (config 1 per-op)",
        "value": round(value, 3), "unit": "Mkeys/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total_dev, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic (SplitMix64 seed 1)",
        "config": {"workload": "plain: FlashMap<u64,u64> with_capacity(2^20); from empty: bulk_insert 500k, "
                               "bulk_get 500k (50% hits), bulk_remove 100k, bulk_get of the same queries",
                   "keys_per_step": sum(n.values()), "parallelism": "single GPU",
                   "timing": "per op: kernel events on the map's stream summed (device_ms) and wall clock "
                             "around the synchronised call (wall_ms); medians over the timed steps"},
        "roofline": {"bound": "hbm", "op": dom, "achieved": ops[dom]["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ops[dom]["frac"], "traffic": ops[dom].get("traffic_bytes")},
        "stream_copy": copy,
        "ops": ops,
        "get_autotune": get_tune,
        "cpu_baseline": cpu and {"value": cpu["port"]["sequence_mkeys_s"], "unit": "Mkeys/s",
                                 "cores": cpu["port"]["cores"], "kind": "port",
                                 "sample": f"the whole config-1 sequence, one warm-up then the median of 5 per op; "
                                           f"{host_text()}", "detail": cpu},
        "check": check,
    }
    return out_line, check is None or check["bit_exact"]


# ---------------------------------------------------------------------------
# Config 3: join-probe index, FlashMap<[u8;32],u64> with_capacity(50M)
# (2^26 slots).  One step: bulk-load 40M keys -> row id into the empty
# table, then bulk_get 100M probes (90% Zipf(1.0) over the loaded keys, 10%
# misses).
# ---------------------------------------------------------------------------
def run_c3(args):
    from flashmap_amd.workloads import config3

    s = args.scale
    log("c3: generating workload")
    t0 = time.perf_counter()
    c = config3(n_load=int(40_000_000 * s), n_probe=int(100_000_000 * s), capacity=int(50_000_000 * s))
    log(f"c3: generated in {time.perf_counter() - t0:.1f} s")
    K, V, F = 32, 8, 4
    nl, npb = len(c.keys), len(c.probes)
    m = fm.FlashMap.builder(c.capacity, K, V).require_gpu().build()
    assert m.backend() == fm.Backend.GPU
    copy = stream_copy_peak(m)
    log(f"c3: stream copy {copy}")
    dk, dv = m.upload(c.keys), m.upload(c.row_ids.view(np.uint8))
    dp = m.upload(c.probes)
    out, found = m.alloc_device(npb * V), m.alloc_device(npb)
    hit = float(c.probe_hit.mean())
    get_tune = None
    dev = {"load": [], "probe": []}
    wall = {"load": [], "probe": []}
    kern = {"load": {}, "probe": {}}
    check = None
    for i in range(args.warmup + args.steps):
        m.clear()
        r, d, w, ks = timed_op(m, lambda: m.bulk_insert_device(dk, dv, nl))
        assert r == nl, f"load inserted {r} of {nl}"
        if i == 0 and not args.no_autotune:
            get_tune = pinned_get_shape(args)
            if get_tune is None:
                get_tune = m.autotune_get(dp, min(npb, 10_000_000), reps=3)
                log(f"c3: get autotune {get_tune}")
            m.set_tuning(get_group=get_tune["get_group"], get_block=get_tune["get_block"])
        _, d2, w2, ks2 = timed_op(m, lambda: m.bulk_get_device(dp, npb, out, found))
        if i == 0 and not args.no_check:
            # Property check: every inserted probe returns the row id its key
            # was loaded with, every miss is not found, the map holds nl keys.
            f = m.download(found).astype(bool)
            v = m.download(out).view("<u8")
            exp_hit = c.probe_row >= 0
            ok = (len(m) == nl and np.array_equal(f, exp_hit) and
                  np.array_equal(v[exp_hit], c.probe_row[exp_hit].astype(np.uint64)) and not v[~exp_hit].any())
            check = {"rows_exact": bool(ok), "probes_checked": npb, "hits": int(f.sum())}
            if not ok:
                log("PARITY FAILURE c3")
        if i >= args.warmup:
            for k, dd, ww, kk in (("load", d, w, ks), ("probe", d2, w2, ks2)):
                dev[k].append(dd)
                wall[k].append(ww)
                for kn, ms in kk.items():
                    kern[k].setdefault(kn, []).append(ms)
    pmc = load_pmc("join_probe_50M_pmc.json") or {}
    ops = {}
    for k, n, bk, lk in (("load", nl, b_ins(K, V, F), line_ins(K, V)),
                         ("probe", npb, b_get(K, V, F, hit), line_get(K, V, hit))):
        f = op_figures(n, bk, lk, dev[k], wall[k], HBM_PEAK_GBS, copy["best_gbs"])
        f["kernels_us"] = {kn: round(float(np.median(v)) * 1e3, 2) for kn, v in sorted(kern[k].items())}
        tr = pmc.get(k)
        if tr:
            f["hbm_measured_gbs"] = round(tr["hbm_bytes"] / (f["device_ms"] / 1e3) / 1e9, 1)
            f["traffic_bytes"] = tr["hbm_bytes"]
        ops[k] = f

    cpu = None
    if not args.no_cpu_baseline:
        from oracle.binding import RayonPortMap
        threads = cpu_threads(args)
        n_cpu_probe = min(npb, int(20_000_000 * max(s, 0.01)))
        tls, tps = [], []
        for rep in range(4):  # one warm-up, then the median of 3, each into a fresh map
            p = RayonPortMap(c.capacity, K, V, 0, threads=threads)
            t0 = time.perf_counter()
            p.bulk_insert(c.keys, c.row_ids.view(np.uint8))
            tl = time.perf_counter() - t0
            t0 = time.perf_counter()
            p.bulk_get(c.probes[:n_cpu_probe])
            tp = time.perf_counter() - t0
            p.close()
            if rep:
                tls.append(tl)
                tps.append(tp)
        tl, tp = float(np.median(tls)), float(np.median(tps))
        cpu = {"value": round(n_cpu_probe / tp / 1e6, 3), "unit": "Mkeys/s (probe)", "cores": threads,
               "kind": "port", "load_mkeys_s": round(nl / tl / 1e6, 3),
               "sample": f"load of all {nl} keys ({tl:.2f} s) then the first {n_cpu_probe} probes ({tp:.2f} s), "
                         f"one warm-up then the median of 3, each into a fresh map; {host_text()}"}
    probe = ops["probe"]
    out_line = {
        "metric": "bulk_get Mkeys/s + achieved HBM GB/s (join-probe index)",
        "value": probe["mkeys_s"], "unit": "Mkeys/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ops["load"]["device_ms"] + probe["device_ms"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic (SplitMix64 seed 3)",
        "config": {"workload": "join-probe: FlashMap<[u8;32],u64> with_capacity(50M)=2^26 slots; bulk-load "
                               f"{nl} keys -> row id, then {npb} bulk_get probes (90% Zipf(1.0) hits, 10% misses)",
                   "parallelism": "single GPU",
                   "timing": "per op: kernel events on the map's stream summed (device_ms) and wall clock around "
                             "the synchronised call (wall_ms); medians over the timed steps"},
        "roofline": {"bound": "hbm", "op": "probe", "achieved": probe["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": probe["frac"], "traffic": probe.get("traffic_bytes")},
        "stream_copy": copy,
        "ops": ops,
        "get_autotune": get_tune,
        "cpu_baseline": cpu,
        "check": check,
    }
    return out_line, check is None or check["rows_exact"]


# ---------------------------------------------------------------------------
# Config 4: 512 independent FlashMap<u64,[u8;32]> tables, one step =
# bulk_insert -> bulk_get -> bulk_remove on every table.  The fused path (one
# hipGraph launch for the whole step over all tables) against the plain path
# (one bulk call per table-op on 512 separate maps), both from the same
# preloaded state, results compared table by table.
# ---------------------------------------------------------------------------
def run_c4(args):
    from flashmap_amd.workloads import config4_table

    n = args.tables
    log(f"c4: generating {n} tables")
    tabs = [config4_table(t, cap_hi=max(1 << 10, int((1 << 22) * args.scale))) for t in range(n)]
    caps = [t.capacity for t in tabs]
    V = 32

    def rows64(a):
        return np.ascontiguousarray(a, dtype=np.uint64).view(np.uint8).reshape(-1, 8)

    def cat(parts, w):
        return np.concatenate([np.ascontiguousarray(p).reshape(-1, w) for p in parts]) if parts else \
            np.zeros((0, w), np.uint8)

    snap = fm.FlashMapGroup(caps, 8, V)
    work = fm.FlashMapGroup(caps, 8, V)
    log(f"c4: preloading {sum(len(t.preload_keys) for t in tabs)} keys into {sum(caps)} slots (fused)")
    pk, pv = snap.upload(cat([rows64(t.preload_keys) for t in tabs], 8)), snap.upload(cat([t.preload_vals for t in tabs], V))
    new, status = snap.bulk_insert_device(pk, pv, [len(t.preload_keys) for t in tabs])
    assert all(int(new[i]) == len(tabs[i].preload_keys) for i in range(n))
    pk.free()
    pv.free()
    ic, gc, rc = [len(t.ins_keys) for t in tabs], [len(t.get_keys) for t in tabs], [len(t.rem_keys) for t in tabs]
    ik, iv = work.upload(cat([rows64(t.ins_keys) for t in tabs], 8)), work.upload(cat([t.ins_vals for t in tabs], V))
    gk, rk = work.upload(cat([rows64(t.get_keys) for t in tabs], 8)), work.upload(cat([rows64(t.rem_keys) for t in tabs], 8))
    out, found = work.alloc_device(max(1, sum(gc)) * V), work.alloc_device(max(1, sum(gc)))
    keys_per_step = sum(ic) + sum(gc) + sum(rc)

    # launch shapes: coordinate-descent search, each run on `work` restored
    # from the preloaded snapshot (fm_group_autotune), then used by `work`
    tune = None
    if not args.no_autotune:
        log("c4: autotuning the step's launch shapes")
        t0 = time.perf_counter()
        tune = snap.autotune(ik, iv, ic, gk, gc, out, found, rk, rc, reps=3, scratch=work)
        tune["search_s"] = round(time.perf_counter() - t0, 2)
        work.set_tuning(**tune["tuning"])

    # fused: warm up (captures the graph), then time restored repetitions
    reps = max(1, args.steps)
    fused_ms = []
    for r in range(reps + 1):
        work.copy_from(snap)
        work.download(out, 1)  # waits for the restore on the group's stream
        t0 = time.perf_counter()
        f_new, f_status, f_rem = work.step_device(ik, iv, ic, gk, gc, out, found, rk, rc)
        dt = time.perf_counter() - t0
        if r:
            fused_ms.append(dt * 1e3)
    f_vals = work.download(out, sum(gc) * V).reshape(-1, V)
    f_found = work.download(found, sum(gc)).astype(bool)
    f_slots = [hashlib.sha256(b"".join(a.tobytes() for a in work.export_slots(t))).hexdigest() for t in range(n)]

    # plain: 512 separate maps, one call per table-op
    log("c4: building plain per-table maps")
    maps = []
    for t, w in enumerate(tabs):
        m = fm.FlashMap.builder(w.capacity, 8, V).require_gpu().build()
        dk, dv = m.upload(rows64(w.preload_keys)), m.upload(w.preload_vals)
        if len(w.preload_keys):
            m.bulk_insert_device(dk, dv, len(w.preload_keys))
        dk.free()
        dv.free()
        bufs = dict(ik=m.upload(rows64(w.ins_keys)), iv=m.upload(w.ins_vals), gk=m.upload(rows64(w.get_keys)),
                    rk=m.upload(rows64(w.rem_keys)), out=m.alloc_device(max(1, len(w.get_keys)) * V),
                    found=m.alloc_device(max(1, len(w.get_keys))))
        maps.append((m, bufs))
    m0, _ = maps[0]
    p_new, p_status, p_rem = [0] * n, [0] * n, [0] * n
    t0 = time.perf_counter()
    for t, (m, b) in enumerate(maps):
        try:
            p_new[t] = m.bulk_insert_device(b["ik"], b["iv"], ic[t]) if ic[t] else 0
        except fm.TableFull:
            p_status[t] = fm.TABLE_FULL
        if gc[t]:
            m.bulk_get_device(b["gk"], gc[t], b["out"], b["found"])
        p_rem[t] = m.bulk_remove_device(b["rk"], rc[t]) if rc[t] else 0
    for m, _ in maps:
        m.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3

    # table-by-table comparison of the two paths
    exact = True
    off = 0
    for t, (m, b) in enumerate(maps):
        v = m.download(b["out"], gc[t] * V).reshape(-1, V) if gc[t] else np.zeros((0, V), np.uint8)
        f = m.download(b["found"], gc[t]).astype(bool) if gc[t] else np.zeros(0, bool)
        same = (int(f_new[t]) == p_new[t] and int(f_status[t]) == p_status[t] and int(f_rem[t]) == p_rem[t]
                and np.array_equal(v, f_vals[off:off + gc[t]]) and np.array_equal(f, f_found[off:off + gc[t]])
                and hashlib.sha256(b"".join(a.tobytes() for a in m.export_slots())).hexdigest() == f_slots[t])
        off += gc[t]
        if not same:
            exact = False
            log(f"c4: table {t} differs between fused and plain paths")
    # oracle on every table: the same three ops from the same preload, slot
    # arrays compared by digest (TableFull tables included)
    from oracle.binding import OracleMap
    oracle_ok, checked = True, 0
    if not args.no_check:
        for t in range(n):
            w = tabs[t]
            o = OracleMap(w.capacity, 8, V, 0)
            if len(w.preload_keys):
                o.bulk_insert(rows64(w.preload_keys), w.preload_vals)
            try:
                en, es = o.bulk_insert(rows64(w.ins_keys), w.ins_vals), 0
            except Exception:
                en, es = 0, fm.TABLE_FULL
            ov, of = o.bulk_get(rows64(w.get_keys))
            er = o.bulk_remove(rows64(w.rem_keys))
            ok = (en, es, er) == (int(f_new[t]), int(f_status[t]), int(f_rem[t])) and \
                hashlib.sha256(b"".join(a.tobytes() for a in o.export())).hexdigest() == f_slots[t]
            goff = int(sum(gc[:t]))
            ok = ok and np.array_equal(ov.reshape(-1, V), f_vals[goff:goff + gc[t]]) and \
                np.array_equal(of.astype(bool), f_found[goff:goff + gc[t]])
            oracle_ok &= ok
            checked += 1
            if not ok:
                log(f"c4: table {t} differs from the oracle")
            o.close()
    fused = float(np.median(fused_ms))
    cpu = None
    if not args.no_cpu_baseline:
        # Rayon restatement on every 8th table: each preloaded once, then its
        # insert -> get -> remove timed from the preloaded state (restored
        # untimed before each repetition); one warm-up, then the median of 5
        # of the summed step time.
        from oracle.binding import RayonPortMap, TableFull as OTF
        threads = cpu_threads(args)
        picked = list(range(0, n, 8))
        rep_s = [0.0] * 6
        cpu_keys = 0
        for t in picked:
            w = tabs[t]
            ps = RayonPortMap(w.capacity, 8, V, 0, threads=threads)
            if len(w.preload_keys):
                ps.bulk_insert(rows64(w.preload_keys), w.preload_vals)
            p = RayonPortMap(w.capacity, 8, V, 0, threads=threads)
            for rep in range(6):
                p.copy_from(ps)
                t0 = time.perf_counter()
                if len(w.ins_keys):
                    try:
                        p.bulk_insert(rows64(w.ins_keys), w.ins_vals)
                    except OTF:
                        pass
                if len(w.get_keys):
                    p.bulk_get(rows64(w.get_keys))
                if len(w.rem_keys):
                    p.bulk_remove(rows64(w.rem_keys))
                rep_s[rep] += time.perf_counter() - t0
            cpu_keys += len(w.ins_keys) + len(w.get_keys) + len(w.rem_keys)
            p.close()
            ps.close()
        cpu_s = float(np.median(rep_s[1:]))
        cpu = {"value": round(cpu_keys / cpu_s / 1e6, 3), "unit": "Mkeys/s", "cores": threads, "kind": "port",
               "sample": f"step of {len(picked)} of the {n} tables (every 8th): {cpu_keys} key-ops, one warm-up "
                         f"then the median of 5 ({cpu_s:.3f} s each), each table restored untimed from its "
                         f"preloaded state; {host_text()}"}
    out_line = {
        "metric": "batched multi-table step Mkeys/s (fused hipGraph vs plain per-table launches)",
        "value": round(keys_per_step / fused / 1e3, 3), "unit": "Mkeys/s", "n_gpus": 1,
        "plain_mkeys_s": round(keys_per_step / plain_ms / 1e3, 3),
        "speedup_vs_plain": round(plain_ms / fused, 2),
        "fused_ms": round(fused, 3), "plain_ms": round(plain_ms, 3), "reps": reps,
        "config": {"workload": f"{n} x FlashMap<u64,[u8;32]>, capacities log-uniform [2^10,2^22], load 0.5-0.85; "
                               "step = insert (log-uniform [1,2^16], 30% overwrite) + get (log-uniform [1,2^16], "
                               "70% hits) + remove (log-uniform [1,2^12]) per table",
                   "keys_per_step": keys_per_step, "slots": int(sum(caps)),
                   "table_full_batches": int(sum(1 for s in f_status if s))},
        "tuning": {k: getattr(work.tuning(), k) for k in ("group", "block", "grid", "coop_arc")},
        "autotune": tune,
        "cpu_baseline": cpu,
        "check": {"fused_equals_plain": exact, "oracle_tables": checked, "oracle_exact": bool(oracle_ok)},
    }
    return out_line, bool(exact and oracle_ok)


# ---------------------------------------------------------------------------
# Config 5: serving loop over one FlashMap<[u8;32],u64> (2^25 slots,
# 20M preloaded), Poisson arrivals at several offered loads.
# ---------------------------------------------------------------------------
def run_c5(args):
    from flashmap_amd.workloads import config5_preload, config5_requests

    s = args.scale
    cap, pk, pv = config5_preload(n_preload=int(20_000_000 * s), capacity=int(24_000_000 * s))
    snap = fm.FlashMap.builder(cap, 32, 8).require_gpu().build()
    m = fm.FlashMap.builder(cap, 32, 8).require_gpu().build()
    log(f"c5: preloading {len(pk)} keys")
    dk, dv = snap.upload(pk), snap.upload(pv)
    snap.bulk_insert_device(dk, dv, len(pk))
    dk.free()
    dv.free()
    lines, pageable = [], []
    osnap = None
    for ri, rate in enumerate(float(x) for x in args.rates.split(",")):
        reqs = config5_requests(pk, rate, args.requests)
        # Pageable request buffers: the server copies them through its
        # pinned staging ring.  Reported beside the main (pinned) figures.
        m.copy_from(snap)
        m.synchronize()
        log(f"c5: serving {len(reqs)} requests at {rate:.0f}/s (pageable request buffers)")
        _, stp = fm.serve_log(m, reqs, staging_bytes=1 << 30, max_batch_keys=1 << 21, realtime=True)
        pageable.append({"offered_req_s": rate, "achieved_req_s": round(len(reqs) / stp.wall_s, 1),
                         "throughput_mkeys_s": round(stp.keys / stp.wall_s / 1e6, 2),
                         "p50_ms": round(stp.p50_ms, 3), "p99_ms": round(stp.p99_ms, 3)})
        # Request buffers in pinned memory (fm_host_alloc, filled before the
        # first arrival): keys, values and results move by DMA.
        m.copy_from(snap)
        m.synchronize()
        log(f"c5: serving {len(reqs)} requests at {rate:.0f}/s (pinned request buffers)")
        results, st = fm.serve_log(m, reqs, staging_bytes=1 << 30, max_batch_keys=1 << 21, realtime=True,
                                   pinned=True)
        check = None
        if not args.no_check:
            # every offered load's log replayed in arrival order by the
            # sequential oracle, from the same preloaded state
            from oracle.binding import OracleMap, TableFull as OTF
            if osnap is None:
                o = OracleMap(cap, 32, 8, 0)
                o.bulk_insert(pk, pv)
                osnap = (o.export(), len(o))
            else:
                o = OracleMap(cap, 32, 8, 0)
                o.import_slots(*osnap[0], osnap[1])
            bad = 0
            for r, res in zip(reqs, results):
                if r.op == "get":
                    v, f = o.bulk_get(r.keys)
                    bad += int(not (np.array_equal(v.reshape(-1, 8), res.values) and
                                    np.array_equal(f.astype(bool), res.found.astype(bool))))
                elif r.op == "insert":
                    try:
                        bad += int(o.bulk_insert(r.keys, r.vals) != res.count or res.status != 0)
                    except OTF:
                        bad += int(res.status != fm.TABLE_FULL)
                else:
                    bad += int(o.bulk_remove(r.keys) != res.count)
            check = {"requests_checked": len(reqs), "mismatches": bad}
            o.close()
        lines.append({"offered_req_s": rate, "requests": len(reqs), "keys": int(st.keys),
                      "throughput_mkeys_s": round(st.keys / st.wall_s / 1e6, 2),
                      "achieved_req_s": round(len(reqs) / st.wall_s, 1), "p50_ms": round(st.p50_ms, 3),
                      "p99_ms": round(st.p99_ms, 3), "mean_ms": round(st.mean_ms, 3), "batches": int(st.batches),
                      "max_inflight_batches": int(st.max_inflight_batches), "staging_waits": int(st.staging_waits),
                      "check": check})
    cpu = None
    if not args.no_cpu_baseline:
        # Rayon restatement applying the first load's requests back to back in
        # arrival order (no arrival gaps: its throughput ceiling), from the
        # preloaded state restored untimed before each repetition; one
        # warm-up, then the median of 5.
        from oracle.binding import RayonPortMap, TableFull as OTF
        threads = cpu_threads(args)
        reqs = config5_requests(pk, float(args.rates.split(",")[0]), args.requests)
        psnap = RayonPortMap(cap, 32, 8, 0, threads=threads)
        psnap.bulk_insert(pk, pv)
        p = RayonPortMap(cap, 32, 8, 0, threads=threads)
        keys = sum(len(r.keys) for r in reqs)
        times = []
        for rep in range(6):
            p.copy_from(psnap)
            t0 = time.perf_counter()
            for r in reqs:
                if r.op == "get":
                    p.bulk_get(r.keys)
                elif r.op == "insert":
                    try:
                        p.bulk_insert(r.keys, r.vals)
                    except OTF:
                        pass
                else:
                    p.bulk_remove(r.keys)
            if rep:
                times.append(time.perf_counter() - t0)
        p.close()
        psnap.close()
        dt = float(np.median(times))
        cpu = {"value": round(len(reqs) / dt, 1), "unit": "req/s", "mkeys_s": round(keys / dt / 1e6, 2),
               "cores": threads, "kind": "port",
               "sample": f"the {len(reqs)} requests of the first offered load applied back to back in arrival "
                         f"order ({keys} keys) on the 20M-key preload, one warm-up then the median of 5 "
                         f"({dt:.2f} s each); {host_text()}"}
    ok = all(l["check"] is None or l["check"]["mismatches"] == 0 for l in lines)
    return ({"metric": "serving p50/p99 latency and throughput at offered loads", "n_gpus": 1,
                      "config": {"workload": "FlashMap<[u8;32],u64> with_capacity(24M)=2^25, preload 20M; requests "
                                             "70% get / 20% insert / 10% remove, log-uniform [64,262144] keys, gets "
                                             "80% hits Zipf(0.9)", "staging_bytes": 1 << 30,
                                 "request_buffers": "loads: pinned (fm_host_alloc, DMA both ways); "
                                                    "loads_pageable: pageable (host copy through the staging ring)"},
             "loads": lines, "loads_pageable": pageable, "cpu_baseline": cpu}, ok)



# ---------------------------------------------------------------------------
# Host-buffer calls on the shape of the reference's published figures
# (README.md:9-12: 1M inserts, 1M lookups; its examples use
# [u8;32] -> [u8;128]).  Wall clock around fm_bulk_insert / fm_bulk_get with
# host buffers, PCIe transfers and the count / result read-back included
# (the reference's gpu.rs:111-273 path).  Never the metric value.
# ---------------------------------------------------------------------------
def run_host_api(args):
    from flashmap_amd.workloads import Stream

    n = int(1_000_000 * args.scale)
    st = Stream(6)
    keys = st.bytes(n, 32)
    vals = st.bytes(n, 128)
    out = {"pairs": n, "key_bytes": 32, "value_bytes": 128,
           "timing": "wall clock around the host-buffer call (H2D, kernels, D2H, sync), cleared table before "
                     "each insert; one warm-up, then the median of 5",
           "buffers": "pageable: numpy arrays, results allocated by bulk_get; pinned: keys, values and "
                      "the get's results in one PinnedArena (fm_host_alloc)",
           "reference_h100_ms": {"insert": 60, "get": 35}}
    ok = True
    for label, capacity in (("cap_1M", 1_000_000), ("cap_2M", 2_000_000)):
        m = fm.FlashMap.builder(capacity, 32, 128).require_gpu().build()
        assert m.backend() == fm.Backend.GPU
        res = {}
        for mode in ("pageable", "pinned"):
            arena = None
            hk, hv, outs = keys, vals, {}
            if mode == "pinned":  # fm_host_alloc buffers both ways: moved by DMA, no staging copy
                arena = fm.PinnedArena(keys.nbytes + 2 * vals.nbytes + n + 4096)
                hk, hv = arena.copy_of(keys), arena.copy_of(vals)
                outs = {"out": arena.take(vals.shape, fill=None), "found": arena.take((n,), fill=None).view(np.bool_)}
            ti, tg = [], []
            for rep in range(6):
                m.clear()
                m.synchronize()
                t0 = time.perf_counter()
                new = m.bulk_insert(hk, hv)
                t1 = time.perf_counter()
                gv, gf = m.bulk_get(hk, **outs)
                t2 = time.perf_counter()
                if rep == 0:
                    good = new == n and bool(gf.all()) and np.array_equal(gv, vals)
                    ok &= good
                    res[f"{mode}_exact"] = bool(good)
                else:
                    ti.append((t1 - t0) * 1e3)
                    tg.append((t2 - t1) * 1e3)
            res[f"{mode}_insert_ms"] = round(float(np.median(ti)), 3)
            res[f"{mode}_get_ms"] = round(float(np.median(tg)), 3)
            if arena is not None:
                del hk, hv, outs, gv, gf
                arena.close()
        res["slots"] = m.capacity()
        res["load_after_insert"] = round(n / m.capacity(), 3)
        out[label] = res
        del m
        gc.collect()
    return out, ok


# ---------------------------------------------------------------------------
# Default run: the account-state metric line plus one compact section per
# other configuration (the driver keeps the line's text, so it stays short;
# --detail PATH writes every configuration's full line).
# ---------------------------------------------------------------------------
def _r(x, nd=3):
    return None if x is None else round(x, nd)


def summarize(name, d):
    if name == "c1":
        o = d["ops"]
        s = {"device_us": {k: _r(o[k]["device_ms"] * 1e3, 1) for k in o},
             "mkeys_s": {k: _r(o[k]["mkeys_s"], 0) for k in o},
             "hbm_frac": {k: _r(o[k]["frac"], 4) for k in o},
             "traffic_x_alg": {k: o[k].get("traffic_x_alg") for k in o},
             "host_api_ms": {k: o[k]["host_api_ms"] for k in o},
             "check": d["check"]}
        if d.get("cpu_baseline"):
            s["cpu"] = {"mkeys_s": d["cpu_baseline"]["value"], "cores": d["cpu_baseline"]["cores"],
                        "per_op": d["cpu_baseline"]["detail"]["port"]["mkeys_s"],
                        "seq_1_thread_mkeys_s": d["cpu_baseline"]["detail"]["sequential"]["sequence_mkeys_s"]}
        return s
    if name == "c3":
        o = d["ops"]
        s = {k: {"device_ms": o[k]["device_ms"], "mkeys_s": _r(o[k]["mkeys_s"], 1), "hbm_frac": _r(o[k]["frac"], 4),
                 "fabric_gbs": o[k].get("hbm_measured_gbs")} for k in ("load", "probe")}
        s["check"] = d["check"]
        if d.get("cpu_baseline"):
            s["cpu"] = {"probe_mkeys_s": d["cpu_baseline"]["value"], "load_mkeys_s": d["cpu_baseline"]["load_mkeys_s"],
                        "cores": d["cpu_baseline"]["cores"]}
        return s
    if name == "c4":
        s = {k: d[k] for k in ("fused_ms", "plain_ms", "speedup_vs_plain", "value", "plain_mkeys_s")}
        s["fused_mkeys_s"] = s.pop("value")
        s["keys_per_step"] = d["config"]["keys_per_step"]
        s["table_full_batches"] = d["config"]["table_full_batches"]
        s["tuning"] = d["tuning"]
        s["check"] = d["check"]
        if d.get("cpu_baseline"):
            s["cpu"] = {"mkeys_s": d["cpu_baseline"]["value"], "cores": d["cpu_baseline"]["cores"],
                        "tables": "every 8th"}
        return s
    if name == "c5":
        s = {"loads": [{"offered": l["offered_req_s"], "served_req_s": l["achieved_req_s"],
                        "mkeys_s": l["throughput_mkeys_s"], "p50_ms": l["p50_ms"], "p99_ms": l["p99_ms"],
                        "checked": l["check"] and l["check"]["requests_checked"],
                        "mismatches": l["check"] and l["check"]["mismatches"]} for l in d["loads"]],
             "pageable": [{"offered": l["offered_req_s"], "served_req_s": l["achieved_req_s"],
                           "p50_ms": l["p50_ms"], "p99_ms": l["p99_ms"]} for l in d["loads_pageable"]],
             "buffers": "pinned (fm_host_alloc); pageable beside"}
        if d.get("cpu_baseline"):
            s["cpu"] = {"req_s": d["cpu_baseline"]["value"], "mkeys_s": d["cpu_baseline"]["mkeys_s"],
                        "cores": d["cpu_baseline"]["cores"]}
        return s
    if name == "host_api":
        return {k: v for k, v in d.items() if k not in ("timing", "buffers")}
    return d


def run_all(args):
    t_start = time.perf_counter()
    line, ok = run_c2(args)
    detail = {"c2": dict(line)}
    # the c2 line carries its full per-kernel table in --detail only
    line["kernels_ms"] = {k: v for k, v in line["kernels_ms"].items()
                          if k in ("k_get", "k_ins_dry", "k_ins_leader", "k_ins_big", "k_ins_bits", "k_rem_probe")}
    sections = (("c1", run_c1, {"steps": 10, "warmup": 2}), ("c3", run_c3, {"steps": 3, "warmup": 1}),
                ("c4", run_c4, {"steps": 10}), ("c5", run_c5, {}), ("host_api", run_host_api, {}))
    for name, fn, over in sections:
        gc.collect()
        sargs = argparse.Namespace(**{**vars(args), **over})
        t0 = time.perf_counter()
        log(f"section {name}")
        try:
            d, sok = fn(sargs)
            detail[name] = d
            sec = summarize(name, d)
        except Exception as e:  # reported in the line; the other sections still run
            log(f"section {name} failed: {type(e).__name__}: {e}")
            sec, sok = {"error": f"{type(e).__name__}: {e}"}, False
        sec["ok"] = bool(sok)
        sec["s"] = round(time.perf_counter() - t0, 1)
        line[name] = sec
        ok = ok and sok
    try:
        c1, c4, c5 = line["c1"], line["c4"], line["c5"]
        line["targets"] = {
            "c1_hbm_frac": {"get": c1["hbm_frac"]["get"], "insert": c1["hbm_frac"]["insert"], "target": 0.5},
            "c4_speedup_vs_plain": {"value": c4["speedup_vs_plain"], "target": 3.0},
            "c5_p99_ms": {str(int(l["offered"])): l["p99_ms"] for l in c5["loads"]},
            "all_checks_pass": bool(ok)}
    except (KeyError, TypeError):
        line["targets"] = {"all_checks_pass": bool(ok)}
    line["bench_s"] = round(time.perf_counter() - t_start, 1)
    if args.detail:
        with open(args.detail, "w") as f:
            json.dump(detail, f, indent=1)
    print(json.dumps(line, separators=(",", ":")), flush=True)
    sys.exit(0 if ok else 1)

if __name__ == "__main__":
    main()
