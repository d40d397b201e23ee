#!/usr/bin/env python3
"""FlashMap bulk-op benchmark on one MI355X.

Workload (the account-state configuration): FlashMap<[u8;32],[u8;128]>
with_capacity(10,000,000) (2^24 slots), preloaded with 8,000,000 random
pairs.  One step applies, in order, 20 bulk_insert batches of 100,000 pairs
(50% overwrites of existing keys, 50% new keys, no duplicates inside a batch),
one bulk_get of 1,000,000 queries at a 90% hit rate and one bulk_remove of
200,000 existing keys: 3,200,000 key-ops.  Every step starts from the
preloaded state: two working tables alternate, and while one runs a step the
other is restored from a device snapshot on its own stream.  Inputs are
resident in HBM before the timed region; results stay on the device.

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
import os
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


def cpu_baseline(cfg, threads):
    """Rayon-port timing of one step (after an untimed preload)."""
    from oracle.binding import RayonPortMap

    p = RayonPortMap(cfg.capacity, K_BYTES, V_BYTES, 0, threads=threads)
    p.bulk_insert(cfg.preload_keys, cfg.preload_vals)
    t0 = time.perf_counter()
    for k, v in cfg.batches:
        p.bulk_insert(k, v)
    p.bulk_get(cfg.get_keys)
    p.bulk_remove(cfg.remove_keys)
    dt = time.perf_counter() - t0
    keys = sum(len(k) for k, _ in cfg.batches) + len(cfg.get_keys) + len(cfg.remove_keys)
    p.close()
    return keys / dt / 1e6, dt


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, host cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the workload (testing only)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the table would shard by "
              "slot range; see the sharding design). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

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
    if not args.no_check:
        log("checking one step against the oracle")
        A.copy_from(snap)
        new, removed = step.run(A)
        vals = A.download(step.out_vals)
        found = A.download(step.out_found)
        exp = oracle_step(cfg)
        got_sha = hashlib.sha256(vals.tobytes() + found.tobytes()).hexdigest()
        ok = new == exp["new"] and removed == exp["removed"] and got_sha == exp["get_sha"] and len(A) == exp["len"]
        check = {"bit_exact": bool(ok), "hits": exp["hits"], "new_per_batch": exp["new"][:3],
                 "removed": removed}
        if not ok:
            log(f"PARITY FAILURE: gpu new={new[:3]} removed={removed} sha={got_sha[:12]} len={len(A)}; "
                f"oracle new={exp['new'][:3]} removed={exp['removed']} sha={exp['get_sha'][:12]} len={exp['len']}")
            sys.exit(1)

    maps = [A, B]
    A.copy_from(snap)
    B.copy_from(snap)
    for i in range(args.warmup):
        m = maps[i % 2]
        step.run(m)
        m.copy_from(snap)
    for m in maps:
        m.synchronize()
        m.profile(True)
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    for i in range(args.steps):
        m = maps[(args.warmup + i) % 2]
        step.run(m)
        m.copy_from(snap)
    for m in maps:
        m.synchronize()
    elapsed = time.perf_counter() - t0
    kstats = {}
    for m in maps:
        for name, (n, ms) in m.profile_read().items():
            a = kstats.setdefault(name, [0, 0.0])
            a[0] += n
            a[1] += ms
        m.profile(False)

    keys = step.keys_per_step * args.steps
    value = keys / elapsed / 1e6
    ms_per_step = elapsed * 1e3 / args.steps

    # Dominant kernel (largest summed device time, restore copies excluded).
    per_launch = {"k_get": B_GET * step.n_get, "k_ins_replay": B_INS * len(cfg.batches[0][0]),
                  "k_ins_dry": B_INS * len(cfg.batches[0][0]), "k_ins_closure": B_INS * len(cfg.batches[0][0]),
                  "ins_radix_sort": B_INS * len(cfg.batches[0][0]), "ins_bitmap_copy": B_INS * len(cfg.batches[0][0]),
                  "k_rem_probe": B_REM * step.n_rem, "k_rem_check": B_REM * step.n_rem,
                  "k_rem_apply": B_REM * step.n_rem}
    ops_only = {k: v for k, v in kstats.items() if k in per_launch}
    dom = max(ops_only, key=lambda k: ops_only[k][1])
    dn, dms = ops_only[dom]
    avg_s = dms / dn / 1e3
    achieved = per_launch[dom] / avg_s / 1e9
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", f"{dom}_c2_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    def op_rate(names, keys_per_launch, launches_key):
        n = kstats.get(launches_key, [0, 0])[0]
        ms = sum(kstats.get(k, [0, 0.0])[1] for k in names)
        return (keys_per_launch * n / (ms / 1e3) / 1e6) if ms else None

    ins_names = ["ins_bitmap_copy", "k_ins_dry", "k_ins_closure", "ins_radix_sort", "k_ins_replay"]
    rem_names = ["k_rem_probe", "k_rem_check", "k_rem_apply", "k_rem_runs", "rem_radix_sort", "k_rem_replay"]
    get_rate = op_rate(["k_get"], step.n_get, "k_get")
    ins_rate = op_rate(ins_names, len(cfg.batches[0][0]), "k_ins_replay")
    rem_rate = op_rate(rem_names, step.n_rem, "k_rem_probe")

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"CPU baseline (Rayon restatement, {threads} threads)")
        rate, dt = cpu_baseline(cfg, threads)
        cpu = {"value": round(rate, 3), "unit": "Mkeys/s", "cores": threads, "kind": "port",
               "sample": f"one step ({step.keys_per_step} key-ops) after an untimed 8M preload; "
                         f"{dt:.2f} s; host '{host_cpu_model()}'"}

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
                   "keys_per_step": step.keys_per_step, "parallelism": "single GPU"},
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
        "check": check,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
