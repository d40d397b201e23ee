"""The device sort that groups keys by closure for the sort-grouped insert
(csrc/fm_sort.h; a batch that may fill the table, forced here with
FM_INSERT_SORT_PATH=1), compared with the oracle, and large remove batches
with duplicate keys (their conflicting runs replayed in batch order),
counted and inside the serving loop.

Sizes cover the one-workgroup LDS sort (n <= 16384), and radix sorts of one
pass (ids below 2^11), two passes (even: the first pass writes the scratch
buffer) and three passes (odd: the first pass writes the output), with
count matrices of one and of many scan chunks."""
import sys
from types import SimpleNamespace

import numpy as np
import pytest

import flashmap_amd as fm
from helpers import OracleAdapter, apply_and_compare

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_make(cap, K, V):
    m = fm.FlashMap.builder(cap, K, V).require_gpu().build()
    assert m.backend() == fm.Backend.GPU
    return m


def rows(rng, n, K, space):
    ks = rng.integers(0, space, size=n).astype(np.uint64)
    keys = np.zeros((n, K), dtype=np.uint8)
    keys[:, :min(8, K)] = ks.view(np.uint8).reshape(-1, 8)[:, :min(8, K)]
    return keys


@pytest.mark.parametrize("cap,n", [(1 << 12, 3000), (1 << 15, 16384), (1 << 18, 100_000), (1 << 22, 300_000)])
def test_sort_grouped_insert_large_batches(cap, n, monkeypatch):
    # closure ids in [0, cap]: the two small batches sort in LDS, then end bits
    # 19 (two radix passes) and 23 (three)
    say(f"sort-grouped insert cap={cap} n={n}")
    monkeypatch.setenv("FM_INSERT_SORT_PATH", "1")
    m = gpu_make(cap, 8, 16)
    monkeypatch.delenv("FM_INSERT_SORT_PATH")
    o = OracleAdapter(cap, 8, 16)
    rng = np.random.default_rng(cap + n)
    space = cap * 2
    ops = []
    for b in range(3):  # collision-heavy keys with in-batch duplicates, then removes and refills
        k = rows(rng, n // (b + 1), 8, space)
        ops.append(("ins", k, rng.integers(0, 256, (len(k), 16), dtype=np.uint8)))
        ops.append(("rem", k[: len(k) // 3]))
    ops.append(("get", rows(rng, n, 8, space)))
    m.profile(True)
    apply_and_compare(m, o, ops)
    assert m.profile_read().get("ins_radix_sort", (0, 0))[0] >= 3  # every insert took the sorted path


@pytest.mark.parametrize("cap,n", [(1 << 10, 20_000), (1 << 16, 12_000), (1 << 16, 60_000), (1 << 20, 200_000)])
def test_remove_conflict_replay_large(cap, n):
    # duplicate removes force the exact replay of their runs (k_rem_check
    # flags a conflict)
    say(f"remove conflict replay cap={cap} n={n}")
    m, o = gpu_make(cap, 8, 8), OracleAdapter(cap, 8, 8)
    rng = np.random.default_rng(n ^ cap)
    fill = rows(rng, int(cap * 0.7), 8, cap * 4)
    ops = [("ins", fill, rng.integers(0, 256, (len(fill), 8), dtype=np.uint8))]
    rk = fill[rng.integers(0, len(fill), n)]  # many keys removed twice or more
    ops += [("rem", rk), ("get", fill)]
    more = rows(rng, cap // 4, 8, cap * 4)
    ops += [("ins", more, rng.integers(0, 256, (len(more), 8), dtype=np.uint8)), ("rem", np.concatenate([more, more]))]
    apply_and_compare(m, o, ops)


def test_served_removes_with_duplicates_large():
    # serving removes run the gated replay with no readback
    say("served large duplicate removes")
    cap = 1 << 18
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 256, (150_000, 32), dtype=np.uint8)
    vals = rng.integers(0, 256, (150_000, 8), dtype=np.uint8)
    m, o = gpu_make(cap, 32, 8), OracleAdapter(cap, 32, 8)
    assert m.bulk_insert(keys, vals) == o.bulk_insert(keys, vals)
    rem = keys[rng.integers(0, len(keys), 40_000)]
    reqs = [SimpleNamespace(op="remove", keys=rem, arrival_s=0.0), SimpleNamespace(op="get", keys=keys, arrival_s=0.0),
            SimpleNamespace(op="remove", keys=np.concatenate([rem[:9000], rem[:9000]]), arrival_s=0.0),
            SimpleNamespace(op="get", keys=keys, arrival_s=0.0)]
    results, _ = fm.serve_log(m, reqs, staging_bytes=64 << 20, max_batch_keys=1 << 20, realtime=False)
    for r, res in zip(reqs, results):
        if r.op == "remove":
            assert res.count == o.bulk_remove(r.keys)
        else:
            v, f = o.bulk_get(r.keys)
            assert np.array_equal(res.found.astype(bool), f.astype(bool))
            assert np.array_equal(res.values, v.reshape(-1, 8))
