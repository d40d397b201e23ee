"""Exact removes with conflicts on the GPU, against the oracle.

A remove batch conflicts where an earlier key tombstones the slot a later
key's pre-batch probe ended on (a duplicate key matching the same slot, or
the occupied terminator of a later key).  Keys interact only inside the run
that ends at the first EMPTY slot at/after their home, so k_rem_check marks
the conflicting runs dirty, k_rem_apply applies every clean run and lists
the dirty runs' keys, and k_rem_replay replays each dirty run in batch order
(kRunKeys keys at a time for long runs); with no EMPTY slot anywhere the
whole batch is one run.  Every case compares counts, lengths and the whole
slot array with the oracle after each op, through the counted call and
through the serving loop (asynchronous removes, counts from outcome flags).
"""
import sys
from types import SimpleNamespace

import numpy as np
import pytest

import flashmap_amd as fm
from helpers import OracleAdapter, apply_and_compare, assert_same_slots, le

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_make(cap, K, V):
    m = fm.FlashMap.builder(cap, K, V).require_gpu().build()
    assert m.backend() == fm.Backend.GPU
    return m


def keys_of(ids, K):
    return le(ids, K)


def distinct(rng, n, space):
    ids = np.unique(rng.integers(0, space, n * 2 + 64))
    return rng.permutation(ids)[:n]


@pytest.mark.parametrize("cap,K,V", [(1 << 12, 8, 8), (1 << 12, 32, 8), (1 << 14, 32, 128)])
def test_long_dirty_runs(cap, K, V):
    # 95% full: runs of hundreds of slots; every key removed twice, so every
    # run is dirty and holds far more than 64 keys
    say(f"long dirty runs cap={cap} K={K} V={V}")
    rng = np.random.default_rng(cap + K)
    m, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    n = int(cap * 0.95)
    ids = distinct(rng, n, 1 << 30)
    ks = keys_of(ids, K)
    vs = rng.integers(0, 256, (n, V), dtype=np.uint8)
    ops = [("ins", ks, vs)]
    twice = np.concatenate([ks, ks])[rng.permutation(2 * n)]
    ops += [("rem", twice), ("get", ks)]
    back = ks[: n // 2]
    ops += [("ins", back, vs[: n // 2]), ("rem", np.concatenate([back[::3], back[::3], ks[n // 2:]])), ("get", ks)]
    apply_and_compare(m, o, ops)


def test_no_empty_slot_is_one_run():
    # every slot OCCUPIED or TOMBSTONE: a conflict makes the batch one run
    say("no EMPTY slot")
    cap = 1024
    rng = np.random.default_rng(3)
    m, o = gpu_make(cap, 8, 8), OracleAdapter(cap, 8, 8)
    ids = distinct(rng, cap, 1 << 20)
    ks = keys_of(ids, 8)
    vs = rng.integers(0, 256, (cap, 8), dtype=np.uint8)
    ops = [("ins", ks, vs), ("rem", np.concatenate([ks[:100], ks[50:150]])), ("get", ks)]
    more = keys_of(distinct(rng, 150, 1 << 20) + (1 << 21), 8)
    ops += [("ins", more, vs[:150]), ("rem", np.concatenate([more[::2], more[::2], ks[200:300]])), ("get", ks),
            ("get", more)]
    apply_and_compare(m, o, ops)


def test_few_dirty_runs_among_clean_ones():
    # a large batch with a handful of duplicates: most runs are applied by
    # k_rem_apply, a few replayed
    say("few dirty runs")
    cap = 1 << 20
    rng = np.random.default_rng(11)
    m, o = gpu_make(cap, 32, 8), OracleAdapter(cap, 32, 8)
    n = 600_000
    ks = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    vs = rng.integers(0, 256, (n, 8), dtype=np.uint8)
    rem = ks[rng.permutation(n)[:200_000]]
    dup = rem[rng.integers(0, len(rem), 40)]
    batch = np.concatenate([rem, dup])[rng.permutation(len(rem) + 40)]
    apply_and_compare(m, o, [("ins", ks, vs), ("rem", batch), ("get", ks[:300_000])])


@pytest.mark.parametrize("seed", range(6))
def test_collision_heavy_remove_histories(seed):
    # tiny tables and key spaces: duplicate matches and removed terminators
    # in nearly every batch, counted and served
    say(f"collision-heavy removes seed={seed}")
    rng = np.random.default_rng(100 + seed)
    cap = int(rng.choice([64, 128, 256, 512]))
    space = cap * 2
    m, o = gpu_make(cap, 8, 8), OracleAdapter(cap, 8, 8)
    for step in range(40):
        n = int(rng.integers(1, cap))
        ks = keys_of(rng.integers(0, space, n), 8)
        if rng.random() < 0.5 and len(o) + n <= cap:
            vs = rng.integers(0, 256, (n, 8), dtype=np.uint8)
            assert m.bulk_insert(ks, vs) == o.bulk_insert(ks, vs), f"step {step} insert"
        else:
            assert m.bulk_remove(ks) == o.bulk_remove(ks), f"step {step} remove"
        assert len(m) == len(o)
        assert_same_slots(m, o, f"step {step}")
    # the same kind of history through the serving loop
    reqs = []
    for step in range(40):
        n = int(rng.integers(1, cap // 2))
        ks = keys_of(rng.integers(0, space, n), 8)
        reqs.append(SimpleNamespace(op="remove" if rng.random() < 0.6 else "get", keys=ks, arrival_s=0.0))
    results, _ = fm.serve_log(m, reqs, staging_bytes=8 << 20, max_batch_keys=1 << 16, realtime=False)
    for i, (r, res) in enumerate(zip(reqs, results)):
        if r.op == "remove":
            assert res.count == o.bulk_remove(r.keys), f"served remove {i}"
        else:
            v, f = o.bulk_get(r.keys)
            assert np.array_equal(res.found.astype(bool), f.astype(bool)), f"served get {i}"
            assert np.array_equal(res.values, v.reshape(-1, 8)), f"served get {i} values"
    assert len(m) == len(o)
    assert_same_slots(m, o, "after the served log")
