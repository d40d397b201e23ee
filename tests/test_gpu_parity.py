"""GPU (HIP, gfx950) backend parity with the oracle.  Needs an MI355X.

Every map here is built with require_gpu() and asserted to be the GPU arm,
so a silent CPU fallback cannot pass.  Results are compared bit for bit with
the sequential CPU oracle: counts, found flags, values, and the whole logical
slot array (flags with distances, keys and values of occupied slots).
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import flashmap_amd as fm
from helpers import (ANOMALY_CASES, REFERENCE_TESTS, OracleAdapter, apply_and_compare,
                     assert_same_slots, random_history, run_history)

pytestmark = pytest.mark.gpu
GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_make(cap, K, V, mode=0, group=None, block=None):
    m = fm.FlashMap.builder(cap, K, V).hash_strategy(mode).require_gpu().build()
    assert m.backend() == fm.Backend.GPU
    if group or block:
        m.set_tuning(get_group=group, remove_group=group, get_block=block, insert_block=block)
    return m


@pytest.mark.parametrize("test", REFERENCE_TESTS, ids=lambda t: t.__name__[2:])
def test_reference_tests_on_gpu(test):
    say(f"reference test {test.__name__}")
    test(lambda c, K, V, mode: gpu_make(c, K, V, mode))


@pytest.mark.parametrize("name", sorted(ANOMALY_CASES))
def test_anomaly_histories_gpu(name):
    ops, expected = ANOMALY_CASES[name]
    assert run_history(gpu_make(16, 8, 8), ops) == expected


@pytest.mark.parametrize("seed", range(16))
def test_gpu_vs_oracle_random_histories(seed):
    rng = np.random.default_rng(1000 + seed)
    cap = [16, 64, 256, 1024, 4096][seed % 5]
    K, V = [(8, 8), (32, 128), (32, 8), (8, 32), (3, 5), (12, 20), (32, 64), (16, 16)][seed % 8]
    mode = (seed // 2) % 2
    group = [1, 2, 4, 8, 16][seed % 5]
    say(f"random history seed={seed} cap={cap} K={K} V={V} mode={mode} group={group}")
    ops = random_history(rng, cap, K, V, 24, max(1, cap // 2), cap * 2)
    apply_and_compare(gpu_make(cap, K, V, mode, group=group), OracleAdapter(cap, K, V, mode), ops)


@pytest.mark.parametrize("seed", range(6))
def test_gpu_sort_grouped_insert_path(seed, monkeypatch):
    """The radix-sort grouping (used when a batch may fill the table) on
    ordinary histories, forced with FM_INSERT_SORT_PATH=1 at map creation."""
    monkeypatch.setenv("FM_INSERT_SORT_PATH", "1")
    rng = np.random.default_rng(2000 + seed)
    cap = [64, 256, 1024][seed % 3]
    K, V = [(8, 8), (32, 128), (3, 5)][seed % 3]
    say(f"sort path seed={seed} cap={cap} K={K} V={V}")
    ops = random_history(rng, cap, K, V, 20, max(1, cap // 2), cap * 2)
    m = gpu_make(cap, K, V, seed % 2)
    monkeypatch.delenv("FM_INSERT_SORT_PATH")
    apply_and_compare(m, OracleAdapter(cap, K, V, seed % 2), ops)


@pytest.mark.parametrize("seed", range(10))
def test_gpu_wave_walk_all_closures(seed):
    """insert_coop_arc = 0 sends every closure through the wave-cooperative
    walk (windows of 64 slots, LDS carry across windows)."""
    rng = np.random.default_rng(3000 + seed)
    cap = [64, 256, 1024, 4096, 1 << 14][seed % 5]
    K, V = [(8, 8), (32, 128), (32, 8), (8, 32), (16, 16)][seed % 5]
    say(f"wave walk seed={seed} cap={cap} K={K} V={V}")
    ops = random_history(rng, cap, K, V, 20, max(1, cap // 2), cap * 2)
    m = gpu_make(cap, K, V, seed % 2)
    m.set_tuning(insert_coop_arc=0)
    apply_and_compare(m, OracleAdapter(cap, K, V, seed % 2), ops)


def test_gpu_wave_walk_crowded_and_config1():
    from flashmap_amd.workloads import config1

    say("wave walk crowded + config1")
    rng = np.random.default_rng(78)
    cap = 1 << 13
    base = rng.integers(0, 1 << 40, size=3200, dtype=np.int64).astype(np.uint64)
    rows = np.unique((base << np.uint64(13)) | (base & np.uint64(0x3FF)))[:3000].view(np.uint8).reshape(-1, 8)
    m = gpu_make(cap, 8, 8)
    m.set_tuning(insert_coop_arc=0)
    apply_and_compare(m, OracleAdapter(cap, 8, 8), [("ins", rows[:1500], rows[:1500]), ("rem", rows[:500]),
                                                    ("ins", rows[1500:], rows[:1500]), ("get", rows)])
    c = config1(n_insert=200_000, n_query=100_000, n_remove=20_000, capacity=1 << 18)
    m = gpu_make(c.capacity, 8, 8)
    m.set_tuning(insert_coop_arc=0)
    apply_and_compare(m, OracleAdapter(c.capacity, 8, 8), [("ins", c.insert_keys, c.insert_vals), ("get", c.queries),
                                                           ("rem", c.remove_keys), ("ins", c.remove_keys, c.remove_keys)])


@pytest.mark.parametrize("seed", range(4))
def test_gpu_fill_to_full_no_barrier(seed):
    """Batches that consume every free slot (single closure, no barrier)."""
    rng = np.random.default_rng(50 + seed)
    cap = 256
    say(f"fill to full seed={seed}")
    m, o = gpu_make(cap, 8, 8, seed % 2), OracleAdapter(cap, 8, 8, seed % 2)
    ks = rng.choice(100_000, size=cap + 100, replace=False).astype(np.uint64)
    rows = ks.view(np.uint8).reshape(-1, 8)
    # 100 in, 30 out (tombstones), then exactly the 186 free slots' worth
    ops = [("ins", rows[:100], rows[:100]),
           ("rem", rows[:30]),
           ("ins", rows[100:100 + cap - 70], rows[100:100 + cap - 70]),
           ("get", rows)]
    apply_and_compare(m, o, ops)
    assert len(m) == cap


def test_gpu_crowded_homes():
    """Heavy collisions: 3000 keys whose homes crowd into 1024 of 8192 slots
    (one long cluster, long eviction chains, tombstones inside it)."""
    rng = np.random.default_rng(77)
    cap = 1 << 13
    say("crowded homes 8k")
    m, o = gpu_make(cap, 8, 8), OracleAdapter(cap, 8, 8)
    base = rng.integers(0, 1 << 40, size=3200, dtype=np.int64).astype(np.uint64)
    ks = np.unique((base << np.uint64(13)) | (base & np.uint64(0x3FF)))[:3000]
    rows = ks.view(np.uint8).reshape(-1, 8)
    ops = [("ins", rows[:1500], rows[:1500]), ("rem", rows[:500]), ("ins", rows[1500:], rows[:1500]),
           ("get", rows), ("rem", rows[1000:2000]), ("get", rows), ("ins", rows[:1200], rows[1200:2400])]
    apply_and_compare(m, o, ops)


def test_config1_full_bit_exact():
    """Config 1 end to end on the GPU vs the oracle and the golden digests."""
    from flashmap_amd.workloads import config1

    say("config1 full")
    c = config1()
    with open(os.path.join(GOLDEN_DIR, "config1_digest.json")) as f:
        g = json.load(f)
    m = gpu_make(c.capacity, 8, 8)
    o = OracleAdapter(c.capacity, 8, 8)
    assert m.bulk_insert(c.insert_keys, c.insert_vals) == g["inserted"] == o.bulk_insert(c.insert_keys, c.insert_vals)
    assert_same_slots(m, o, "config1 after insert")
    v1, f1 = m.bulk_get(c.queries)
    assert hashlib.sha256(v1.tobytes() + f1.astype(np.uint8).tobytes()).hexdigest() == g["get1_sha256"]
    assert m.bulk_remove(c.remove_keys) == g["removed"]
    v2, f2 = m.bulk_get(c.queries)
    assert int(f2.sum()) == g["hits2"]
    assert hashlib.sha256(v2.tobytes() + f2.astype(np.uint8).tobytes()).hexdigest() == g["get2_sha256"]
    assert hashlib.sha256(m.export_slots()[0].tobytes()).hexdigest() == g["flags_sha256"]


@pytest.mark.parametrize("group", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("block", [64, 256, 1024])
def test_launch_shapes_agree(group, block):
    from flashmap_amd.workloads import config1

    say(f"launch shape group={group} block={block}")
    c = config1(n_insert=50_000, n_query=50_000, n_remove=10_000, capacity=1 << 17)
    m = gpu_make(c.capacity, 8, 8, group=group, block=block)
    o = OracleAdapter(c.capacity, 8, 8)
    ops = [("ins", c.insert_keys, c.insert_vals), ("get", c.queries), ("rem", c.remove_keys), ("get", c.queries)]
    apply_and_compare(m, o, ops)


def test_device_resident_api():
    say("device-resident api")
    rng = np.random.default_rng(5)
    cap = 1 << 12
    m = gpu_make(cap, 32, 8)
    o = OracleAdapter(cap, 32, 8)
    keys = rng.integers(0, 256, (2000, 32), dtype=np.uint8)
    vals = rng.integers(0, 256, (2000, 8), dtype=np.uint8)
    dk, dv = m.upload(keys), m.upload(vals)
    assert m.bulk_insert_device(dk, dv, 1500) == o.bulk_insert(keys[:1500], vals[:1500])
    dout, dfound = m.alloc_device(2000 * 8), m.alloc_device(2000)
    m.bulk_get_device(dk, 2000, dout, dfound)
    ev, ef = o.bulk_get(keys)
    assert np.array_equal(m.download(dout).reshape(-1, 8), ev)
    assert np.array_equal(m.download(dfound).astype(bool), ef)
    dout2 = m.alloc_device(2000 * 8)
    m.bulk_get_values_device(dk, 2000, dout2)
    assert np.array_equal(m.download(dout2).reshape(-1, 8), ev)
    # uncounted insert: len += count (gpu.rs:670), recount restores the truth
    dk2, dv2 = m.upload(keys[1000:]), m.upload(vals[1000:])
    m.bulk_insert_device_uncounted(dk2, dv2, 1000)
    o.bulk_insert(keys[1000:], vals[1000:])
    assert len(m) == 2500
    assert m.recount() == len(o) == 2000
    assert m.bulk_remove_device(dk, 700) == o.bulk_remove(keys[:700])
    assert_same_slots(m, o, "device api")
    with pytest.raises(fm.FlashMapError):
        m.bulk_get_device(dk, 10, fm.DeviceBuffer(m, 0, 0), dfound)


def test_config2_reduced_bit_exact():
    """Config 2 shape ([u8;32]->[u8;128]) at 1/16 size vs the oracle."""
    from flashmap_amd.workloads import config2

    say("config2 reduced")
    c = config2(n_preload=500_000, n_batches=4, batch=6_250, n_get=62_500, n_remove=12_500, capacity=625_000)
    m, o = gpu_make(c.capacity, 32, 128), OracleAdapter(c.capacity, 32, 128)
    ops = [("ins", c.preload_keys, c.preload_vals)] + [("ins", k, v) for k, v in c.batches] + \
          [("get", c.get_keys), ("rem", c.remove_keys), ("get", c.get_keys)]
    apply_and_compare(m, o, ops)


def test_gpu_config2_shape_history():
    """The config 2 shape at reduced size: preload, sparse batches, removes,
    re-inserts of removed keys, gets.  The retired cooperative insert is
    refused by the tuning call."""
    from flashmap_amd.workloads import config2

    say("config2 shape history")
    c = config2(n_preload=200_000, n_batches=3, batch=5_000, n_get=20_000, n_remove=5_000, capacity=400_000)
    o = OracleAdapter(c.capacity, 32, 128)
    ops = [("ins", c.preload_keys, c.preload_vals)] + [("ins", k, v) for k, v in c.batches] + \
          [("rem", c.remove_keys), ("ins", c.remove_keys[:2000], c.batches[0][1][:2000]), ("get", c.get_keys)]
    m = gpu_make(c.capacity, 32, 128)
    with pytest.raises(fm.FlashMapError):
        m.set_tuning(insert_fused=1)
    apply_and_compare(m, o, ops)


@pytest.mark.parametrize("structure_only", [False, True], ids=["all", "structure"])
def test_gpu_copy_then_ops(structure_only):
    """copy_from leaves the destination's claim bitmap equal to its free
    bitmap, so the next inserts skip the copy; results must not change."""
    rng = np.random.default_rng(91)
    cap, K, V = 4096, 32, 8
    say(f"copy then ops structure_only={structure_only}")
    ops1 = random_history(rng, cap, K, V, 8, cap // 2, cap * 2)
    ops2 = random_history(rng, cap, K, V, 8, cap // 4, cap)
    src, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    apply_and_compare(src, o, ops1)
    dst = gpu_make(cap, K, V)
    if structure_only:  # values are not copied: dst first holds the same values
        apply_and_compare(dst, OracleAdapter(cap, K, V), ops1)
    dst.copy_from(src, structure_only=structure_only)
    apply_and_compare(dst, o, ops2)


def test_gpu_shift_past_duplicate_key():
    """A shift whose carried resident meets its own duplicate (left by the
    tombstone anomaly): the sequential insert updates the duplicate's value
    and claims nothing.  Identity hash, home = key mod 1024."""
    cap, h = 1024, 500
    say("shift past duplicate")

    def key(home, tag):
        return np.array([home + cap * tag], dtype=np.uint64).view(np.uint8).reshape(1, 8)

    W, X, Y, Z = key(h - 1, 3), key(h, 1), key(h, 2), key(h - 1, 4)
    v = lambda x: np.full((1, 8), x, dtype=np.uint8)  # noqa: E731
    ops = [("ins", np.concatenate([W, X, Y]), np.concatenate([v(1), v(2), v(3)])),
           ("rem", X),                      # tombstone at h
           ("ins", Y, v(4)),                # claims the tombstone: Y at h and at h+1
           ("ins", Z, v(5)),                # evicts Y at h; the carried Y meets its duplicate
           ("get", np.concatenate([W, X, Y, Z]))]
    m = gpu_make(cap, 8, 8)
    apply_and_compare(m, OracleAdapter(cap, 8, 8), ops)


def _hot_batch(rng, pool, n, K):
    """n keys drawn with heavy repetition from `pool` (rows), skewed."""
    ranks = np.minimum((rng.pareto(0.8, size=n) * 3).astype(np.int64), len(pool) - 1)
    return pool[ranks]


@pytest.mark.parametrize("seed", range(8))
def test_gpu_insert_dedup_hot_keys(seed):
    """insert_dedup = 1: batches with many copies of hot keys (new and
    existing, with tombstones in the table) fold later copies into the first
    one; counts, gets, removes and the slot array must equal the oracle's."""
    rng = np.random.default_rng(5000 + seed)
    cap = [1 << 12, 1 << 14, 1 << 16, 256][seed % 4]
    K, V = [(8, 8), (32, 8), (32, 128), (16, 16)][seed % 4]
    say(f"insert dedup seed={seed} cap={cap} K={K} V={V}")
    pool = rng.integers(0, 256, size=(max(8, cap // 16), K), dtype=np.uint8)
    m = gpu_make(cap, K, V, seed % 2)
    m.set_tuning(insert_dedup=1, insert_coop_arc=[64, 0][seed % 2])
    ops = []
    for r in range(8):
        n = int(rng.integers(1, cap // 3))
        ks = _hot_batch(rng, pool, n, K)
        ops.append(("ins", ks, rng.integers(0, 256, size=(n, V), dtype=np.uint8)))
        if r % 3 == 1:
            ops.append(("rem", _hot_batch(rng, pool, max(1, n // 4), K)))
        ops.append(("get", pool))
    apply_and_compare(m, OracleAdapter(cap, K, V, seed % 2), ops)


@pytest.mark.parametrize("seed", range(3))
def test_gpu_insert_dedup_fill_to_full(seed):
    """Deduplicated batches on the sort-grouped path (the batch could fill
    the table), including a batch that fills it exactly."""
    rng = np.random.default_rng(5100 + seed)
    cap = 512
    say(f"insert dedup fill seed={seed}")
    ks = rng.choice(1 << 30, size=cap + 200, replace=False).astype(np.uint64).view(np.uint8).reshape(-1, 8)
    m, o = gpu_make(cap, 8, 8, seed % 2), OracleAdapter(cap, 8, 8, seed % 2)
    m.set_tuning(insert_dedup=1)
    batch = np.concatenate([ks[:cap - 100], ks[:50], ks[10:60]])
    ops = [("ins", ks[:100], ks[:100]), ("rem", ks[:20]),
           ("ins", batch, rng.integers(0, 256, size=(len(batch), 8), dtype=np.uint8)),
           ("get", ks)]
    apply_and_compare(m, o, ops)


def test_gpu_serve_hot_insert_batch_matches_oracle():
    """fm_serve_log turns deduplication on: a served hot-key insert batch and
    gets after it agree with the oracle."""
    rng = np.random.default_rng(5200)
    cap, K, V = 1 << 16, 32, 8
    say("serve hot insert batch")
    pool = rng.integers(0, 256, size=(400, K), dtype=np.uint8)
    m, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    ins = _hot_batch(rng, pool, 20_000, K)
    vals = rng.integers(0, 256, size=(len(ins), V), dtype=np.uint8)
    from flashmap_amd.workloads import Request

    reqs = [Request("insert", 0.0, ins, vals), Request("get", 0.0, pool)]
    results, _ = fm.serve_log(m, reqs, staging_bytes=1 << 26, max_batch_keys=1 << 20, realtime=False)
    assert results[0].count == o.bulk_insert(ins, vals)
    ev, ef = o.bulk_get(pool)
    assert np.array_equal(results[1].values, ev) and np.array_equal(results[1].found.astype(bool), ef)
    assert_same_slots(m, o, "after served hot inserts")


def test_config3_reduced_bit_exact():
    """Config 3 shape ([u8;32] -> u64 row id, zipf probes) at 1/200 size vs the oracle."""
    from flashmap_amd.workloads import config3

    say("config3 reduced")
    c = config3(n_load=200_000, n_probe=500_000, capacity=250_000)
    rows = c.row_ids.view(np.uint8).reshape(-1, 8)
    m, o = gpu_make(c.capacity, 32, 8), OracleAdapter(c.capacity, 32, 8)
    apply_and_compare(m, o, [("ins", c.keys, rows), ("get", c.probes)])


def test_config3_full_row_ids():
    """Config 3 at full load: 40M keys -> row id in a 50M map, 20M zipf probes.

    The oracle cannot run this size in test time, so the check is the
    property the workload fixes: every hit returns exactly the row id its key
    was loaded with, every miss is not found, and the map holds 40M entries.
    """
    from flashmap_amd.workloads import config3

    say("config3 full: generating")
    c = config3(n_probe=20_000_000)
    n, m_probe = len(c.keys), len(c.probes)
    m = gpu_make(c.capacity, 32, 8)
    dk = m.upload(c.keys)
    dv = m.upload(c.row_ids.view(np.uint8))
    say("config3 full: insert")
    assert m.bulk_insert_device(dk, dv, n) == n
    assert len(m) == n
    del dk, dv
    dp = m.upload(c.probes)
    dout, dfound = m.alloc_device(m_probe * 8), m.alloc_device(m_probe)
    say("config3 full: get")
    m.bulk_get_device(dp, m_probe, dout, dfound)
    found = m.download(dfound).astype(bool)
    vals = m.download(dout).view("<u8")
    assert np.array_equal(found, c.probe_row >= 0)
    hit = c.probe_row >= 0
    assert np.array_equal(vals[hit], c.probe_row[hit].astype(np.uint64))
    assert not vals[~hit].any()
    say("config3 full: ok")


def test_gpu_autotune_get():
    """The get autotuner leaves the table alone and picks a valid shape;
    gets with it still match the oracle."""
    say("autotune get")
    rng = np.random.default_rng(21)
    cap = 1 << 16
    m, o = gpu_make(cap, 32, 8), OracleAdapter(cap, 32, 8)
    keys = rng.integers(0, 256, (50_000, 32), dtype=np.uint8)
    vals = rng.integers(0, 256, (50_000, 8), dtype=np.uint8)
    assert m.bulk_insert(keys, vals) == o.bulk_insert(keys, vals)
    q = np.concatenate([keys[::3], rng.integers(0, 256, (20_000, 32), dtype=np.uint8)])
    dq = m.upload(q)
    rec = m.autotune_get(dq, len(q), reps=2)
    t = m.tuning()
    assert (t.get_group, t.get_block) == (rec["get_group"], rec["get_block"])
    assert rec["get_group"] in (1, 2, 4, 8, 16) and rec["get_block"] in (64, 128, 256, 512, 1024)
    apply_and_compare(m, o, [("get", q)])


def test_gpu_claim_bitmap_reuse_sequence():
    """The claim bitmap is reused between inserts when the last insert left it
    equal to the free-slot bitmap.  This history crosses every event that
    must keep it equal or force a fresh copy: consecutive inserts, small
    removes (kept in step), large removes (copy next time), duplicate-heavy
    and fill-to-near-full batches, an uncounted insert, clear, slot import
    and table copies, each followed by inserts checked against the oracle."""
    say("claim bitmap reuse")
    rng = np.random.default_rng(77)
    cap = 1 << 17
    m, o = gpu_make(cap, 32, 8), OracleAdapter(cap, 32, 8)
    pool = rng.integers(0, 256, (60_000, 32), dtype=np.uint8)

    def batch(n, dup=False):
        idx = rng.integers(0, len(pool) if not dup else 500, n)
        return pool[idx], rng.integers(0, 256, (n, 8), dtype=np.uint8)

    ops = []
    for _ in range(3):
        ops.append(("ins",) + batch(20_000))
    ops.append(("rem", pool[rng.integers(0, len(pool), 5_000)]))     # small remove: Q kept
    ops.append(("ins",) + batch(20_000, dup=True))
    ops.append(("rem", pool[rng.integers(0, len(pool), 40_000)]))    # large remove: copy next insert
    ops.append(("ins",) + batch(30_000))
    ops.append(("get", pool))
    apply_and_compare(m, o, ops)
    # uncounted insert (no status check), then counted inserts
    k, v = batch(10_000)
    dk, dv = m.upload(k), m.upload(v)
    m.bulk_insert_device_uncounted(dk, dv, len(k))
    o.bulk_insert(k, v)
    m.recount()
    apply_and_compare(m, o, [("ins",) + batch(20_000), ("ins",) + batch(20_000)])
    # a batch that fills the table (sort path), then led inserts again
    apply_and_compare(m, o, [("rem", pool[:30_000])])
    room = cap - len(o)
    fresh = rng.integers(0, 256, (room, 32), dtype=np.uint8)
    apply_and_compare(m, o, [("ins", fresh, np.zeros((room, 8), np.uint8)), ("rem", fresh[:50_000]),
                             ("ins",) + batch(20_000), ("ins",) + batch(20_000)])
    # clear, then slot import and a full table copy
    apply_and_compare(m, o, [("clear",), ("ins",) + batch(20_000), ("ins",) + batch(20_000)])
    flags, keys, vals = o.export_slots()
    m2 = gpu_make(cap, 32, 8)
    m2.import_slots(flags, keys, vals, len(o))
    apply_and_compare(m2, o, [("ins",) + batch(20_000)])
    m.copy_from(m2)
    apply_and_compare(m, o, [("ins",) + batch(20_000), ("rem", pool[rng.integers(0, len(pool), 3_000)]),
                             ("ins",) + batch(20_000)])


@pytest.mark.parametrize("K,V", [(8, 8), (32, 128)])
def test_gpu_long_arcs_in_lds_and_hbm(K, V):
    """Near-full tables leave few barriers, so a batch forms closures whose
    arcs span thousands of slots: the ones that fit the LDS buffer (up to
    7,168 slots at K=8, 2,867 at K=32) replay there, longer ones through HBM,
    and arcs that wrap past the end of the table are among them."""
    say(f"long arcs K={K} V={V}")
    rng = np.random.default_rng(91 + K)
    cap = 1 << 14
    m, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    keys = rng.integers(0, 256, (cap, K), dtype=np.uint8)
    vals = rng.integers(0, 256, (cap, V), dtype=np.uint8)
    ops = [("ins", keys[:cap - 900], vals[:cap - 900]),            # ~94% full
           ("ins", keys[cap - 900:cap - 60], vals[cap - 900:cap - 60]),  # most barriers consumed
           ("rem", keys[:2000:7]),                                  # tombstones inside the runs
           ("ins", keys[:1500:5], vals[4000:5500:5]),               # updates and reinserts past tombstones
           ("ins", np.concatenate([keys[cap - 60:], keys[10:40]]), np.concatenate([vals[:60], vals[:30]])),
           ("ins", keys[:4000:5], vals[4000:8000:5]),               # len + batch > capacity: TableFull
           ("get", keys)]
    apply_and_compare(m, o, ops)


@pytest.mark.parametrize("canon_min", ["1", "0"])
@pytest.mark.parametrize("K,V", [(8, 8), (32, 128)])
def test_gpu_canonical_arc_replay(K, V, canon_min, monkeypatch):
    """Arcs replayed from their groups of equal homes (canon_min 1: every
    canonical arc k_ins_big takes; 0: none, the walk replays) against the
    oracle.  Identity-hashed keys put many keys on few homes, so the arcs are
    long and their groups large: residents updated by the batch, keys
    repeated inside a batch (the last value stays), and groups rotated by
    the inserts before them.  The removes leave tombstones inside runs, and
    the arcs that hold them are not canonical (walk replay either way)."""
    say(f"canonical arc replay K={K} V={V} canon_min={canon_min}")
    monkeypatch.setenv("FM_INSERT_CANON_MIN", canon_min)  # read when a map is built
    rng = np.random.default_rng(4242 + K)
    cap = 1 << 14

    def keys_on(homes):
        ids = rng.permutation(1 << 20)[:len(homes)].astype(np.uint64) + np.uint64(1)
        first = np.asarray(homes, dtype=np.uint64) | (ids << np.uint64(14))
        k = rng.integers(0, 256, (len(homes), K), dtype=np.uint8)
        k[:, :8] = first.view(np.uint8).reshape(-1, 8)[:, :min(8, K)] if K >= 8 else k[:, :8]
        return k

    def vals(n):
        return rng.integers(0, 256, (n, V), dtype=np.uint8)

    a = keys_on(rng.integers(0, 12000, 9000))
    b_new = keys_on(rng.integers(5000, 6000, 1500))
    b = np.concatenate([b_new, a[rng.integers(0, 9000, 1500)], b_new[rng.integers(0, 1500, 1000)]])
    b = b[rng.permutation(len(b))]
    c_new = keys_on(rng.integers(0, cap, 1200))
    c = np.concatenate([c_new, a[:1500:2], a[rng.integers(0, 9000, 600)], c_new[:300]])
    c = c[rng.permutation(len(c))]
    m, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    ops = [("ins", a, vals(len(a))),
           ("ins", b, vals(len(b))),
           ("ins", keys_on(rng.integers(5500, 5600, 40)), vals(40)),   # sparse: leader-queued closures
           ("rem", a[:3000:2]),
           ("ins", c, vals(len(c))),
           ("get", np.concatenate([a, b, c]))]
    apply_and_compare(m, o, ops)


@pytest.mark.parametrize("K,V", [(8, 8), (32, 128)])
def test_gpu_sparse_updates_after_removes(K, V):
    """Sparse insert batches (n * 16 < capacity: planned one-member inserts,
    packed replays) that are half updates, alternating with removes on a
    crowded table.  Removes leave runs that are no longer sorted by home, so
    an update's walk can evict residents before it reaches its equal key:
    the plan then ends at that key with evictions applied before it, each
    evicted slot with its own distance.  The whole slot array is compared
    with the oracle after every op."""
    say(f"sparse updates after removes K={K} V={V}")
    rng = np.random.default_rng(1234 + K)
    cap = 1 << 14
    m, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    keys = rng.integers(0, 256, (cap, K), dtype=np.uint8)
    live = 12_000
    ops = [("ins", keys[:live], rng.integers(0, 256, (live, V), dtype=np.uint8))]
    nxt = live
    for _ in range(12):
        ops.append(("rem", keys[rng.integers(0, nxt, 400)]))
        upd = keys[rng.integers(0, nxt, 250)]
        new = keys[nxt:nxt + 250]
        nxt += 250
        batch = np.concatenate([upd, new])[rng.permutation(500)]
        ops.append(("ins", batch, rng.integers(0, 256, (500, V), dtype=np.uint8)))
    ops.append(("get", keys[:nxt]))
    apply_and_compare(m, o, ops)


@pytest.mark.parametrize("pin", ["pageable", "pinned_keys", "pinned_results"])
def test_gpu_host_calls_in_staged_pieces(pin):
    """Host-buffer calls of several MiB move through the pinned staging
    buffers in pieces of at least 4 MiB that the host copy and the DMA engine
    work on in turn.  Sizes are odd, so each transfer's last piece is short;
    the second insert updates earlier keys, the gets hold misses (zeroed
    values), and one side of a call may be pinned (fm_host_alloc) while the
    other is staged.  Counts, values, found flags and the slot array against
    the oracle."""
    say(f"staged host calls {pin}")
    rng = np.random.default_rng(77)
    cap, K, V, n = 1 << 19, 32, 128, 400_009
    m, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    keys = rng.integers(0, 256, (n, K), dtype=np.uint8)
    vals = rng.integers(0, 256, (n, V), dtype=np.uint8)
    arena = fm.PinnedArena(3 * n * (K + V) + n + (1 << 16)) if pin != "pageable" else None
    host = (lambda a: arena.copy_of(a)) if pin == "pinned_keys" else (lambda a: a)

    def get_and_compare(ks, where):
        outs = {}
        if pin == "pinned_results":
            outs = {"out": arena.take((len(ks), V), fill=7), "found": arena.take((len(ks),), fill=1).view(np.bool_)}
        gv, gf = m.bulk_get(host(ks), **outs)
        ev, ef = o.bulk_get(ks)
        assert np.array_equal(gf, ef), f"{where}: found flags differ"
        assert np.array_equal(gv, ev), f"{where}: values differ"

    assert m.bulk_insert(host(keys[:300_007]), host(vals[:300_007])) == o.bulk_insert(keys[:300_007], vals[:300_007])
    batch = np.concatenate([keys[100_000:150_001], keys[300_007:350_009]])
    bvals = rng.integers(0, 256, (len(batch), V), dtype=np.uint8)
    assert m.bulk_insert(host(batch), host(bvals)) == o.bulk_insert(batch, bvals)
    get_and_compare(keys, "after inserts")
    gone = keys[::3]
    assert m.bulk_remove(host(gone)) == o.bulk_remove(gone)
    get_and_compare(keys, "after remove")
    assert len(m) == len(o)
    assert_same_slots(m, o, "slots")
    if arena is not None:
        arena.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_gpu_bitmap_rebuild_and_per_key_batches_alternate(seed):
    """Insert batches alternate across the size where the bits pass switches
    from per-key atomics to rebuilding every bitmap word (n * 16 >= cap),
    with removes (tombstones) between them: each path must leave the claim
    bitmap the other expects.  The whole slot array is compared after every
    op."""
    say(f"bits rebuild alternation seed={seed}")
    rng = np.random.default_rng(300 + seed)
    cap, K, V = 1 << 12, 8, 8
    m, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    universe = rng.integers(0, 1 << 62, 6000).astype(np.uint64).view(np.uint8).reshape(-1, 8)
    ops = []
    for step in range(16):
        n = int(rng.choice([40, 200, 255, 256, 300, 700]))  # 256 = cap / 16
        idx = rng.integers(0, len(universe), n)
        ops.append(("ins", universe[idx], rng.integers(0, 256, (n, V), dtype=np.uint8)))
        if step % 2:
            ops.append(("rem", universe[rng.integers(0, len(universe), int(rng.integers(20, 400)))]))
        if step % 5 == 4:
            ops.append(("get", universe[rng.integers(0, len(universe), 500)]))
    apply_and_compare(m, o, ops)


@pytest.mark.parametrize("K,V", [(256, 256), (1, 1), (255, 3), (4, 256)])
def test_gpu_size_limits(K, V):
    """Key and value sizes at the GPU limits (1 and 256 bytes; 255 and 3
    take the byte-wise generic kernels) against the oracle, murmur3 and
    identity hashing."""
    for mode in (0, 1):
        say(f"size limits K={K} V={V} mode={mode}")
        rng = np.random.default_rng(K * 1000 + V + mode)
        cap = 256
        ops = random_history(rng, cap, K, V, 16, cap // 2, cap * 2 if K > 1 else 256)
        apply_and_compare(gpu_make(cap, K, V, mode), OracleAdapter(cap, K, V, mode), ops)


def test_gpu_builder_size_above_limit_falls_back():
    """A 257-byte key is beyond the GPU kernels: the builder falls back to
    the CPU arm (lib.rs:414-449), and require_gpu reports the GPU as
    unavailable instead."""
    say("builder above the size limit")
    m = fm.FlashMap.builder(64, 257, 8).quiet().build()
    assert m.backend() == fm.Backend.CPU
    k = np.arange(257 * 3, dtype=np.uint64).astype(np.uint8).reshape(3, 257)
    assert m.bulk_insert(k, np.zeros((3, 8), np.uint8)) == 3
    with pytest.raises(fm.FlashMapError):
        fm.FlashMap.builder(64, 257, 8).require_gpu().build()


def test_gpu_max_capacity_wraps_at_2_31():
    """The largest GPU table (2^31 slots; 16 GiB each of keys and values,
    so every slot offset needs 64 bits).  Identity-hashed keys crowded at the
    last homes of the table run past slot 2^31 - 1 and wrap to slot 0; gets,
    updates and removes there must match the oracle.  One slot more is
    refused when the GPU is required."""
    say("max capacity 2^31")
    cap = 1 << 31
    m = gpu_make(cap, 8, 8)
    assert m.capacity() == cap
    o = OracleAdapter(cap, 8, 8)
    rng = np.random.default_rng(31)
    homes = (cap - 1 - rng.integers(0, 40, 300)).astype(np.uint64)
    lows = rng.integers(0, 20, 100).astype(np.uint64)
    ids = np.concatenate([homes, lows]) + (np.arange(400, dtype=np.uint64) << np.uint64(31))  # same homes, distinct keys
    keys = ids[rng.permutation(ids.size)].view(np.uint8).reshape(-1, 8)
    vals = rng.integers(0, 256, (keys.shape[0], 8), dtype=np.uint8)
    ops = [("ins", keys, vals), ("get", keys), ("ins", keys[::3], vals[::-3][: len(keys[::3])]),
           ("rem", keys[::5]), ("get", keys), ("ins", keys[::5], vals[::5]), ("get", keys)]
    apply_and_compare(m, o, ops, check_slots=False)
    m.close()
    with pytest.raises(fm.FlashMapError):
        fm.FlashMap.builder(cap + 1, 8, 8).require_gpu().build()


@pytest.mark.parametrize("next_bits", [True, False])
@pytest.mark.parametrize("K,V", [(8, 8), (32, 128)])
def test_gpu_sparse_batches_that_queue_big_closures(K, V, next_bits, monkeypatch):
    """Counted inserts of sparse batches (n * 16 < cap) launch no k_ins_big up
    front: when the leader queues closures, its hand-over reports them and
    the host runs k_ins_big.  The claims either go to the next bitmaps as
    the replay makes them (no settle pass at all: the leader or k_ins_big
    hands the count over) or, with FM_INSERT_NEXT_BITS=0, are published by
    the settle pass k_ins_bits (once, or again after k_ins_big).  In a near-full
    table such batches queue closures (batches of updates, with repeated
    keys, sit between them); in a lightly loaded one a sparse batch queues
    none.  The whole slot array is compared with the oracle after every op,
    and the launch counts show which branch ran."""
    say(f"sparse batches with queued closures K={K} V={V} next_bits={next_bits}")
    monkeypatch.setenv("FM_INSERT_NEXT_BITS", "1" if next_bits else "0")  # read when a map is built
    rng = np.random.default_rng(700 + K)
    cap = 1 << 13
    m, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    keys = rng.integers(0, 256, (cap, K), dtype=np.uint8)
    vals = rng.integers(0, 256, (cap, V), dtype=np.uint8)
    lo = cap - 900
    ops = [("ins", keys[:lo], vals[:lo])]  # dense: k_ins_big launched directly
    for n in [8, 300, 1, 200, 50, 40]:  # every n * 16 < cap
        ops.append(("ins", keys[lo:lo + n], vals[lo:lo + n]))
        lo += n
        ops.append(("ins", keys[rng.integers(0, lo, n)], vals[rng.integers(0, cap, n)]))
    ops.append(("rem", keys[:3000:3]))
    ops.append(("ins", keys[:400:2], vals[:200]))
    ops.append(("get", keys))
    apply_and_compare(m, o, ops[:1])
    m.profile(True)  # the sparse inserts only
    apply_and_compare(m, o, ops[1:])
    prof = m.profile_read()
    sparse = sum(1 for op in ops[1:] if op[0] == "ins")
    settles = prof.get("k_ins_bits", (0, 0))[0]
    bigs = prof.get("k_ins_big", (0, 0))[0]
    assert bigs >= 1, prof  # some batch queued closures: deferred, then replayed
    if next_bits:
        assert settles == 0, (bigs, settles)  # k_ins_big hands the count over after each replay
    else:
        assert sparse < settles <= 2 * sparse, (sparse, settles)  # ... and settled a second time
    # the other branch: a sparse batch into a lightly loaded table queues
    # nothing and settles in one pass (or none)
    m2, o2 = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    apply_and_compare(m2, o2, [("ins", keys[:2000], vals[:2000])])
    m2.profile(True)
    apply_and_compare(m2, o2, [("ins", keys[2000:2100], vals[2000:2100])])
    prof2 = m2.profile_read()
    assert prof2.get("k_ins_bits", (0, 0))[0] == (0 if next_bits else 1), prof2
    assert prof2.get("k_ins_big", (0, 0))[0] == 0, prof2


def test_profile_only_times_one_kernel():
    """fm_profile_only limits the timing events to one kernel's launches (the
    benchmark's timed steps), "" times every kernel again, and an unknown
    name is rejected; results are unaffected."""
    say("profile_only")
    rng = np.random.default_rng(77)
    cap, K, V = 1 << 14, 32, 128
    m, o = gpu_make(cap, K, V), OracleAdapter(cap, K, V)
    keys = rng.integers(0, 256, (3000, K), dtype=np.uint8)
    vals = rng.integers(0, 256, (3000, V), dtype=np.uint8)
    apply_and_compare(m, o, [("ins", keys[:2000], vals[:2000])])
    m.profile_only("k_ins_leader")
    m.profile(True)
    apply_and_compare(m, o, [("ins", keys[2000:2600], vals[2000:2600]), ("get", keys), ("rem", keys[:300])])
    prof = m.profile_read()
    assert set(prof) == {"k_ins_leader"} and prof["k_ins_leader"][0] == 1, prof
    m.profile_only("")
    m.profile(True)
    apply_and_compare(m, o, [("ins", keys[2600:], vals[2600:]), ("get", keys)])
    prof = m.profile_read()
    assert {"k_ins_dry", "k_ins_leader", "k_get"} <= set(prof), prof
    with pytest.raises(fm.FlashMapError):
        m.profile_only("no_such_kernel")
    m.profile(False)

