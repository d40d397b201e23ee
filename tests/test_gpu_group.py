"""Table groups (fused multi-table launches, hipGraph step) vs the oracle.

Each table of a group must behave exactly like its own FlashMap applying the
same batches: insert counts and TableFull statuses, every get result, remove
counts and the whole logical slot array.  Shapes follow config 4
(u64 -> [u8;32]) at reduced sizes, plus hand-made tables that hit TableFull
and tables whose batch fills them exactly (serial replay).
"""
import sys

import numpy as np
import pytest

import flashmap_amd as fm
from flashmap_amd.workloads import config4_table
from helpers import OracleAdapter

pytestmark = pytest.mark.gpu
V = 32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _concat(parts, width):
    if not parts:
        return np.zeros((0, width), dtype=np.uint8)
    return np.concatenate([np.ascontiguousarray(p).reshape(-1, width) for p in parts])


def _rows64(a):
    return np.ascontiguousarray(a, dtype=np.uint64).view(np.uint8).reshape(-1, 8)


def _tables(n, seed, cap_hi, ins_hi, get_hi, rem_hi):
    tabs = [config4_table(seed * 100 + t, seed=seed, cap_lo=16, cap_hi=cap_hi, ins_hi=ins_hi, get_hi=get_hi,
                          rem_hi=rem_hi) for t in range(n)]
    # a table whose insert batch overflows it (TableFull) and one whose batch
    # fills it exactly (no barrier: serial replay)
    full = config4_table(seed * 100 + n, seed=seed, cap_lo=64, cap_hi=64, ins_hi=2, get_hi=8, rem_hi=2)
    full.ins_keys = np.arange(10_000, 10_000 + full.capacity, dtype=np.uint64)
    full.ins_vals = np.zeros((full.capacity, V), dtype=np.uint8)
    fill = config4_table(seed * 100 + n + 1, seed=seed, cap_lo=128, cap_hi=128, ins_hi=2, get_hi=8, rem_hi=2)
    room = fill.capacity - len(fill.preload_keys)
    fill.ins_keys = np.arange(20_000, 20_000 + room, dtype=np.uint64)
    fill.ins_vals = np.full((room, V), 7, dtype=np.uint8)
    return tabs + [full, fill]


def _run_and_check(tabs, use_step, tuning=None, steps=1, autotune=None):
    n = len(tabs)
    g = fm.FlashMapGroup([t.capacity for t in tabs], 8, V)
    if tuning:
        g.set_tuning(**tuning)
    oracles = [OracleAdapter(t.capacity, 8, V) for t in tabs]
    # preload through the group insert path
    pk = g.upload(_concat([_rows64(t.preload_keys) for t in tabs], 8))
    pv = g.upload(_concat([t.preload_vals for t in tabs], V))
    new, status = g.bulk_insert_device(pk, pv, [len(t.preload_keys) for t in tabs])
    for i, t in enumerate(tabs):
        exp = oracles[i].bulk_insert(_rows64(t.preload_keys), t.preload_vals)
        assert (int(new[i]), int(status[i])) == (exp, 0), f"preload table {i}"
    ic = [len(t.ins_keys) for t in tabs]
    gc = [len(t.get_keys) for t in tabs]
    rc = [len(t.rem_keys) for t in tabs]
    ik = g.upload(_concat([_rows64(t.ins_keys) for t in tabs], 8))
    iv = g.upload(_concat([t.ins_vals for t in tabs], V))
    gk = g.upload(_concat([_rows64(t.get_keys) for t in tabs], 8))
    rk = g.upload(_concat([_rows64(t.rem_keys) for t in tabs], 8))
    out = g.alloc_device(max(1, sum(gc)) * V)
    found = g.alloc_device(max(1, sum(gc)))
    if autotune is not None:
        # the search steps a scratch copy: the steps below must still see the
        # preloaded state, and run with the tuning it picked
        rec = g.autotune(ik, iv, ic, gk, gc, out, found, rk, rc, **autotune)
        say(f"autotune: {rec}")
        assert {k: getattr(g.tuning(), k) for k in rec["tuning"]} == rec["tuning"]
        assert rec["evaluated"] >= 1 or rec["cached"]
    for s in range(steps):
        if use_step:
            new, status, removed = g.step_device(ik, iv, ic, gk, gc, out, found, rk, rc)
        else:
            new, status = g.bulk_insert_device(ik, iv, ic)
            g.bulk_get_device(gk, gc, out, found)
            removed = g.bulk_remove_device(rk, rc)
        vals = g.download(out, sum(gc) * V).reshape(-1, V)
        fnd = g.download(found, sum(gc)).astype(bool)
        off = 0
        for i, t in enumerate(tabs):
            o = oracles[i]
            try:
                exp_new, exp_st = o.bulk_insert(_rows64(t.ins_keys), t.ins_vals), 0
            except fm.TableFull:
                exp_new, exp_st = 0, fm.TABLE_FULL
            assert (int(new[i]), int(status[i])) == (exp_new, exp_st), f"step {s} table {i} insert"
            ev, ef = o.bulk_get(_rows64(t.get_keys))
            assert np.array_equal(fnd[off:off + gc[i]], ef), f"step {s} table {i} found"
            assert np.array_equal(vals[off:off + gc[i]], ev), f"step {s} table {i} values"
            off += gc[i]
            assert int(removed[i]) == o.bulk_remove(_rows64(t.rem_keys)), f"step {s} table {i} remove"
            assert g.len(i) == len(o), f"step {s} table {i} len"
        for i in range(n):
            fa, ka, va = g.export_slots(i)
            fb, kb, vb = oracles[i].export_slots()
            assert np.array_equal(fa, fb) and np.array_equal(ka, kb) and np.array_equal(va, vb), f"slots table {i}"


@pytest.mark.parametrize("use_step", [False, True], ids=["ops", "graph"])
def test_group_small_tables(use_step):
    say(f"group small tables step={use_step}")
    _run_and_check(_tables(24, 11, 1 << 12, 1 << 10, 1 << 10, 1 << 8), use_step)


def test_group_graph_replays_many_steps():
    """The captured step replayed 4 times; every replay checked (mutating
    batches: later steps see overwrites, tombstones and TableFull)."""
    say("group graph replays")
    _run_and_check(_tables(16, 12, 1 << 11, 1 << 9, 1 << 9, 1 << 7), True, steps=4)


@pytest.mark.parametrize("tuning", [dict(group=1, block=64, grid=64, coop_arc=0), dict(group=8, block=512, grid=2048),
                                    dict(group=16, block=128, grid=300)], ids=["g1b64", "g8b512", "g16b128"])
def test_group_tunings_agree(tuning):
    say(f"group tuning {tuning}")
    _run_and_check(_tables(12, 13, 1 << 13, 1 << 12, 1 << 11, 1 << 9), True, tuning=tuning)


def test_group_large_closures():
    """High-load tables with large batches: closures beyond the leader lane's
    capacity go through the workgroup path."""
    say("group large closures")
    tabs = _tables(6, 14, 1 << 15, 1 << 15, 1 << 10, 1 << 8)
    _run_and_check(tabs, True)


def test_group_autotune_search():
    """Coordinate-descent search on a scratch group, then two checked steps
    with the chosen shape."""
    say("group autotune search")
    _run_and_check(_tables(12, 15, 1 << 13, 1 << 12, 1 << 11, 1 << 9), True, steps=2, autotune=dict(reps=1))


def test_group_autotune_candidates_and_cache(tmp_path):
    """An explicit candidate list with a caller-provided scratch group; the
    result is cached and a second group of the same shape takes it from the
    cache without measuring."""
    say("group autotune cache")
    tabs = _tables(8, 16, 1 << 12, 1 << 10, 1 << 10, 1 << 8)
    cache = str(tmp_path / "tune.json")
    cand = [dict(group=2, block=128, grid=512, coop_arc=16), dict(group=8, block=256, grid=1024, coop_arc=64),
            dict(group=4, block=64, grid=256, coop_arc=0)]
    caps = [t.capacity for t in tabs]
    scratch = fm.FlashMapGroup(caps, 8, V)
    _run_and_check(tabs, True, autotune=dict(reps=2, candidates=cand, scratch=scratch, cache=cache))
    import json
    with open(cache) as f:
        known = json.load(f)
    assert len(known) == 1
    rec = next(iter(known.values()))
    assert rec["evaluated"] == 3 and rec["tuning"] in cand
    _run_and_check(tabs, True, autotune=dict(reps=2, candidates=cand, cache=cache))
