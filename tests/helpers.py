"""Shared test helpers: oracle adapter, restated reference tests, differential runner."""
from __future__ import annotations

import numpy as np

import flashmap_amd as fm
from oracle.binding import OracleMap, TableFull as OracleTableFull


class OracleAdapter:
    """Gives the C oracle the same surface as flashmap_amd.FlashMap."""

    def __init__(self, capacity, key_size=8, value_size=8, mode=0):
        self.o = OracleMap(capacity, key_size, value_size, mode)
        self.key_size, self.value_size = key_size, value_size

    def bulk_get(self, keys):
        k = np.ascontiguousarray(keys).reshape(-1).view(np.uint8)
        vals, found = self.o.bulk_get(k)
        return vals.reshape(-1, self.value_size), found.astype(bool)

    def bulk_insert(self, keys, values):
        try:
            return self.o.bulk_insert(np.ascontiguousarray(keys).reshape(-1).view(np.uint8),
                                      np.ascontiguousarray(values).reshape(-1).view(np.uint8))
        except OracleTableFull:
            n, c = len(self.o), self.o.capacity
            raise fm.TableFull(fm.TABLE_FULL, "table full", n, c, n / c * 100.0)

    def bulk_remove(self, keys):
        return self.o.bulk_remove(np.ascontiguousarray(keys).reshape(-1).view(np.uint8))

    def __len__(self):
        return len(self.o)

    def len(self):
        return len(self.o)

    def is_empty(self):
        return len(self.o) == 0

    def capacity(self):
        return self.o.capacity

    def load_factor(self):
        return len(self.o) / self.o.capacity

    def clear(self):
        self.o.clear()

    def export_slots(self):
        return self.o.export()

    def __repr__(self):
        n, c = len(self.o), self.o.capacity
        return f'FlashMap {{ backend: "oracle", len: {n}, capacity: {c}, load_factor: "{n / c * 100:.1f}%" }}'


def le(xs, width=8):
    """Little-endian byte rows of integers (i.to_le_bytes() in the reference tests)."""
    x = np.array(list(xs), dtype=np.uint64).view(np.uint8).reshape(-1, 8)
    a = np.zeros((x.shape[0], width), dtype=np.uint8)
    w = min(8, width)
    a[:, :w] = x[:, :w]
    return a


def u64s(vals):
    return np.asarray(vals).view(np.uint64).reshape(-1)


# ---------------------------------------------------------------------------
# tests/correctness.rs restated.  `make(capacity, K, V, mode)` builds a map.
# ---------------------------------------------------------------------------
def t_insert_and_get(make):  # correctness.rs:4-22
    m = make(1024, 8, 8, 0)
    assert m.bulk_insert(le(range(100)), le([i * 100 for i in range(100)])) == 100
    assert len(m) == 100
    v, f = m.bulk_get(le(range(100)))
    assert f.all() and (u64s(v) == np.arange(100, dtype=np.uint64) * 100).all()


def t_get_missing_keys(make):  # :25-39
    m = make(1024, 8, 8, 0)
    m.bulk_insert(le(range(10)), le(range(10)))
    _, f = m.bulk_get(le(range(100, 110)))
    assert not f.any()


def t_update_existing_keys(make):  # :42-65
    m = make(1024, 8, 8, 0)
    assert m.bulk_insert(le(range(10)), le(range(10))) == 10
    assert m.bulk_insert(le(range(10)), le([i + 1000 for i in range(10)])) == 0
    assert len(m) == 10
    v, f = m.bulk_get(le(range(10)))
    assert f.all() and (u64s(v) == np.arange(10, dtype=np.uint64) + 1000).all()


def t_remove_keys(make):  # :68-95
    m = make(1024, 8, 8, 0)
    m.bulk_insert(le(range(100)), le(range(100)))
    assert len(m) == 100
    assert m.bulk_remove(le(range(50))) == 50
    assert len(m) == 50
    assert not m.bulk_get(le(range(50)))[1].any()
    assert m.bulk_get(le(range(50, 100)))[1].all()


def t_insert_after_remove_reuses_tombstones(make):  # :98-124
    m = make(64, 8, 8, 0)
    m.bulk_insert(le(range(10)), le(range(10)))
    m.bulk_remove(le(range(10)))
    assert len(m) == 0
    assert m.bulk_insert(le(range(100, 110)), le([i * 2 for i in range(100, 110)])) == 10
    v, f = m.bulk_get(le(range(100, 110)))
    assert f.all() and (u64s(v) == np.arange(100, 110, dtype=np.uint64) * 2).all()


def t_clear_resets_map(make):  # :127-145
    m = make(1024, 8, 8, 0)
    m.bulk_insert(le(range(100)), le(range(100)))
    assert len(m) == 100
    m.clear()
    assert len(m) == 0 and m.is_empty()
    assert not m.bulk_get(le(range(100)))[1].any()


def t_murmur3_hash_strategy(make):  # :148-164
    m = make(1024, 8, 8, 1)
    m.bulk_insert(le(range(100)), le([i * 7 for i in range(100)]))
    v, f = m.bulk_get(le(range(100)))
    assert f.all() and (u64s(v) == np.arange(100, dtype=np.uint64) * 7).all()


def t_large_keys_and_values(make):  # :167-191
    m = make(4096, 32, 64, 0)
    keys = le(range(500), 32)
    vals = le([i * 42 for i in range(500)], 64)
    m.bulk_insert(keys, vals)
    assert len(m) == 500
    v, f = m.bulk_get(keys)
    assert f.all() and (v == vals).all()


def t_load_factor_and_capacity(make):  # :194-199
    m = make(1024, 8, 8, 0)
    assert m.load_factor() == 0.0
    assert m.capacity() >= 1024 and (m.capacity() & (m.capacity() - 1)) == 0


def t_empty_operations_are_no_ops(make):  # :202-211
    m = make(64, 8, 8, 0)
    e = np.zeros((0, 8), dtype=np.uint8)
    assert m.bulk_insert(e, e) == 0
    v, f = m.bulk_get(e)
    assert f.size == 0
    assert m.bulk_remove(e) == 0


def t_table_full_error(make):  # :214-227
    import pytest

    m = make(16, 8, 8, 0)
    m.bulk_insert(le(range(16)), le(range(16)))
    with pytest.raises(fm.TableFull) as ei:
        m.bulk_insert(le([99]), le([99]))
    assert ei.value.occupied == 16 and ei.value.capacity == 16 and ei.value.load_factor == 100.0


def t_debug_formatting(make):  # :230-236
    s = repr(make(256, 8, 8, 0))
    assert "FlashMap" in s and "len" in s and "capacity" in s


def t_u64_key_value_types(make):  # :239-250
    m = make(1024, 8, 8, 0)
    ks = np.arange(50, dtype=np.uint64)
    m.bulk_insert(ks, ks * ks)
    v, f = m.bulk_get(ks)
    assert f.all() and (u64s(v) == ks * ks).all()


def t_robin_hood_high_load_correctness(make):  # :257-285
    for pct in (80, 90, 95):
        n = 1024 * pct // 100
        m = make(1024, 8, 8, 0)
        assert m.bulk_insert(le(range(n)), le([i * 7 for i in range(n)])) == n
        assert len(m) == n
        v, f = m.bulk_get(le(range(n)))
        assert f.all() and (u64s(v) == np.arange(n, dtype=np.uint64) * 7).all()


def t_robin_hood_miss_exits_early(make):  # :288-309
    n = 1024 * 90 // 100
    m = make(1024, 8, 8, 0)
    m.bulk_insert(le(range(n)), le(range(n)))
    assert not m.bulk_get(le(range(10_000, 10_500)))[1].any()


def t_robin_hood_insert_remove_reinsert(make):  # :312-362
    n = 512 * 80 // 100
    m = make(512, 8, 8, 0)
    m.bulk_insert(le(range(n)), le(range(n)))
    assert m.bulk_remove(le(range(n // 2))) == n // 2
    assert m.bulk_insert(le(range(10_000, 10_000 + n // 2)),
                         le([i * 3 for i in range(10_000, 10_000 + n // 2)])) == n // 2
    v, f = m.bulk_get(le(range(n // 2, n)))
    assert f.all() and (u64s(v) == np.arange(n // 2, n, dtype=np.uint64)).all()
    v, f = m.bulk_get(le(range(10_000, 10_000 + n // 2)))
    assert f.all() and (u64s(v) == np.arange(10_000, 10_000 + n // 2, dtype=np.uint64) * 3).all()


def t_robin_hood_update_at_high_load(make):  # :365-392
    n = 1024 * 45 // 100
    m = make(1024, 8, 8, 0)
    m.bulk_insert(le(range(n)), le(range(n)))
    assert m.bulk_insert(le(range(n)), le([i + 999 for i in range(n)])) == 0
    assert len(m) == n
    v, f = m.bulk_get(le(range(n)))
    assert f.all() and (u64s(v) == np.arange(n, dtype=np.uint64) + 999).all()


REFERENCE_TESTS = [
    t_insert_and_get, t_get_missing_keys, t_update_existing_keys, t_remove_keys,
    t_insert_after_remove_reuses_tombstones, t_clear_resets_map, t_murmur3_hash_strategy,
    t_large_keys_and_values, t_load_factor_and_capacity, t_empty_operations_are_no_ops,
    t_table_full_error, t_debug_formatting, t_u64_key_value_types,
    t_robin_hood_high_load_correctness, t_robin_hood_miss_exits_early,
    t_robin_hood_insert_remove_reinsert, t_robin_hood_update_at_high_load,
]


# ---------------------------------------------------------------------------
# Appendix-style anomaly histories on 16-slot identity tables (home = k & 15).
# Expected results are the sequential CPU path's.
# ---------------------------------------------------------------------------
ANOMALY_CASES = {
    # duplicates in one batch: last value wins, counted once
    "dup_in_batch": ([("ins", [(7, 1), (7, 2), (7, 3)]), ("get", [7])], [1, [3]]),
    # an update that passes a tombstone duplicates the key; removing one copy
    # resurfaces the stale one
    "update_past_tombstone": ([("ins", [(4, 1), (20, 2)]), ("rem", [4]), ("ins", [(20, 99)]),
                               ("get", [20]), ("rem", [20]), ("get", [20])],
                              [2, 1, 1, [99], 1, [2]]),
    # a duplicate remove in one batch removes the stale copy too
    "double_remove": ([("ins", [(4, 1), (20, 2)]), ("rem", [4]), ("ins", [(20, 99)]),
                       ("rem", [20, 20]), ("get", [20])], [2, 1, 1, 2, [None]]),
    # the physical order of equal-home keys is observable
    "order_a": ([("ins", [(4, 0), (20, 0), (36, 0)]), ("rem", [4]), ("ins", [(5, 0)]), ("get", [36])],
                [3, 1, 1, [0]]),
    "order_b": ([("ins", [(20, 0), (4, 0), (36, 0)]), ("rem", [4]), ("ins", [(5, 0)]), ("get", [36])],
                [3, 1, 1, [None]]),
    # a carried (evicted) key merges into its stale duplicate
    "carry_merges_stale": ([("ins", [(4, 1), (20, 2)]), ("rem", [4]), ("ins", [(20, 99)]),
                            ("ins", [(3, 30), (19, 190)])], [2, 1, 1, 1]),
}


def run_history(m, ops):
    out = []
    for op, arg in ops:
        if op == "ins":
            out.append(m.bulk_insert(np.array([k for k, _ in arg], dtype=np.uint64),
                                     np.array([v for _, v in arg], dtype=np.uint64)))
        elif op == "rem":
            out.append(m.bulk_remove(np.array(arg, dtype=np.uint64)))
        else:
            v, f = m.bulk_get(np.array(arg, dtype=np.uint64))
            vv = u64s(v)
            out.append([int(vv[i]) if f[i] else None for i in range(len(arg))])
    return out


# ---------------------------------------------------------------------------
# Differential runner: apply the same op sequence to a map and the oracle,
# compare every result and the whole logical slot array after each op.
# ---------------------------------------------------------------------------
def assert_same_slots(a, b, where=""):
    fa, ka, va = a.export_slots()
    fb, kb, vb = b.export_slots()
    if not np.array_equal(fa, fb):
        bad = np.nonzero(fa != fb)[0][:8]
        raise AssertionError(f"{where}: flags differ at {bad.tolist()}: {fa[bad].tolist()} vs {fb[bad].tolist()}")
    assert np.array_equal(ka, kb), f"{where}: keys differ"
    assert np.array_equal(va, vb), f"{where}: values differ"


def apply_and_compare(m, o, ops, check_slots=True):
    """ops: list of (name, keys[, values]) with numpy byte rows."""
    for step, op in enumerate(ops):
        name = op[0]
        if name == "ins":
            ea = eb = None
            try:
                ra = m.bulk_insert(op[1], op[2])
            except fm.TableFull as e:
                ra, ea = None, (e.occupied, e.capacity)
            try:
                rb = o.bulk_insert(op[1], op[2])
            except fm.TableFull as e:
                rb, eb = None, (e.occupied, e.capacity)
            assert (ra, ea) == (rb, eb), f"step {step} insert: {ra, ea} vs {rb, eb}"
        elif name == "rem":
            ra, rb = m.bulk_remove(op[1]), o.bulk_remove(op[1])
            assert ra == rb, f"step {step} remove: {ra} vs {rb}"
        elif name == "get":
            va, fa = m.bulk_get(op[1])
            vb, fb = o.bulk_get(op[1])
            assert np.array_equal(fa, fb), f"step {step} get: found differs at {np.nonzero(fa != fb)[0][:8]}"
            assert np.array_equal(va, vb), f"step {step} get: values differ"
        elif name == "clear":
            m.clear()
            o.clear()
        assert len(m) == len(o), f"step {step}: len {len(m)} vs {len(o)}"
        if check_slots:
            assert_same_slots(m, o, f"after step {step} ({name})")


def random_history(rng, cap, K, V, n_ops, max_batch, key_space):
    """Collision-heavy op mix: keys from a small space, duplicates allowed."""
    ops = []
    for _ in range(n_ops):
        r = rng.random()
        n = int(rng.integers(0, max_batch + 1))
        ks = rng.integers(0, key_space, size=n).astype(np.uint64)
        keys = np.zeros((n, K), dtype=np.uint8)
        keys[:, :min(8, K)] = ks.view(np.uint8).reshape(-1, 8)[:, :min(8, K)]
        if r < 0.5:
            vals = rng.integers(0, 256, size=(n, V), dtype=np.uint8)
            ops.append(("ins", keys, vals))
        elif r < 0.75:
            ops.append(("rem", keys))
        elif r < 0.98:
            ops.append(("get", keys))
        else:
            ops.append(("clear",))
    return ops
