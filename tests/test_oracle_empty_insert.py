"""A closed form for the oracle's insert into an empty table.

A bulk_insert into a table that holds nothing fixes, from the batch alone,
which slots end up occupied and at what probe distance.  Robin Hood
insertion (rayon_cpu.rs:133-270) keeps every cluster ordered by home slot,
cyclically, so with the distinct keys sorted by home and r the rank of a key
in that order, the slot of rank r on a first lap is

    pos_r = max(home_r, pos_{r-1} + 1) = r + max_{j <= r}(home_j - j)

and the keys that run past the end of the table push the low slots on a
second lap that starts from the first lap's last position minus the
capacity:

    pos_r = r + max(M_r, last - cap + 1),   slot = pos_r mod cap,
    dist = pos_r - home_r.

The order of keys that share a home is not fixed by this: an arriving key
swaps only with a strictly richer resident, so an evicted resident walks past
the later keys of its own home and ends behind them, and the order inside a
home group depends on the eviction history.  The test therefore pins the
oracle's flags (state and distance of every slot), the set of keys in each
home group, and that every key holds the value of its last copy in the batch.
"""
import numpy as np
import pytest

from oracle.binding import OracleMap, hash_many


def closed_form(keys, vals, cap, mode):
    n, K = keys.shape
    homes = (hash_many(np.ascontiguousarray(keys).reshape(-1), K, mode) & np.uint64(cap - 1)).astype(np.int64)
    order = np.argsort(homes, kind="stable")
    hs = homes[order]
    first, last = {}, {}
    for i in range(n):
        b = keys[i].tobytes()
        first.setdefault(b, i)
        last[b] = i
    dflag = np.array([first[keys[i].tobytes()] == i for i in order], dtype=bool)
    vsrc = np.array([last[keys[i].tobytes()] for i in order], dtype=np.int64)
    rank = np.cumsum(dflag) - dflag
    v = np.where(dflag, hs - rank, np.iinfo(np.int64).min)
    M = np.maximum.accumulate(v)
    D = int(dflag.sum())
    lap_end = (D - 1) + int(M[-1])
    pos = rank + np.maximum(M, lap_end - cap + 1)
    flags = np.zeros(cap, dtype=np.uint32)
    groups = {}  # home -> set of key bytes
    value_of = {}  # key bytes -> value bytes of its last copy
    for p in np.nonzero(dflag)[0]:
        slot = int(pos[p]) % cap
        dist = int(pos[p]) - int(hs[p])
        assert 0 <= dist < cap and flags[slot] == 0
        flags[slot] = (dist << 4) | 1
        kb = keys[order[p]].tobytes()
        groups.setdefault(int(hs[p]), set()).add(kb)
        value_of[kb] = vals[vsrc[p]].tobytes()
    return D, flags, groups, value_of


def check(keys, vals, cap, mode=0):
    K, V = keys.shape[1], vals.shape[1]
    D, flags, groups, value_of = closed_form(keys, vals, cap, mode)
    o = OracleMap(cap, K, V, mode)
    assert o.bulk_insert(keys, vals) == D
    of, ok, ov = o.export()
    o.close()
    assert np.array_equal(of, flags), f"flags differ at {np.nonzero(of != flags)[0][:8]}"
    got = {}
    for slot in np.nonzero(of)[0]:
        kb = ok[slot * K:(slot + 1) * K].tobytes()
        home = (int(slot) - int(of[slot] >> 4)) % cap
        got.setdefault(home, set()).add(kb)
        assert ov[slot * V:(slot + 1) * V].tobytes() == value_of[kb]
    assert got == groups


def u64_keys(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64)).view(np.uint8).reshape(-1, 8)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("mode", [0, 1])
def test_random_batches_with_copies(seed, mode):
    rng = np.random.default_rng(seed)
    cap = 1 << int(rng.integers(4, 12))
    n = int(rng.integers(1, cap))
    keys = u64_keys(rng.integers(0, max(2, n // 2), n))  # many later copies
    vals = rng.integers(0, 256, (n, 8), dtype=np.uint8)
    check(keys, vals, cap, mode)


@pytest.mark.parametrize("cap", [16, 64, 1024])
def test_one_free_slot_left(cap):
    rng = np.random.default_rng(cap)
    keys = u64_keys(rng.integers(0, 1 << 62, cap - 1))
    vals = rng.integers(0, 256, (cap - 1, 8), dtype=np.uint8)
    check(keys, vals, cap, 1)


def test_wrapping_cluster_identity_hash():
    """Homes crowded just below the end of the table: the cluster runs past
    slot cap-1 and pushes keys whose homes are at the start."""
    cap = 256
    rng = np.random.default_rng(7)
    high = cap - 1 - rng.integers(0, 12, 60)   # homes in [cap-12, cap-1]
    low = rng.integers(0, 20, 40)              # homes in [0, 19]
    ids = np.concatenate([high, low]).astype(np.uint64)
    ids = ids + (np.arange(ids.size, dtype=np.uint64) << np.uint64(20))  # distinct keys, same homes
    perm = rng.permutation(ids.size)
    keys = u64_keys(ids[perm])
    vals = rng.integers(0, 256, (ids.size, 8), dtype=np.uint8)
    check(keys, vals, cap, 0)


def test_single_home_and_wide_keys():
    cap = 128
    rng = np.random.default_rng(11)
    # 32-byte keys that all hash (identity: first 8 bytes) to home 5
    keys = rng.integers(0, 256, (100, 32), dtype=np.uint8)
    keys[:, :8] = u64_keys(np.full(100, 5 + cap * 3))
    keys[50:60] = keys[10:20]  # later copies
    vals = rng.integers(0, 256, (100, 16), dtype=np.uint8)
    check(keys, vals, cap, 0)
