"""fm_bulk_insert_device_noalloc (the insert that allocates nothing) against
fm_bulk_insert_device: the same inputs through the C ABI on two maps give
the same counts, lengths and logical slot arrays, call after call (each
second call smaller than the first), on the led path, on the deferred path
whose closures need the big-closure replay, and on the sort-grouped path of
a batch that may fill the table; sizes above the fixed scratch are refused,
and so are full tables."""
import sys

import numpy as np
import pytest

import flashmap_amd as fm
from helpers import OracleAdapter, assert_same_slots

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pair(cap, K, V):
    a = fm.FlashMap.builder(cap, K, V).require_gpu().build()
    b = fm.FlashMap.builder(cap, K, V).require_gpu().build()
    assert a.backend() == fm.Backend.GPU and b.backend() == fm.Backend.GPU
    return a, b


def both(a, b, o, keys, vals):
    """One batch through each entry point (and the oracle); same count."""
    dka, dva = a.upload(keys), a.upload(vals)
    dkb, dvb = b.upload(keys), b.upload(vals)
    na = a.bulk_insert_device(dka, dva, len(keys))
    nb = b.bulk_insert_device_noalloc(dkb, dvb, len(keys))
    no = o.bulk_insert(keys, vals)
    assert na == nb == no, (na, nb, no)
    assert len(a) == len(b) == len(o)


@pytest.mark.parametrize("K,V", [(8, 8), (32, 128), (32, 8)])
def test_noalloc_matches_plain_insert(K, V):
    say(f"noalloc vs plain K={K} V={V}")
    cap = 1 << 20
    rng = np.random.default_rng(K * 1000 + V)
    a, b = pair(cap, K, V)
    o = OracleAdapter(cap, K, V)
    pool = rng.integers(0, 256, (700_000, K), dtype=np.uint8)
    # first call large (dense: n * 16 >= cap), the second smaller, then sparse
    # batches of overwrites and new keys (settled per key, closures deferred)
    sizes = [400_000, 150_000, 40_000, 20_000, 5_000]
    start = 0
    for n in sizes:
        over = pool[rng.integers(0, max(1, start), n // 2)] if start else pool[start:start + n // 2]
        new = pool[start:start + n - n // 2]
        start += n - n // 2
        keys = np.concatenate([over, new])[rng.permutation(n)]
        vals = rng.integers(0, 256, (n, V), dtype=np.uint8)
        both(a, b, o, keys, vals)
        assert_same_slots(a, b, f"after a batch of {n}")
    assert_same_slots(b, o, "against the oracle")


def test_noalloc_deferred_big_closures():
    # A near-full small table: sparse batches queue closures for the
    # big-closure replay (the settle pass hands over and the host runs it),
    # whose queue lives in the fixed scratch.
    say("noalloc deferred big closures")
    cap = 1 << 16
    rng = np.random.default_rng(5)
    a, b = pair(cap, 8, 8)
    o = OracleAdapter(cap, 8, 8)
    k0 = rng.integers(0, 1 << 62, 60_000, dtype=np.uint64).view(np.uint8).reshape(-1, 8)
    both(a, b, o, k0, rng.integers(0, 256, (len(k0), 8), dtype=np.uint8))
    for n in (3000, 1500, 700):
        k = rng.integers(0, 1 << 62, n, dtype=np.uint64).view(np.uint8).reshape(-1, 8)
        both(a, b, o, k, rng.integers(0, 256, (n, 8), dtype=np.uint8))
        assert_same_slots(a, b, f"deferred batch {n}")
    assert_same_slots(b, o, "against the oracle")


def test_noalloc_sort_grouped_fill_and_limits():
    # A batch that may fill the table takes the sort-grouped path (home,
    # closure ids, sort buffers in the fixed scratch); then a smaller one.
    say("noalloc sort-grouped path and limits")
    cap = 1 << 17
    rng = np.random.default_rng(11)
    a, b = pair(cap, 32, 8)
    o = OracleAdapter(cap, 32, 8)
    k0 = rng.integers(0, 256, (100_000, 32), dtype=np.uint8)
    both(a, b, o, k0, rng.integers(0, 256, (len(k0), 8), dtype=np.uint8))
    room = cap - len(o)
    k1 = np.concatenate([rng.integers(0, 256, (room - 5000, 32), dtype=np.uint8), k0[:5000]])
    both(a, b, o, k1, rng.integers(0, 256, (len(k1), 8), dtype=np.uint8))
    assert_same_slots(a, b, "fill")
    k2 = k0[rng.integers(0, len(k0), 1000)]
    both(a, b, o, k2, rng.integers(0, 256, (len(k2), 8), dtype=np.uint8))
    assert_same_slots(a, b, "after fill")
    assert_same_slots(b, o, "against the oracle")
    # a full table refuses a batch with TableFull, before any change
    extra = rng.integers(0, 256, (cap - len(b) + 1, 32), dtype=np.uint8)
    de, dv = b.upload(extra), b.upload(np.zeros((len(extra), 8), np.uint8))
    with pytest.raises(fm.TableFull):
        b.bulk_insert_device_noalloc(de, dv, len(extra))
    # a batch above the fixed scratch is refused with a clear error
    big = fm.FlashMap.builder(1 << 22, 8, 8).require_gpu().build()
    n = fm.NOALLOC_MAX_KEYS + 1
    dk, dv = big.alloc_device(n * 8), big.alloc_device(n * 8)
    with pytest.raises(fm.FlashMapError, match="at most"):
        big.bulk_insert_device_noalloc(dk, dv, n)
    assert len(big) == 0
    assert big.bulk_insert_device_noalloc(dk, dv, 0) == 0
