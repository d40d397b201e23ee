"""Serving loop (fm_serve_log) vs a sequential replay of the same log.

The server coalesces adjacent same-kind requests, admits them by staging
memory and overlaps copies with kernels; every request's results must equal
(a) the oracle applying the log one request at a time in arrival order and
(b) the same request run alone through the plain bulk API on a second GPU
map ("the same alone and in the mix").  Shapes follow config 5
([u8;32] -> u64) at reduced sizes.
"""
import sys
from types import SimpleNamespace

import numpy as np
import pytest

import flashmap_amd as fm
from flashmap_amd.workloads import Stream
from helpers import OracleAdapter, assert_same_slots

pytestmark = pytest.mark.gpu
K, V = 32, 8


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _log(seed, cap, n_pre, n_req, size_hi, rate, p_get=0.5, p_ins=0.3):
    st = Stream(seed)
    pre = st.bytes(n_pre, K)
    pre_v = st.u64(n_pre)
    gaps = -np.log1p(-st.unit(n_req)) / rate
    arr = np.cumsum(gaps)
    reqs = []
    for i in range(n_req):
        u = float(st.unit(1)[0])
        n = int(st.log_uniform(1, size_hi, 1)[0])
        if u < p_get:
            hit = pre[st.below(n_pre, n).astype(np.int64)]
            ks = np.concatenate([hit[: n - n // 4], st.bytes(n // 4, K)])
            reqs.append(SimpleNamespace(op="get", keys=ks[st.permutation(n)], arrival_s=arr[i]))
        elif u < p_get + p_ins:
            # overwrites (with duplicates inside the request) and new keys
            over = pre[st.below(n_pre, n // 2 + 1).astype(np.int64)][: n // 2]
            ks = np.concatenate([over, st.bytes(n - n // 2, K)])
            reqs.append(SimpleNamespace(op="insert", keys=ks[st.permutation(n)], vals=st.u64(n), arrival_s=arr[i]))
        else:
            ks = np.concatenate([pre[st.below(n_pre, n).astype(np.int64)][: n - n // 5], st.bytes(n // 5, K)])
            reqs.append(SimpleNamespace(op="remove", keys=ks[st.permutation(n)], arrival_s=arr[i]))
    return pre, pre_v, reqs


def _expected(cap, pre, pre_v, reqs):
    o = OracleAdapter(cap, K, V)
    o.bulk_insert(pre, pre_v)
    out = []
    for r in reqs:
        if r.op == "get":
            v, f = o.bulk_get(r.keys)
            out.append(("get", v, f, int(f.sum()), fm.OK))
        elif r.op == "insert":
            try:
                out.append(("insert", None, None, o.bulk_insert(r.keys, r.vals), fm.OK))
            except fm.TableFull:
                out.append(("insert", None, None, 0, fm.TABLE_FULL))
        else:
            out.append(("remove", None, None, o.bulk_remove(r.keys), fm.OK))
    return o, out


def _check(results, expected, what):
    for i, (res, exp) in enumerate(zip(results, expected)):
        op, ev, ef, count, status = exp
        assert (res.count, res.status) == (count, status), f"{what}: request {i} ({op})"
        if op == "get":
            assert np.array_equal(res.found.astype(bool), ef), f"{what}: request {i} found"
            assert np.array_equal(res.values, ev), f"{what}: request {i} values"


@pytest.mark.parametrize("realtime,rate,staging,pinned", [
    (False, 1.0, 64 << 20, False), (True, 20_000.0, 64 << 20, False), (False, 1.0, 1 << 20, False),
    (False, 1.0, 64 << 20, True), (True, 20_000.0, 64 << 20, True), (False, 1.0, 1 << 20, "inputs"),
    (False, 1.0, 64 << 20, "outputs")],
    ids=["burst", "poisson20k", "tiny-staging", "burst-pinned", "poisson20k-pinned", "tiny-staging-pinned-inputs",
         "burst-pinned-outputs"])
def test_serve_log_matches_sequential_replay(realtime, rate, staging, pinned):
    cap = 1 << 16
    pre, pre_v, reqs = _log(5, cap, 30_000, 300, 4096, rate)
    say(f"serve realtime={realtime} rate={rate} staging={staging} pinned={pinned}")
    m = fm.FlashMap.builder(cap, K, V).require_gpu().build()
    m.bulk_insert(pre, pre_v)
    results, stats = fm.serve_log(m, reqs, staging_bytes=staging, max_batch_keys=1 << 14, realtime=realtime,
                                  pinned=pinned)
    o, expected = _expected(cap, pre, pre_v, reqs)
    _check(results, expected, "served")
    assert len(m) == len(o)
    assert_same_slots(m, o, "after serving")
    assert stats.requests == len(reqs) and stats.batches <= len(reqs)
    if not realtime:
        assert stats.batches < len(reqs), "adjacent same-kind requests should coalesce"


@pytest.mark.parametrize("pinned", [False, True], ids=["pageable", "pinned"])
def test_serve_alone_equals_in_the_mix_and_table_full(pinned):
    """Each request run alone through the bulk API equals its served result;
    the log drives the table to TableFull so that path is exercised too."""
    cap = 1 << 13
    pre, pre_v, reqs = _log(7, cap, 5_000, 200, 1024, 1.0, p_get=0.4, p_ins=0.5)
    say(f"serve alone vs mix pinned={pinned}")
    served_map = fm.FlashMap.builder(cap, K, V).require_gpu().build()
    alone_map = fm.FlashMap.builder(cap, K, V).require_gpu().build()
    served_map.bulk_insert(pre, pre_v)
    alone_map.bulk_insert(pre, pre_v)
    results, _ = fm.serve_log(served_map, reqs, staging_bytes=8 << 20, max_batch_keys=4096, realtime=False,
                              pinned=pinned)
    saw_full = False
    for i, (r, res) in enumerate(zip(reqs, results)):
        if r.op == "get":
            v, f = alone_map.bulk_get(r.keys)
            assert np.array_equal(res.values, v) and np.array_equal(res.found.astype(bool), f), f"request {i}"
        elif r.op == "insert":
            try:
                assert (res.count, res.status) == (alone_map.bulk_insert(r.keys, r.vals), fm.OK), f"request {i}"
            except fm.TableFull:
                saw_full = True
                assert res.status == fm.TABLE_FULL, f"request {i}"
        else:
            assert res.count == alone_map.bulk_remove(r.keys), f"request {i}"
    assert saw_full, "log should reach TableFull"
    assert_same_slots(served_map, alone_map, "served vs alone")


def test_host_alloc_and_free():
    """fm_host_alloc hands out pinned memory the serving loop recognises;
    fm_host_free accepts exactly those pointers, once."""
    import ctypes
    from flashmap_amd import lib
    say("host alloc/free")
    a = fm.PinnedArena(1 << 20)
    x = a.take(1000, fill=7)
    assert x.sum() == 7000
    inner = ctypes.c_void_p(a.ptr + 256)
    assert lib().fm_host_free(inner) == fm.INVALID_ARGUMENT
    a.close()
    assert lib().fm_host_free(ctypes.c_void_p(12345)) == fm.INVALID_ARGUMENT
    assert lib().fm_host_free(ctypes.c_void_p(0)) == fm.OK


@pytest.mark.parametrize("pinned", [False, True], ids=["pageable", "pinned"])
def test_serve_empty_requests_in_the_mix(pinned):
    """Requests with no keys (of every kind) sit between ordinary ones: they
    complete with count 0 and leave the neighbours' results exact."""
    cap = 1 << 14
    pre, pre_v, reqs = _log(11, cap, 6_000, 60, 2048, 1.0)
    for i in range(0, len(reqs), 7):
        r = reqs[i]
        r.keys = r.keys[:0]
        if r.op == "insert":
            r.vals = r.vals[:0]
    say(f"serve with empty requests pinned={pinned}")
    m = fm.FlashMap.builder(cap, K, V).require_gpu().build()
    m.bulk_insert(pre, pre_v)
    results, _ = fm.serve_log(m, reqs, staging_bytes=8 << 20, max_batch_keys=4096, realtime=False, pinned=pinned)
    o, expected = _expected(cap, pre, pre_v, reqs)
    _check(results, expected, "served")
    assert all(results[i].count == 0 for i in range(0, len(reqs), 7))
    assert len(m) == len(o)
    assert_same_slots(m, o, "after serving")


def test_host_buffer_calls_with_pinned_buffers():
    """fm_bulk_insert / fm_bulk_get / fm_bulk_remove with host buffers of
    more than 1 MiB that come from fm_host_alloc (moved by DMA directly) and
    from ordinary memory (copied through pinned staging): both equal the
    oracle."""
    import ctypes
    from flashmap_amd import lib
    say("host-buffer calls with pinned buffers")
    cap, n = 1 << 19, 200_000
    rng = np.random.default_rng(77)
    keys = rng.integers(0, 1 << 62, n).astype(np.uint64).view(np.uint8).reshape(-1, 8)
    vals = rng.integers(0, 256, (n, 8), dtype=np.uint8)
    q = np.concatenate([keys[::2], rng.integers(0, 1 << 62, n // 2).astype(np.uint64).view(np.uint8).reshape(-1, 8)])
    arena = fm.PinnedArena(6 * n * 8 + n + (1 << 16))
    pk, pv, pq = arena.copy_of(keys), arena.copy_of(vals), arena.copy_of(q)
    pout, pfound = arena.take((len(q), 8)), arena.take(len(q))
    o = OracleAdapter(cap, 8, 8)
    m = fm.FlashMap.builder(cap, 8, 8).require_gpu().build()
    vp = ctypes.c_void_p
    new = ctypes.c_uint64(0)
    assert lib().fm_bulk_insert(m._h, vp(pk.ctypes.data), vp(pv.ctypes.data), n, ctypes.byref(new)) == fm.OK
    assert new.value == o.bulk_insert(keys, vals)
    assert lib().fm_bulk_get(m._h, vp(pq.ctypes.data), len(q), vp(pout.ctypes.data), vp(pfound.ctypes.data)) == fm.OK
    ev, ef = o.bulk_get(q)
    assert np.array_equal(pout, ev) and np.array_equal(pfound.astype(bool), ef)
    rem = ctypes.c_uint64(0)
    rk = np.ascontiguousarray(keys[: 3 * n // 4])  # pageable, staged
    assert lib().fm_bulk_remove(m._h, vp(rk.ctypes.data), len(rk), ctypes.byref(rem)) == fm.OK
    assert rem.value == o.bulk_remove(rk)
    rk2 = arena.copy_of(np.concatenate([keys[n // 2:], keys[n // 2:]]))  # pinned, with duplicates
    assert lib().fm_bulk_remove(m._h, vp(rk2.ctypes.data), len(rk2), ctypes.byref(rem)) == fm.OK
    assert rem.value == o.bulk_remove(rk2)
    v2, f2 = m.bulk_get(q)  # pageable buffers, staged
    ev, ef = o.bulk_get(q)
    assert np.array_equal(v2, ev) and np.array_equal(f2, ef)
    assert len(m) == len(o)
    arena.close()
