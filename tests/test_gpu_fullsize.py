"""Production-scale parity on the GPU: the account-state step, the 512-table
batching step and the serving log at their full sizes, against the
sequential oracle.

  * account state: FlashMap<[u8;32],[u8;128]>, 2^24 slots, 8M preloaded,
    then 20 insert batches of 100k, 1M gets, 200k removes; every count, every
    get result and the whole logical slot array after the step;
  * batching: 512 tables (2^10..2^22 slots, 356M slots in all), one fused
    hipGraph step against one plain call per table-op and against the
    oracle, table by table (slot-array digests, counts, TableFull, gets);
  * serving: 2^25 slots, 20M preloaded, 1000 Poisson requests at 10k/s with
    pinned request buffers, every request against the oracle's replay of the
    log in arrival order.
"""
import hashlib
import os
import sys

import numpy as np
import pytest

import flashmap_amd as fm
from oracle.binding import OracleMap

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bench():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import bench
    return bench


def test_account_state_full_step():
    from flashmap_amd.workloads import config2
    say("account state full step: generating")
    cfg = config2()
    m = fm.FlashMap.builder(cfg.capacity, 32, 128).require_gpu().build()
    assert m.backend() == fm.Backend.GPU and m.capacity() == 1 << 24
    o = OracleMap(cfg.capacity, 32, 128, 0)
    say("preloading 8M")
    dk, dv = m.upload(cfg.preload_keys), m.upload(cfg.preload_vals)
    assert m.bulk_insert_device(dk, dv, len(cfg.preload_keys)) == len(cfg.preload_keys)
    del dk, dv
    assert o.bulk_insert(cfg.preload_keys, cfg.preload_vals) == len(cfg.preload_keys)
    say("20 insert batches")
    for b, (k, v) in enumerate(cfg.batches):
        dk, dv = m.upload(k), m.upload(v)
        assert m.bulk_insert_device(dk, dv, len(k)) == o.bulk_insert(k, v), f"batch {b}"
    say("1M gets, 200k removes")
    dq = m.upload(cfg.get_keys)
    out, found = m.alloc_device(len(cfg.get_keys) * 128), m.alloc_device(len(cfg.get_keys))
    m.bulk_get_device(dq, len(cfg.get_keys), out, found)
    ov, of = o.bulk_get(cfg.get_keys)
    assert np.array_equal(m.download(found).astype(bool), of.astype(bool))
    assert np.array_equal(m.download(out).reshape(-1), ov.reshape(-1))
    dr = m.upload(cfg.remove_keys)
    assert m.bulk_remove_device(dr, len(cfg.remove_keys)) == o.bulk_remove(cfg.remove_keys)
    assert len(m) == len(o)
    say("comparing the 2^24-slot arrays")
    gf, gk, gv = m.export_slots()
    of_, ok_, ov_ = o.export()
    assert np.array_equal(gf, of_)
    assert np.array_equal(gk, ok_)
    assert np.array_equal(gv, ov_)


def test_batching_512_tables_full_step():
    b = _bench()
    say("512-table step: fused vs plain vs oracle on every table")
    line, ok = b.run_c4(b.parse_args(["--config", "c4", "--steps", "1", "--no-autotune", "--no-cpu-baseline"]))
    assert line["config"]["slots"] > 300_000_000 and line["config"]["table_full_batches"] > 0
    assert line["check"] == {"fused_equals_plain": True, "oracle_tables": 512, "oracle_exact": True}
    assert ok


def test_serving_full_size_log():
    b = _bench()
    say("serving 2^25 slots, 20M preloaded, 1000 requests at 10k/s, pinned buffers")
    line, ok = b.run_c5(b.parse_args(["--config", "c5", "--rates", "10000", "--no-cpu-baseline"]))
    (load,) = line["loads"]
    assert load["requests"] == 1000 and load["check"] == {"requests_checked": 1000, "mismatches": 0}
    assert ok
