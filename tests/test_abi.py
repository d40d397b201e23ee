"""C-ABI library (no GPU needed): exports, builder fallback, CPU backend parity.

The CPU backend is the builder's fallback arm (FlashMapBackend::Rayon,
src/lib.rs:431-441); it is checked against the oracle like the GPU backend.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import flashmap_amd as fm
from helpers import (ANOMALY_CASES, REFERENCE_TESTS, OracleAdapter, apply_and_compare,
                     random_history, run_history)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            with open(os.path.join(ROOT, "include", name)) as f:
                text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
            syms |= set(re.findall(r"\b(fm_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    fm.lib()  # loads (raises when the library is missing)
    out = subprocess.run(["nm", "-D", "--defined-only", fm.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = _declared_symbols()
    assert declared, "no declarations parsed"
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    for s in declared:
        assert hasattr(fm.lib(), s)


def cpu_make(cap, K, V, mode):
    return fm.FlashMap.builder(cap, K, V).hash_strategy(mode).force_cpu().build()


@pytest.mark.parametrize("test", REFERENCE_TESTS, ids=lambda t: t.__name__[2:])
def test_reference_tests_on_cpu_backend(test):
    test(cpu_make)


@pytest.mark.parametrize("name", sorted(ANOMALY_CASES))
def test_anomaly_histories_cpu_backend(name):
    ops, expected = ANOMALY_CASES[name]
    assert run_history(cpu_make(16, 8, 8, 0), ops) == expected


@pytest.mark.parametrize("seed", range(8))
def test_cpu_backend_vs_oracle_random(seed):
    rng = np.random.default_rng(300 + seed)
    cap = [16, 64, 256, 1024][seed % 4]
    K, V = [(8, 8), (32, 128), (3, 5), (12, 20)][seed % 4]
    mode = seed % 2
    ops = random_history(rng, cap, K, V, 30, cap // 2, cap * 2)
    apply_and_compare(cpu_make(cap, K, V, mode), OracleAdapter(cap, K, V, mode), ops)


def test_builder_semantics_without_gpu():
    import ctypes

    n = ctypes.c_int(0)
    has_gpu = False
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        has_gpu = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        pass
    if has_gpu:
        pytest.skip("a GPU is present")
    # default build falls back to the CPU arm (lib.rs:431-441)
    m = fm.FlashMap.builder(100, 8, 8).quiet().build()
    assert m.backend() == fm.Backend.CPU and m.capacity() == 128
    # require_gpu reports the GPU error instead of falling back
    with pytest.raises(fm.FlashMapError) as e:
        fm.FlashMap.builder(100, 8, 8).require_gpu().build()
    assert e.value.status == fm.DEVICE_INIT


def test_error_paths_cpu():
    m = cpu_make(0, 8, 8, 0)  # Rayon tier maps 0 to 16 (rayon_cpu.rs:64)
    assert m.capacity() == 16
    with pytest.raises(fm.GpuRequired):
        m.upload(b"12345678")
    with pytest.raises(fm.GpuRequired):
        m.recount()
    out = ctypes.c_uint64(0)  # the insert that allocates nothing is GPU-only too
    assert fm.lib().fm_bulk_insert_device_noalloc(m._h, None, None, 1, ctypes.byref(out)) == fm.GPU_REQUIRED
    with pytest.raises(ValueError):
        m.bulk_insert(np.zeros(3, dtype=np.uint64), np.zeros(2, dtype=np.uint64))
    with pytest.raises(fm.FlashMapError) as e:
        fm.FlashMap.builder(16, 0, 8).build()
    assert e.value.status == fm.INVALID_ARGUMENT
    assert "FlashMap" in repr(m) and 'backend: "CPU"' in repr(m)


def test_bulk_get_into_caller_arrays_cpu():
    m = cpu_make(64, 8, 8, 0)
    ks = np.arange(20, dtype=np.uint64)
    m.bulk_insert(ks[:10], ks[:10] * 3)
    out = np.full((20, 8), 9, dtype=np.uint8)
    found = np.zeros(20, dtype=np.bool_)
    gv, gf = m.bulk_get(ks, out=out, found=found)
    assert gv is out and gf is found
    ev, ef = m.bulk_get(ks)
    assert np.array_equal(out, ev) and np.array_equal(found, ef)
    assert found[:10].all() and not found[10:].any() and not out[10:].any()  # misses zeroed
    for bad in (np.zeros((20, 8), dtype=np.uint16), np.zeros((19, 8), dtype=np.uint8),
                np.zeros((8, 20), dtype=np.uint8).T):
        with pytest.raises(ValueError):
            m.bulk_get(ks, out=bad)
    with pytest.raises(ValueError):
        m.bulk_get(ks, found=np.zeros(20, dtype=np.uint8))


def test_table_full_all_update_batch_cpu():
    # TableFull fires on len + n > cap even when every key is an update
    # (rayon_cpu.rs:134-142)
    m = cpu_make(16, 8, 8, 0)
    ks = np.arange(10, dtype=np.uint64)
    m.bulk_insert(ks, ks)
    with pytest.raises(fm.TableFull) as e:
        m.bulk_insert(ks, ks)
    assert (e.value.occupied, e.value.capacity) == (10, 16)
    assert abs(e.value.load_factor - 62.5) < 1e-12


def test_async_wrapper_cpu():
    import asyncio

    m = cpu_make(1024, 8, 8, 0)
    am = fm.AsyncFlashMap(m)

    async def go():
        ks = np.arange(100, dtype=np.uint64)
        assert await am.bulk_insert(ks, ks * 2) == 100
        res = await asyncio.gather(*[am.bulk_get(ks) for _ in range(8)])
        for v, f in res:
            assert f.all() and (v.view(np.uint64).reshape(-1) == ks * 2).all()
        assert await am.bulk_remove(ks[:10]) == 10
        assert am.len() == 90 and am.clone().capacity() == 1024
        await am.clear()
        assert am.is_empty()

    asyncio.run(go())
