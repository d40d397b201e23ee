"""ctypes binding of the C oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests, the smoke check and the
benchmark's CPU-baseline leg, as the checker.  Arrays are numpy uint8/uint32.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compile liboracle.so with the committed Makefile (gcc only)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        u8p = ctypes.c_void_p
        L.or_identity_hash.restype = ctypes.c_uint64
        L.or_identity_hash.argtypes = [u8p, ctypes.c_uint32]
        L.or_murmur3_hash.restype = ctypes.c_uint64
        L.or_murmur3_hash.argtypes = [u8p, ctypes.c_uint32]
        L.or_hash_many.restype = None
        L.or_hash_many.argtypes = [u8p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, u8p]
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_bulk_get.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint64, u8p, u8p]
        L.or_bulk_insert.restype = ctypes.c_int
        L.or_bulk_insert.argtypes = [ctypes.c_void_p, u8p, u8p, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_uint64)]
        L.or_bulk_remove.restype = ctypes.c_uint64
        L.or_bulk_remove.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint64]
        L.or_clear.argtypes = [ctypes.c_void_p]
        L.or_len.restype = ctypes.c_uint64
        L.or_len.argtypes = [ctypes.c_void_p]
        L.or_capacity.restype = ctypes.c_uint64
        L.or_capacity.argtypes = [ctypes.c_void_p]
        L.or_export.argtypes = [ctypes.c_void_p, u8p, u8p, u8p]
        L.or_import.argtypes = [ctypes.c_void_p, u8p, u8p, u8p, ctypes.c_uint64]
        L.or_par_create.restype = ctypes.c_void_p
        L.or_par_create.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.c_int]
        L.or_par_destroy.argtypes = [ctypes.c_void_p]
        L.or_par_bulk_get.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint64, u8p, u8p]
        L.or_par_bulk_insert.restype = ctypes.c_int
        L.or_par_bulk_insert.argtypes = [ctypes.c_void_p, u8p, u8p, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_uint64)]
        L.or_par_bulk_remove.restype = ctypes.c_uint64
        L.or_par_bulk_remove.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint64]
        L.or_par_len.restype = ctypes.c_uint64
        L.or_par_copy_from.restype = ctypes.c_int
        L.or_par_copy_from.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.or_par_len.argtypes = [ctypes.c_void_p]
        _LIB = L
    return _LIB


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else ctypes.c_void_p(0)


def _u8(a):
    return np.ascontiguousarray(a).view(np.uint8).reshape(-1)


def hash_bytes(key: bytes, mode: int) -> int:
    buf = np.frombuffer(bytes(key) + b"\0", dtype=np.uint8)
    f = lib().or_identity_hash if mode == 0 else lib().or_murmur3_hash
    return int(f(_ptr(buf), len(key)))


def hash_many(keys: np.ndarray, key_size: int, mode: int) -> np.ndarray:
    k = _u8(keys)
    n = k.size // key_size
    out = np.zeros(n, dtype=np.uint64)
    lib().or_hash_many(_ptr(k), n, key_size, mode, _ptr(out))
    return out


class TableFull(Exception):
    pass


class OracleMap:
    """Sequential oracle map; keys/values are numpy arrays of n*K / n*V bytes."""

    def __init__(self, capacity, key_size, value_size, mode=0):
        self.K, self.V, self.mode = key_size, value_size, mode
        self.h = lib().or_create(capacity, key_size, value_size, mode)
        if not self.h:
            raise MemoryError("oracle allocation failed")

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def capacity(self):
        return int(lib().or_capacity(self.h))

    def __len__(self):
        return int(lib().or_len(self.h))

    def bulk_get(self, keys):
        k = _u8(keys)
        n = k.size // self.K
        vals = np.zeros(n * self.V, dtype=np.uint8)
        found = np.zeros(n, dtype=np.uint8)
        lib().or_bulk_get(self.h, _ptr(k), n, _ptr(vals), _ptr(found))
        return vals, found

    def bulk_insert(self, keys, values):
        k, v = _u8(keys), _u8(values)
        n = k.size // self.K
        assert v.size == n * self.V
        out = ctypes.c_uint64(0)
        rc = lib().or_bulk_insert(self.h, _ptr(k), _ptr(v), n, ctypes.byref(out))
        if rc != 0:
            raise TableFull(len(self), self.capacity)
        return int(out.value)

    def bulk_remove(self, keys):
        k = _u8(keys)
        return int(lib().or_bulk_remove(self.h, _ptr(k), k.size // self.K))

    def clear(self):
        lib().or_clear(self.h)

    def export(self):
        cap = self.capacity
        flags = np.zeros(cap, dtype=np.uint32)
        keys = np.zeros(cap * self.K, dtype=np.uint8)
        vals = np.zeros(cap * self.V, dtype=np.uint8)
        lib().or_export(self.h, _ptr(flags), _ptr(keys), _ptr(vals))
        return flags, keys, vals

    def import_slots(self, flags, keys, vals, length):
        lib().or_import(self.h, _ptr(flags), _ptr(_u8(keys)), _ptr(_u8(vals)), length)


class RayonPortMap:
    """Multi-threaded restatement of rayon_cpu.rs (CPU baseline)."""

    def __init__(self, capacity, key_size, value_size, mode=0, threads=None):
        self.K, self.V = key_size, value_size
        self.threads = threads or os.cpu_count() or 1
        self.h = lib().or_par_create(capacity, key_size, value_size, mode, self.threads)
        if not self.h:
            raise MemoryError("rayon-port allocation failed")

    def close(self):
        if self.h:
            lib().or_par_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __len__(self):
        return int(lib().or_par_len(self.h))

    def bulk_get(self, keys):
        k = _u8(keys)
        n = k.size // self.K
        vals = np.zeros(n * self.V, dtype=np.uint8)
        found = np.zeros(n, dtype=np.uint8)
        lib().or_par_bulk_get(self.h, _ptr(k), n, _ptr(vals), _ptr(found))
        return vals, found

    def bulk_insert(self, keys, values):
        k, v = _u8(keys), _u8(values)
        out = ctypes.c_uint64(0)
        rc = lib().or_par_bulk_insert(self.h, _ptr(k), _ptr(v), k.size // self.K, ctypes.byref(out))
        if rc != 0:
            raise TableFull()
        return int(out.value)

    def bulk_remove(self, keys):
        k = _u8(keys)
        return int(lib().or_par_bulk_remove(self.h, _ptr(k), k.size // self.K))

    def copy_from(self, src: "RayonPortMap", structure_only: bool = False):
        if lib().or_par_copy_from(self.h, src.h, 1 if structure_only else 0) != 0:
            raise ValueError("rayon-port maps differ in shape")
