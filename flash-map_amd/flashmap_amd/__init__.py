"""Python mirror of the reference FlashMap API over the C ABI.

Mirrors src/lib.rs (FlashMap, FlashMapBuilder), src/hash.rs (HashStrategy),
src/error.rs (FlashMapError) and src/async_map.rs (AsyncFlashMap).  Keys and
values are fixed-size byte strings: any numpy array is viewed as raw bytes,
``n = nbytes // key_size`` (bytemuck::cast_slice, src/gpu.rs:120).

The shared library ``libflashmap_amd.so`` next to this file must be built
(``make -C flash-map_amd``); importing fails loudly when it is missing.
"""
from __future__ import annotations

import asyncio
import ctypes
import os
import threading
from typing import List, Optional, Sequence, Tuple

import numpy as np

__all__ = [
    "FlashMap", "FlashMapBuilder", "HashStrategy", "Backend", "FlashMapError",
    "TableFull", "GpuRequired", "AsyncFlashMap", "lib", "LIB_PATH",
]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libflashmap_amd.so")
NOALLOC_MAX_KEYS = 1 << 20  # FM_NOALLOC_MAX_KEYS in include/flashmap.h

# fm_status (include/flashmap.h) == FlashMapError variants (src/error.rs:4-39)
OK, DEVICE_INIT, GPU_ALLOC, KERNEL_LAUNCH, TRANSFER, TABLE_FULL, ZERO_CAPACITY, \
    NO_BACKEND, GPU_REQUIRED, LOCK_POISONED, ASYNC_JOIN, INVALID_ARGUMENT = range(12)


class HashStrategy:
    """src/hash.rs:3-27 (numbering = HashStrategy::to_mode)."""
    IDENTITY = 0
    MURMUR3 = 1


class Backend:
    """enum FlashMapBackend (src/lib.rs:98-103)."""
    NONE = 0
    GPU = 1
    CPU = 2


class FlashMapError(Exception):
    """src/error.rs:4-39; ``status`` is the fm_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class TableFull(FlashMapError):
    """FlashMapError::TableFull (error.rs:17-24); load_factor in percent."""

    def __init__(self, status, message, occupied, capacity, load_factor):
        super().__init__(status, message)
        self.occupied = occupied
        self.capacity = capacity
        self.load_factor = load_factor


class GpuRequired(FlashMapError):
    """FlashMapError::GpuRequired (error.rs:32-33)."""


OP_GET, OP_INSERT, OP_REMOVE = 0, 1, 2


class ServeRequest(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("n", ctypes.c_uint64),
                ("keys", ctypes.c_void_p), ("values", ctypes.c_void_p), ("out_values", ctypes.c_void_p),
                ("out_found", ctypes.c_void_p), ("out_flags", ctypes.c_void_p), ("arrival_s", ctypes.c_double),
                ("count", ctypes.c_uint64), ("status", ctypes.c_int32), ("batch", ctypes.c_int32),
                ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("staging_bytes", ctypes.c_uint64), ("max_batch_keys", ctypes.c_uint64),
                ("realtime", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class ServeStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("keys", ctypes.c_uint64),
                ("wall_s", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("max_inflight_batches", ctypes.c_uint64), ("staging_waits", ctypes.c_uint64)]


class GroupTuning(ctypes.Structure):
    _fields_ = [("group", ctypes.c_uint32), ("block", ctypes.c_uint32), ("grid", ctypes.c_uint32),
                ("coop_arc", ctypes.c_uint32)]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double)]


class _Builder(ctypes.Structure):
    _fields_ = [
        ("capacity", ctypes.c_uint64),
        ("key_size", ctypes.c_uint32),
        ("value_size", ctypes.c_uint32),
        ("hash_strategy", ctypes.c_uint32),
        ("device_id", ctypes.c_int32),
        ("force_cpu", ctypes.c_uint32),
        ("require_gpu", ctypes.c_uint32),
        ("quiet", ctypes.c_uint32),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("get_group", ctypes.c_uint32),
        ("get_block", ctypes.c_uint32),
        ("remove_group", ctypes.c_uint32),
        ("insert_block", ctypes.c_uint32),
        ("insert_coop_arc", ctypes.c_uint32),
        ("insert_fused", ctypes.c_uint32),
        ("insert_dedup", ctypes.c_uint32),
    ]


_LIB = None
_LIB_LOCK = threading.Lock()


def lib() -> ctypes.CDLL:
    """Loads libflashmap_amd.so (raises if it was not built)."""
    global _LIB
    with _LIB_LOCK:
        if _LIB is not None:
            return _LIB
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C flash-map_amd` "
                              "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
        pu64 = ctypes.POINTER(ctypes.c_uint64)
        sig = {
            "fm_builder_init": (None, [ctypes.POINTER(_Builder), u64, u32, u32]),
            "fm_build": (ctypes.c_int, [ctypes.POINTER(_Builder), ctypes.POINTER(vp)]),
            "fm_with_capacity": (ctypes.c_int, [u64, u32, u32, ctypes.POINTER(vp)]),
            "fm_destroy": (None, [vp]),
            "fm_bulk_get": (ctypes.c_int, [vp, vp, u64, vp, vp]),
            "fm_bulk_insert": (ctypes.c_int, [vp, vp, vp, u64, pu64]),
            "fm_bulk_remove": (ctypes.c_int, [vp, vp, u64, pu64]),
            "fm_len": (u64, [vp]),
            "fm_capacity": (u64, [vp]),
            "fm_load_factor": (ctypes.c_double, [vp]),
            "fm_clear": (ctypes.c_int, [vp]),
            "fm_backend": (ctypes.c_int, [vp]),
            "fm_key_size": (u32, [vp]),
            "fm_value_size": (u32, [vp]),
            "fm_debug_string": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t]),
            "fm_last_error_message": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
            "fm_last_table_full": (ctypes.c_int, [pu64, pu64, ctypes.POINTER(ctypes.c_double)]),
            "fm_device_stream": (vp, [vp]),
            "fm_device_id": (i32, [vp]),
            "fm_alloc_device": (ctypes.c_int, [vp, u64, ctypes.POINTER(vp)]),
            "fm_free_device": (ctypes.c_int, [vp, vp]),
            "fm_upload": (ctypes.c_int, [vp, vp, u64, ctypes.POINTER(vp)]),
            "fm_download": (ctypes.c_int, [vp, vp, u64, vp]),
            "fm_bulk_get_device": (ctypes.c_int, [vp, vp, u64, vp, vp]),
            "fm_bulk_get_values_device": (ctypes.c_int, [vp, vp, u64, vp]),
            "fm_bulk_insert_device": (ctypes.c_int, [vp, vp, vp, u64, pu64]),
            "fm_bulk_insert_device_uncounted": (ctypes.c_int, [vp, vp, vp, u64]),
            "fm_bulk_insert_device_noalloc": (ctypes.c_int, [vp, vp, vp, u64, pu64]),
            "fm_bulk_remove_device": (ctypes.c_int, [vp, vp, u64, pu64]),
            "fm_recount": (ctypes.c_int, [vp, pu64]),
            "fm_synchronize": (ctypes.c_int, [vp]),
            "fm_export_slots": (ctypes.c_int, [vp, vp, vp, vp]),
            "fm_import_slots": (ctypes.c_int, [vp, vp, vp, vp, u64]),
            "fm_copy_table": (ctypes.c_int, [vp, vp, u32]),
            "fm_serve_log": (ctypes.c_int, [vp, ctypes.POINTER(ServeRequest), u64, ctypes.POINTER(ServeConfig),
                                            ctypes.POINTER(ServeStats)]),
            "fm_host_alloc": (ctypes.c_int, [u64, ctypes.POINTER(vp)]),
            "fm_host_free": (ctypes.c_int, [vp]),
            "fm_group_create": (ctypes.c_int, [u32, pu64, u32, u32, u32, i32, ctypes.POINTER(vp)]),
            "fm_group_destroy": (None, [vp]),
            "fm_group_tables": (u32, [vp]),
            "fm_group_len": (u64, [vp, u32]),
            "fm_group_capacity": (u64, [vp, u32]),
            "fm_group_stream": (vp, [vp]),
            "fm_group_alloc_device": (ctypes.c_int, [vp, u64, ctypes.POINTER(vp)]),
            "fm_group_free_device": (ctypes.c_int, [vp, vp]),
            "fm_group_upload": (ctypes.c_int, [vp, vp, u64, ctypes.POINTER(vp)]),
            "fm_group_download": (ctypes.c_int, [vp, vp, u64, vp]),
            "fm_group_bulk_insert_device": (ctypes.c_int, [vp, vp, vp, pu64, pu64, ctypes.POINTER(i32)]),
            "fm_group_bulk_get_device": (ctypes.c_int, [vp, vp, pu64, vp, vp]),
            "fm_group_bulk_remove_device": (ctypes.c_int, [vp, vp, pu64, pu64]),
            "fm_group_step_device": (ctypes.c_int, [vp, vp, vp, pu64, vp, pu64, vp, vp, vp, pu64, pu64,
                                                    ctypes.POINTER(i32), pu64]),
            "fm_group_export_slots": (ctypes.c_int, [vp, u32, vp, vp, vp]),
            "fm_group_copy": (ctypes.c_int, [vp, vp]),
            "fm_group_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(GroupTuning)]),
            "fm_group_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(GroupTuning)]),
            "fm_group_autotune": (ctypes.c_int, [vp, vp, vp, vp, pu64, vp, pu64, vp, vp, vp, pu64,
                                                 ctypes.POINTER(GroupTuning), u32, u32,
                                                 ctypes.POINTER(GroupTuning), ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_uint32)]),
            "fm_profile_enable": (ctypes.c_int, [vp, ctypes.c_int]),
            "fm_profile_only": (ctypes.c_int, [vp, ctypes.c_char_p]),
            "fm_profile_read": (ctypes.c_int, [vp, ctypes.POINTER(KernelStat), u32,
                                               ctypes.POINTER(ctypes.c_uint32)]),
            "fm_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
            "fm_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
            "fm_autotune_get": (ctypes.c_int, [vp, vp, u64, u32, ctypes.POINTER(Tuning),
                                               ctypes.POINTER(ctypes.c_double)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
        return L


def _last_error(status: int) -> FlashMapError:
    L = lib()
    buf = ctypes.create_string_buffer(512)
    L.fm_last_error_message(buf, 512)
    msg = buf.value.decode(errors="replace")
    if status == TABLE_FULL:
        occ, cap, lf = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        L.fm_last_table_full(ctypes.byref(occ), ctypes.byref(cap), ctypes.byref(lf))
        return TableFull(status, msg, occ.value, cap.value, lf.value)
    if status == GPU_REQUIRED:
        return GpuRequired(status, msg)
    return FlashMapError(status, msg)


def _check(status: int):
    if status != OK:
        raise _last_error(status)


def _as_bytes(a, size: int, what: str) -> Tuple[np.ndarray, int]:
    """Contiguous uint8 view of `a` and its element count for `size`-byte items."""
    if isinstance(a, (bytes, bytearray, memoryview)):
        arr = np.frombuffer(bytes(a), dtype=np.uint8)
    elif isinstance(a, np.ndarray):
        arr = np.ascontiguousarray(a).reshape(-1).view(np.uint8)
    else:  # sequence of byte strings
        arr = np.frombuffer(b"".join(bytes(x) for x in a), dtype=np.uint8) if len(a) else \
            np.zeros(0, dtype=np.uint8)
    if arr.size % size:
        raise ValueError(f"{what}: {arr.size} bytes is not a multiple of {size}")
    return arr, arr.size // size


def _result_array(a, shape, dtype, what: str) -> np.ndarray:
    """A fresh zeroed result array, or the caller's `a` once it is checked to
    be a writable C-contiguous array of that shape and dtype."""
    if a is None:
        return np.zeros(shape, dtype=dtype)
    if not isinstance(a, np.ndarray) or a.dtype != dtype or a.shape != tuple(shape) \
            or not a.flags.c_contiguous or not a.flags.writeable:
        raise ValueError(f"{what}: want a writable C-contiguous {np.dtype(dtype).name} array of shape "
                         f"{tuple(shape)}")
    return a


def _ptr(arr: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(arr.ctypes.data if arr.size else 0)


class DeviceBuffer:
    """A device allocation owned by a map (CudaSlice<u8> in the reference)."""

    def __init__(self, owner: "FlashMap", ptr: int, nbytes: int):
        self.owner, self.ptr, self.nbytes = owner, ptr, nbytes

    def free(self):
        if self.ptr and self.owner._h:
            lib().fm_free_device(self.owner._h, ctypes.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FlashMap:
    """FlashMap<K,V> with K and V given as byte sizes (src/lib.rs:94-351)."""

    def __init__(self, handle: int, key_size: int, value_size: int):
        self._h = ctypes.c_void_p(handle)
        self.key_size = key_size
        self.value_size = value_size

    # -- construction (lib.rs:110-117) --
    @staticmethod
    def with_capacity(capacity: int, key_size: int = 8, value_size: int = 8) -> "FlashMap":
        return FlashMapBuilder(capacity, key_size, value_size).build()

    @staticmethod
    def builder(capacity: int, key_size: int = 8, value_size: int = 8) -> "FlashMapBuilder":
        return FlashMapBuilder(capacity, key_size, value_size)

    def close(self):
        if self._h:
            lib().fm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- host-facing API (lib.rs:123-207) --
    def bulk_get(self, keys, out=None, found=None) -> Tuple[np.ndarray, np.ndarray]:
        """Returns (values: (n, V) uint8, found: (n,) bool); misses are zeroed.

        `out` ((n, V) uint8) and `found` ((n,) bool) may be caller arrays that
        receive the results, for instance carved from a PinnedArena, which the
        DMA engines fill without a host copy; they are returned."""
        k, n = _as_bytes(keys, self.key_size, "keys")
        vals = _result_array(out, (n, self.value_size), np.uint8, "out")
        hits = _result_array(found, (n,), np.bool_, "found")
        if n:
            _check(lib().fm_bulk_get(self._h, _ptr(k), n, _ptr(vals), _ptr(hits)))
        return vals, hits

    def bulk_get_options(self, keys) -> List[Optional[bytes]]:
        """Vec<Option<V>> form of bulk_get (lib.rs:124)."""
        vals, found = self.bulk_get(keys)
        return [bytes(vals[i]) if found[i] else None for i in range(len(found))]

    def bulk_insert(self, keys, values) -> int:
        """Number of newly occupied slots (lib.rs:143-150)."""
        k, n = _as_bytes(keys, self.key_size, "keys")
        v, nv = _as_bytes(values, self.value_size, "values")
        if n != nv:
            raise ValueError(f"{n} keys but {nv} values")
        out = ctypes.c_uint64(0)
        _check(lib().fm_bulk_insert(self._h, _ptr(k), _ptr(v), n, ctypes.byref(out)))
        return int(out.value)

    def bulk_insert_pairs(self, pairs: Sequence[Tuple[bytes, bytes]]) -> int:
        """bulk_insert(&[(K,V)]) with the tuples split as gpu.rs:212-217 does."""
        return self.bulk_insert([p[0] for p in pairs], [p[1] for p in pairs])

    def bulk_remove(self, keys) -> int:
        k, n = _as_bytes(keys, self.key_size, "keys")
        out = ctypes.c_uint64(0)
        _check(lib().fm_bulk_remove(self._h, _ptr(k), n, ctypes.byref(out)))
        return int(out.value)

    def __len__(self) -> int:
        return int(lib().fm_len(self._h))

    def len(self) -> int:
        return len(self)

    def is_empty(self) -> bool:
        return len(self) == 0

    def capacity(self) -> int:
        return int(lib().fm_capacity(self._h))

    def load_factor(self) -> float:
        return float(lib().fm_load_factor(self._h))

    def clear(self):
        _check(lib().fm_clear(self._h))

    def backend(self) -> int:
        return int(lib().fm_backend(self._h))

    def __repr__(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _check(lib().fm_debug_string(self._h, buf, 256))
        return buf.value.decode()

    # -- device-resident API (lib.rs:213-350) --
    def device_stream(self) -> int:
        return int(lib().fm_device_stream(self._h) or 0)

    def alloc_device(self, nbytes: int) -> DeviceBuffer:
        p = ctypes.c_void_p()
        _check(lib().fm_alloc_device(self._h, nbytes, ctypes.byref(p)))
        return DeviceBuffer(self, p.value or 0, nbytes)

    def upload(self, data) -> DeviceBuffer:
        arr, _ = _as_bytes(data, 1, "data")
        p = ctypes.c_void_p()
        _check(lib().fm_upload(self._h, _ptr(arr), arr.size, ctypes.byref(p)))
        return DeviceBuffer(self, p.value or 0, arr.size)

    upload_keys = upload
    upload_values = upload

    def download(self, buf: DeviceBuffer, nbytes: Optional[int] = None) -> np.ndarray:
        nbytes = buf.nbytes if nbytes is None else nbytes
        out = np.zeros(nbytes, dtype=np.uint8)
        _check(lib().fm_download(self._h, ctypes.c_void_p(buf.ptr), nbytes, _ptr(out)))
        return out

    def bulk_get_device(self, d_keys: DeviceBuffer, count: int,
                        d_out: DeviceBuffer, d_found: DeviceBuffer):
        _check(lib().fm_bulk_get_device(self._h, ctypes.c_void_p(d_keys.ptr), count,
                                        ctypes.c_void_p(d_out.ptr), ctypes.c_void_p(d_found.ptr)))

    def bulk_get_values_device(self, d_keys: DeviceBuffer, count: int, d_out: DeviceBuffer):
        _check(lib().fm_bulk_get_values_device(self._h, ctypes.c_void_p(d_keys.ptr), count,
                                               ctypes.c_void_p(d_out.ptr)))

    def bulk_insert_device(self, d_keys: DeviceBuffer, d_vals: DeviceBuffer, count: int) -> int:
        out = ctypes.c_uint64(0)
        _check(lib().fm_bulk_insert_device(self._h, ctypes.c_void_p(d_keys.ptr),
                                           ctypes.c_void_p(d_vals.ptr), count, ctypes.byref(out)))
        return int(out.value)

    def bulk_insert_device_noalloc(self, d_keys: DeviceBuffer, d_vals: DeviceBuffer, count: int) -> int:
        """bulk_insert_device with no allocation: per-call buffers in the
        GPU's fixed scratch array, at most NOALLOC_MAX_KEYS keys per call."""
        out = ctypes.c_uint64(0)
        _check(lib().fm_bulk_insert_device_noalloc(self._h, ctypes.c_void_p(d_keys.ptr),
                                                   ctypes.c_void_p(d_vals.ptr), count, ctypes.byref(out)))
        return int(out.value)

    def bulk_insert_device_uncounted(self, d_keys: DeviceBuffer, d_vals: DeviceBuffer, count: int):
        _check(lib().fm_bulk_insert_device_uncounted(self._h, ctypes.c_void_p(d_keys.ptr),
                                                     ctypes.c_void_p(d_vals.ptr), count))

    def bulk_remove_device(self, d_keys: DeviceBuffer, count: int) -> int:
        out = ctypes.c_uint64(0)
        _check(lib().fm_bulk_remove_device(self._h, ctypes.c_void_p(d_keys.ptr), count,
                                           ctypes.byref(out)))
        return int(out.value)

    def recount(self) -> int:
        out = ctypes.c_uint64(0)
        _check(lib().fm_recount(self._h, ctypes.byref(out)))
        return int(out.value)

    def synchronize(self):
        _check(lib().fm_synchronize(self._h))

    # -- logical slot array (test support) --
    def export_slots(self):
        cap = self.capacity()
        flags = np.zeros(cap, dtype=np.uint32)
        keys = np.zeros(cap * self.key_size, dtype=np.uint8)
        vals = np.zeros(cap * self.value_size, dtype=np.uint8)
        _check(lib().fm_export_slots(self._h, _ptr(flags), _ptr(keys), _ptr(vals)))
        return flags, keys, vals

    def import_slots(self, flags, keys, vals, length: int):
        f = np.ascontiguousarray(flags, dtype=np.uint32)
        k = np.ascontiguousarray(keys).view(np.uint8).reshape(-1)
        v = np.ascontiguousarray(vals).view(np.uint8).reshape(-1)
        _check(lib().fm_import_slots(self._h, _ptr(f), _ptr(k), _ptr(v), length))

    def copy_from(self, src: "FlashMap", structure_only: bool = False):
        """Asynchronous device copy of src's table into this map
        (structure_only: flags, keys and bitmaps; value bytes untouched)."""
        _check(lib().fm_copy_table(self._h, src._h, 1 if structure_only else 0))

    def profile(self, enable: bool = True):
        """Turns per-kernel event timing on/off (resets the totals)."""
        _check(lib().fm_profile_enable(self._h, 1 if enable else 0))

    def profile_only(self, kernel: str = ""):
        """Limits event timing to one kernel's launches ("" : every kernel)
        from this call on."""
        _check(lib().fm_profile_only(self._h, kernel.encode()))

    def profile_read(self) -> dict:
        """{kernel name: (launches, total device ms)} since profile()."""
        arr = (KernelStat * 64)()
        n = ctypes.c_uint32(0)
        _check(lib().fm_profile_read(self._h, arr, 64, ctypes.byref(n)))
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms)) for i in range(n.value)}

    def tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().fm_get_tuning(self._h, ctypes.byref(t)))
        return t

    def set_tuning(self, get_group=None, get_block=None, remove_group=None, insert_block=None,
                   insert_coop_arc=None, insert_fused=None, insert_dedup=None):
        t = self.tuning()
        for name, val in (("get_group", get_group), ("get_block", get_block),
                          ("remove_group", remove_group), ("insert_block", insert_block),
                          ("insert_coop_arc", insert_coop_arc), ("insert_fused", insert_fused),
                          ("insert_dedup", insert_dedup)):
            if val is not None:
                setattr(t, name, val)
        _check(lib().fm_set_tuning(self._h, ctypes.byref(t)))

    def autotune_get(self, d_keys: DeviceBuffer, count: int, reps: int = 5) -> dict:
        """Sets the fastest get_group x get_block for gets of these keys
        (fm_autotune_get); the table is not changed."""
        best, us = Tuning(), ctypes.c_double(0)
        _check(lib().fm_autotune_get(self._h, ctypes.c_void_p(d_keys.ptr), count, reps, ctypes.byref(best),
                                     ctypes.byref(us)))
        return {"get_group": best.get_group, "get_block": best.get_block, "get_us": round(us.value, 1)}


class ServeResult:
    """Outputs of one served request."""

    def __init__(self, op, n, value_size, arenas=None):
        # Outputs come from pinned arenas (values, found, flags; DMA writes
        # them directly, packed back to back so that adjacent requests take
        # one copy), or are allocated here with np.full, which writes every
        # page now: np.zeros would map untouched pages that the server's
        # result copy then faults in one by one inside the serving window.
        self.op = op
        self._arenas = arenas
        if arenas is not None:
            vals_a, found_a, flags_a = arenas
            self.values = vals_a.take((n, value_size), align=1) if op == OP_GET else None
            self.found = found_a.take(n, align=1) if op == OP_GET else None
            self.flags = flags_a.take(n, align=1) if op != OP_GET else None
        else:
            self.values = np.full((n, value_size), 0, dtype=np.uint8) if op == OP_GET else None
            self.found = np.full(n, 0, dtype=np.uint8) if op == OP_GET else None
            self.flags = np.full(n, 0, dtype=np.uint8) if op != OP_GET else None
        self.count = 0
        self.status = OK
        self.batch = -1
        self.arrival_s = self.admit_s = self.done_s = 0.0

    @property
    def latency_ms(self):
        return (self.done_s - self.arrival_s) * 1e3


class PinnedArena:
    """One pinned host allocation (fm_host_alloc) carved into numpy arrays.
    Serving requests whose buffers lie in it are moved by DMA, without the
    host copy through the staging ring."""

    def __init__(self, nbytes: int):
        self.nbytes = max(int(nbytes), 256)
        p = ctypes.c_void_p()
        _check(lib().fm_host_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value
        self.buf = np.frombuffer((ctypes.c_uint8 * self.nbytes).from_address(self.ptr), dtype=np.uint8)
        self.used = 0

    def take(self, shape, fill=0, align=256) -> np.ndarray:
        size = int(np.prod(shape))
        start = (self.used + align - 1) // align * align
        if start + size > self.nbytes:
            raise ValueError("pinned arena exhausted")
        self.used = start + size
        a = self.buf[start:start + size].reshape(shape)
        if fill is not None:
            a[...] = fill
        return a

    def copy_of(self, a: np.ndarray, align=256) -> np.ndarray:
        out = self.take(a.shape, fill=None, align=align)
        out[...] = a
        return out

    def close(self):
        if self.ptr:
            self.buf = None
            lib().fm_host_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_log(fmap: "FlashMap", requests, staging_bytes: int = 256 << 20, max_batch_keys: int = 1 << 20,
              realtime: bool = True, pinned=False):
    """Runs a request log through the continuous-batching server
    (fm_serve_log).  `requests`: objects with .op ("get"|"insert"|"remove"),
    .keys ((n, K) bytes), .vals ((n,) or (n, V); inserts) and .arrival_s.
    pinned: False (request buffers in pageable memory, copied through the
    staging ring), True (inputs copied into a pinned arena before serving
    starts and outputs allocated there, so the server moves both by DMA),
    "inputs" or "outputs" (only that side pinned).
    Returns ([ServeResult], ServeStats)."""
    ops = {"get": OP_GET, "insert": OP_INSERT, "remove": OP_REMOVE}
    n = len(requests)
    arr = (ServeRequest * max(n, 1))()
    keep, results = [], []
    pin_in = pinned in (True, "inputs")
    pin_out = pinned in (True, "outputs")
    K, V = fmap.key_size, fmap.value_size
    sizes = [_as_bytes(r.keys, K, "keys")[1] for r in requests]

    def field(pred, per):  # one pinned arena per field, requests packed in arrival order
        return PinnedArena(sum(nk * per for r, nk in zip(requests, sizes) if pred(r)))

    key_a = field(lambda r: True, K) if pin_in else None
    val_a = field(lambda r: r.op == "insert", V) if pin_in else None
    outs = (field(lambda r: r.op == "get", V), field(lambda r: r.op == "get", 1),
            field(lambda r: r.op != "get", 1)) if pin_out else None
    for i, r in enumerate(requests):
        op = ops[r.op]
        k, nk = _as_bytes(r.keys, K, "keys")
        if pin_in:
            k = key_a.copy_of(k, align=1)
        res = ServeResult(op, nk, V, outs)
        res.arrival_s = float(r.arrival_s) if realtime else 0.0
        a = arr[i]
        a.op, a.n, a.keys, a.arrival_s = op, nk, k.ctypes.data if nk else None, float(r.arrival_s)
        keep.append(k)
        if op == OP_INSERT:
            v, nv = _as_bytes(r.vals, fmap.value_size, "values")
            assert nv == nk
            if pin_in:
                v = val_a.copy_of(v, align=1)
            a.values = v.ctypes.data if nv else None
            keep.append(v)
        if op == OP_GET:
            a.out_values = res.values.ctypes.data if nk else None
            a.out_found = res.found.ctypes.data if nk else None
        else:
            a.out_flags = res.flags.ctypes.data if nk else None
        results.append(res)
    cfg = ServeConfig(staging_bytes, max_batch_keys, 1 if realtime else 0, 0)
    stats = ServeStats()
    _check(lib().fm_serve_log(fmap._h, arr, n, ctypes.byref(cfg), ctypes.byref(stats)))
    for i, res in enumerate(results):
        a = arr[i]
        res.count, res.status, res.batch = int(a.count), int(a.status), int(a.batch)
        res.admit_s, res.done_s = float(a.admit_s), float(a.done_s)
    del keep  # pinned input arenas go with this frame; output arenas live on in the results
    return results, stats


class GroupBuffer:
    """A device allocation owned by a FlashMapGroup."""

    def __init__(self, owner: "FlashMapGroup", ptr: int, nbytes: int):
        self.owner, self.ptr, self.nbytes = owner, ptr, nbytes

    def free(self):
        if self.ptr and self.owner._h:
            lib().fm_group_free_device(self.owner._h, ctypes.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _u64s(xs) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(xs, dtype=np.uint64))


class FlashMapGroup:
    """n independent FlashMaps on one GPU driven by fused launches
    (fm_group_* in include/flashmap.h).  Batches are concatenated in table
    order; ``counts[t]`` keys belong to table t."""

    def __init__(self, capacities, key_size: int, value_size: int, hash_strategy: int = 0, device_id: int = 0):
        caps = _u64s(capacities)
        h = ctypes.c_void_p()
        self._h = None
        _check(lib().fm_group_create(caps.size, caps.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), key_size,
                                     value_size, hash_strategy, device_id, ctypes.byref(h)))
        self._h = h
        self.n = caps.size
        self.key_size, self.value_size = key_size, value_size

    def close(self):
        if self._h:
            lib().fm_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def len(self, t: int) -> int:
        return int(lib().fm_group_len(self._h, t))

    def capacity(self, t: int) -> int:
        return int(lib().fm_group_capacity(self._h, t))

    def stream(self) -> int:
        return int(lib().fm_group_stream(self._h) or 0)

    def alloc_device(self, nbytes: int) -> GroupBuffer:
        p = ctypes.c_void_p()
        _check(lib().fm_group_alloc_device(self._h, nbytes, ctypes.byref(p)))
        return GroupBuffer(self, p.value or 0, nbytes)

    def upload(self, data) -> GroupBuffer:
        arr, _ = _as_bytes(data, 1, "data")
        p = ctypes.c_void_p()
        _check(lib().fm_group_upload(self._h, _ptr(arr), arr.size, ctypes.byref(p)))
        return GroupBuffer(self, p.value or 0, arr.size)

    def download(self, buf: GroupBuffer, nbytes: Optional[int] = None) -> np.ndarray:
        nbytes = buf.nbytes if nbytes is None else nbytes
        out = np.zeros(nbytes, dtype=np.uint8)
        _check(lib().fm_group_download(self._h, ctypes.c_void_p(buf.ptr), nbytes, _ptr(out)))
        return out

    @staticmethod
    def _p64(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))

    def bulk_insert_device(self, d_keys: GroupBuffer, d_vals: GroupBuffer, counts):
        c = _u64s(counts)
        new = np.zeros(self.n, dtype=np.uint64)
        status = np.zeros(self.n, dtype=np.int32)
        _check(lib().fm_group_bulk_insert_device(self._h, ctypes.c_void_p(d_keys.ptr), ctypes.c_void_p(d_vals.ptr),
                                                 self._p64(c), self._p64(new),
                                                 status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return new, status

    def bulk_get_device(self, d_keys: GroupBuffer, counts, d_out: GroupBuffer, d_found: GroupBuffer):
        c = _u64s(counts)
        _check(lib().fm_group_bulk_get_device(self._h, ctypes.c_void_p(d_keys.ptr), self._p64(c),
                                              ctypes.c_void_p(d_out.ptr), ctypes.c_void_p(d_found.ptr)))

    def bulk_remove_device(self, d_keys: GroupBuffer, counts):
        c = _u64s(counts)
        rem = np.zeros(self.n, dtype=np.uint64)
        _check(lib().fm_group_bulk_remove_device(self._h, ctypes.c_void_p(d_keys.ptr), self._p64(c), self._p64(rem)))
        return rem

    def step_device(self, ins_keys, ins_vals, ins_counts, get_keys, get_counts, d_out, d_found, rem_keys,
                    rem_counts):
        ic, gc, rc = _u64s(ins_counts), _u64s(get_counts), _u64s(rem_counts)
        new = np.zeros(self.n, dtype=np.uint64)
        status = np.zeros(self.n, dtype=np.int32)
        rem = np.zeros(self.n, dtype=np.uint64)
        vp = ctypes.c_void_p
        _check(lib().fm_group_step_device(self._h, vp(ins_keys.ptr), vp(ins_vals.ptr), self._p64(ic),
                                          vp(get_keys.ptr), self._p64(gc), vp(d_out.ptr), vp(d_found.ptr),
                                          vp(rem_keys.ptr), self._p64(rc), self._p64(new),
                                          status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), self._p64(rem)))
        return new, status, rem

    def export_slots(self, t: int):
        cap = self.capacity(t)
        flags = np.zeros(cap, dtype=np.uint32)
        keys = np.zeros(cap * self.key_size, dtype=np.uint8)
        vals = np.zeros(cap * self.value_size, dtype=np.uint8)
        _check(lib().fm_group_export_slots(self._h, t, _ptr(flags), _ptr(keys), _ptr(vals)))
        return flags, keys, vals

    def copy_from(self, src: "FlashMapGroup"):
        _check(lib().fm_group_copy(self._h, src._h))

    def tuning(self) -> GroupTuning:
        t = GroupTuning()
        _check(lib().fm_group_get_tuning(self._h, ctypes.byref(t)))
        return t

    def set_tuning(self, group=None, block=None, grid=None, coop_arc=None):
        t = self.tuning()
        for name, val in (("group", group), ("block", block), ("grid", grid), ("coop_arc", coop_arc)):
            if val is not None:
                setattr(t, name, val)
        _check(lib().fm_group_set_tuning(self._h, ctypes.byref(t)))

    def tuning_key(self, ins_counts, get_counts, rem_counts) -> str:
        """Cache key of a step shape: sizes, table count, mean load (0.1
        steps) and keys per step (powers of two)."""
        caps = np.array([self.capacity(t) for t in range(self.n)], dtype=np.float64)
        lens = np.array([self.len(t) for t in range(self.n)], dtype=np.float64)
        keys = int(_u64s(ins_counts).sum() + _u64s(get_counts).sum() + _u64s(rem_counts).sum())
        load = float((lens / caps).mean())
        return f"K{self.key_size}-V{self.value_size}-T{self.n}-load{load:.1f}-keys2^{max(0, keys).bit_length()}"

    def autotune(self, ins_keys, ins_vals, ins_counts, get_keys, get_counts, d_out, d_found, rem_keys,
                 rem_counts, reps: int = 3, candidates=None, scratch: "FlashMapGroup" = None,
                 cache: Optional[str] = None) -> dict:
        """Picks the fastest launch shape for this step (fm_group_autotune)
        and sets it.  ``cache``: a JSON file of tunings by step shape
        (tuning_key); a hit sets the cached tuning without measuring, a miss
        measures and records the result.  The tables are left unchanged."""
        import json

        key = self.tuning_key(ins_counts, get_counts, rem_counts)
        if cache and os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
            if key in known:
                self.set_tuning(**known[key]["tuning"])
                return dict(known[key], key=key, cached=True)
        ic, gc, rc = _u64s(ins_counts), _u64s(get_counts), _u64s(rem_counts)
        cand = None
        if candidates:
            cand = (GroupTuning * len(candidates))()
            for i, c in enumerate(candidates):
                cand[i] = GroupTuning(c["group"], c["block"], c["grid"], c.get("coop_arc", 64))
        best, us, ran = GroupTuning(), ctypes.c_double(0), ctypes.c_uint32(0)
        vp = ctypes.c_void_p
        _check(lib().fm_group_autotune(self._h, scratch._h if scratch else None, vp(ins_keys.ptr), vp(ins_vals.ptr),
                                       self._p64(ic), vp(get_keys.ptr), self._p64(gc), vp(d_out.ptr),
                                       vp(d_found.ptr), vp(rem_keys.ptr), self._p64(rc), cand,
                                       len(candidates) if candidates else 0, reps, ctypes.byref(best),
                                       ctypes.byref(us), ctypes.byref(ran)))
        rec = {"tuning": {k: getattr(best, k) for k in ("group", "block", "grid", "coop_arc")},
               "step_us": round(us.value, 1), "evaluated": int(ran.value)}
        if cache:
            known = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    known = json.load(f)
            known[key] = rec
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            tmp = cache + ".tmp"
            with open(tmp, "w") as f:
                json.dump(known, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return dict(rec, key=key, cached=False)


class FlashMapBuilder:
    """src/lib.rs:377-449."""

    def __init__(self, capacity: int, key_size: int = 8, value_size: int = 8):
        self._b = _Builder()
        lib().fm_builder_init(ctypes.byref(self._b), capacity, key_size, value_size)

    def hash_strategy(self, strategy: int) -> "FlashMapBuilder":
        self._b.hash_strategy = strategy
        return self

    def device_id(self, dev: int) -> "FlashMapBuilder":
        self._b.device_id = dev
        return self

    def force_cpu(self) -> "FlashMapBuilder":
        self._b.force_cpu = 1
        return self

    def require_gpu(self) -> "FlashMapBuilder":
        self._b.require_gpu = 1
        return self

    def quiet(self) -> "FlashMapBuilder":
        self._b.quiet = 1
        return self

    def build(self) -> FlashMap:
        h = ctypes.c_void_p()
        _check(lib().fm_build(ctypes.byref(self._b), ctypes.byref(h)))
        return FlashMap(h.value, self._b.key_size, self._b.value_size)


class _RwLock:
    """Reader-writer lock (std::sync::RwLock in src/async_map.rs:33)."""

    def __init__(self):
        self._cv = threading.Condition()
        self._readers = 0
        self._writer = False

    def acquire_read(self):
        with self._cv:
            while self._writer:
                self._cv.wait()
            self._readers += 1

    def release_read(self):
        with self._cv:
            self._readers -= 1
            self._cv.notify_all()

    def acquire_write(self):
        with self._cv:
            while self._writer or self._readers:
                self._cv.wait()
            self._writer = True

    def release_write(self):
        with self._cv:
            self._writer = False
            self._cv.notify_all()


class AsyncFlashMap:
    """src/async_map.rs:32-147: every op runs in a worker thread
    (spawn_blocking); gets share a read lock, mutations take the write lock."""

    def __init__(self, fmap: FlashMap):
        self._map = fmap
        self._lock = _RwLock()

    def _read(self, fn, *a):
        self._lock.acquire_read()
        try:
            return fn(*a)
        finally:
            self._lock.release_read()

    def _write(self, fn, *a):
        self._lock.acquire_write()
        try:
            return fn(*a)
        finally:
            self._lock.release_write()

    async def bulk_get(self, keys):
        return await asyncio.get_running_loop().run_in_executor(None, self._read, self._map.bulk_get, keys)

    async def bulk_insert(self, keys, values) -> int:
        return await asyncio.get_running_loop().run_in_executor(
            None, self._write, self._map.bulk_insert, keys, values)

    async def bulk_remove(self, keys) -> int:
        return await asyncio.get_running_loop().run_in_executor(None, self._write, self._map.bulk_remove, keys)

    async def clear(self):
        return await asyncio.get_running_loop().run_in_executor(None, self._write, self._map.clear)

    def len(self) -> int:
        return self._read(len, self._map)

    def is_empty(self) -> bool:
        return self.len() == 0

    def capacity(self) -> int:
        return self._read(self._map.capacity)

    def load_factor(self) -> float:
        return self._read(self._map.load_factor)

    def clone(self) -> "AsyncFlashMap":
        c = AsyncFlashMap.__new__(AsyncFlashMap)
        c._map, c._lock = self._map, self._lock
        return c
