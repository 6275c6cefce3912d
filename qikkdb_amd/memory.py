"""Device buffers over the library's pooled allocator (qk_malloc / qk_free).

Counterpart of the reference's GPUMemory helpers (QueryEngine/GPUCore/GPUMemory.cuh:89-189):
alloc / free / copies, plus numpy conversions and LSB-first null bit masks
(QueryEngine/GPUCore/NullValues.cu:3-57).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import call

_device_ready = False


def ensure_device(device: int = 0) -> None:
    """Bind the calling thread to `device` (qk_init)."""
    global _device_ready
    if not _device_ready:
        call("qk_init", device)
        _device_ready = True


class DeviceBuffer:
    """A device allocation.  Owned buffers are returned to the pool when freed/collected."""

    __slots__ = ("ptr", "nbytes", "owned")

    def __init__(self, nbytes: int = 0, ptr: int | None = None, owned: bool = True):
        if ptr is None:
            ensure_device()
            p = C.c_void_p()
            call("qk_malloc", C.byref(p), max(1, int(nbytes)))
            ptr = p.value
        self.ptr = ptr
        self.nbytes = int(nbytes)
        self.owned = owned

    @classmethod
    def from_numpy(cls, arr: np.ndarray, stream=None) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        if arr.nbytes:
            call("qk_memcpy_htod", buf.ptr, arr.ctypes.data, arr.nbytes, stream)
        return buf

    @classmethod
    def zeros(cls, nbytes: int, stream=None) -> "DeviceBuffer":
        buf = cls(nbytes)
        call("qk_memset", buf.ptr, 0, max(1, nbytes), stream)
        return buf

    def to_numpy(self, dtype, count: int | None = None, stream=None) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if count:
            call("qk_memcpy_dtoh", out.ctypes.data, self.ptr, out.nbytes, stream)
        return out

    def free(self) -> None:
        if self.owned and self.ptr:
            call("qk_free", self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __int__(self) -> int:
        return self.ptr or 0


def ptr_of(x) -> int | None:
    """Raw device address of a DeviceBuffer / int / None."""
    if x is None:
        return None
    if isinstance(x, DeviceBuffer):
        return x.ptr
    return int(x)


# ---- null bit masks ---------------------------------------------------------------------

def pack_nulls(isnull: np.ndarray) -> np.ndarray:
    """bool per row -> uint64 words, bit (i % 64) of word i // 64 (NullValues.cu:3-57)."""
    isnull = np.asarray(isnull, dtype=bool)
    n = isnull.size
    words = (n + 63) // 64
    padded = np.zeros(words * 64, dtype=bool)
    padded[:n] = isnull
    return np.packbits(padded, bitorder="little").view(np.uint64)


def unpack_nulls(words: np.ndarray, n: int) -> np.ndarray:
    words = np.ascontiguousarray(words, dtype=np.uint64)
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)


class DeviceColumn:
    """A typed device column of n rows with an optional null bit mask."""

    def __init__(self, data: DeviceBuffer, qtype: int, n: int, nulls: DeviceBuffer | None = None):
        self.data, self.qtype, self.n, self.nulls = data, qtype, n, nulls

    @classmethod
    def from_numpy(cls, arr: np.ndarray, isnull: np.ndarray | None = None) -> "DeviceColumn":
        arr = np.ascontiguousarray(arr)
        qtype = _abi.TYPE_OF[arr.dtype]
        nulls = DeviceBuffer.from_numpy(pack_nulls(isnull)) if isnull is not None else None
        return cls(DeviceBuffer.from_numpy(arr), qtype, arr.size, nulls)

    @property
    def ptr(self) -> int:
        return self.data.ptr

    @property
    def null_ptr(self) -> int | None:
        return self.nulls.ptr if self.nulls is not None else None

    def to_numpy(self) -> np.ndarray:
        return self.data.to_numpy(_abi.NP_OF[self.qtype], self.n)


# ---- pool cap and block cache ---------------------------------------------------------------

def set_memory_limit(nbytes: int) -> None:
    """Cap the device memory the pool holds (0 = none): qk_mem_set_limit, the counterpart of
    the reference allocator's GPUMemoryUsagePercent (CudaMemAllocator.cu:109-152)."""
    ensure_device()
    call("qk_mem_set_limit", int(nbytes))


def set_memory_limit_percent(percent: int) -> None:
    ensure_device()
    call("qk_mem_set_limit_percent", int(percent))


def memory_limit() -> int:
    out = C.c_size_t()
    call("qk_mem_get_limit", C.byref(out))
    return out.value


def can_allocate(nbytes: int) -> bool:
    """CudaMemAllocator::CanAllocate (CudaMemAllocator.h:57)."""
    out = C.c_int32()
    call("qk_mem_can_allocate", int(nbytes), C.byref(out))
    return bool(out.value)


def memory_stats() -> tuple[int, int]:
    """(bytes in use, bytes cached) of the pool."""
    u, c = C.c_size_t(), C.c_size_t()
    call("qk_mem_stats", C.byref(u), C.byref(c))
    return u.value, c.value


class BlockCache:
    """LRU cache of device column blocks (qk_cache_*): the reference's GPUMemoryCache
    (QueryEngine/GPUMemoryCache.h:10-141).  get_column returns (DeviceBuffer not owned by the
    caller, bytes, hit)."""

    def __init__(self, max_bytes: int):
        ensure_device()
        self.h = C.c_void_p()
        call("qk_cache_create", C.byref(self.h), int(max_bytes))

    def close(self) -> None:
        if self.h:
            call("qk_cache_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_column(self, database: str, table_column: str, block: int, nbytes: int,
                   load_size: int = -1, load_offset: int = 0) -> tuple[DeviceBuffer, int, bool]:
        p, got, hit = C.c_void_p(), C.c_size_t(), C.c_int32()
        call("qk_cache_get_column", self.h, database.encode(), table_column.encode(), block, int(nbytes),
             load_size, load_offset, C.byref(p), C.byref(got), C.byref(hit))
        return DeviceBuffer(got.value, ptr=p.value, owned=False), got.value, bool(hit.value)

    def contains(self, database: str, table_column: str, block: int, load_size: int = -1,
                 load_offset: int = 0) -> bool:
        out = C.c_int32()
        call("qk_cache_contains", self.h, database.encode(), table_column.encode(), block, load_size,
             load_offset, C.byref(out))
        return bool(out.value)

    def evict(self) -> bool:
        out = C.c_int32()
        call("qk_cache_evict", self.h, C.byref(out))
        return bool(out.value)

    def clear_block(self, database: str, table_column: str, block: int) -> None:
        call("qk_cache_clear_block", self.h, database.encode(), table_column.encode(), block)

    def set_current_block(self, block: int) -> None:
        call("qk_cache_set_current_block", self.h, block)

    @staticmethod
    def set_lock_list(columns) -> None:
        arr = (C.c_char_p * max(1, len(columns)))(*[c.encode() for c in columns])
        call("qk_cache_set_lock_list", len(columns), arr)

    def stats(self) -> dict:
        used, ent, hits, miss, ev = C.c_size_t(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        call("qk_cache_stats", self.h, C.byref(used), C.byref(ent), C.byref(hits), C.byref(miss), C.byref(ev))
        return {"used": used.value, "entries": ent.value, "hits": hits.value, "misses": miss.value,
                "evictions": ev.value}
