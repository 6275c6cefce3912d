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
