"""String key columns and GROUP BY over them (qk_gbs_* in include/qikkdb_amd.h).

A device string column is the reference's GPUString (QueryEngine/GPUCore/GPUMemory.cuh:70-77):
every string's bytes concatenated (`chars`) and `indices[i]` = the end offset of string i.
GPUGroupByString mirrors GPUGroupBy<AGG,O,std::string,V> (GPUGroupByString.cuh:169-785).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _abi
from ._abi import Predicate, call
from .memory import DeviceBuffer, ensure_device, unpack_nulls


def _as_bytes(s) -> bytes:
    return s.encode() if isinstance(s, str) else bytes(s)


class DeviceStrings:
    """A device string column: chars (uint8) and indices (int64 end offsets)."""

    def __init__(self, chars: DeviceBuffer, indices: DeviceBuffer, n: int, nbytes: int):
        self.chars, self.indices, self.n, self.nbytes = chars, indices, n, nbytes

    @classmethod
    def from_list(cls, strings: Sequence) -> "DeviceStrings":
        raw = [_as_bytes(s) for s in strings]
        lens = np.array([len(b) for b in raw], np.int64)
        idx = np.cumsum(lens).astype(np.int64) if len(raw) else np.zeros(0, np.int64)
        blob = np.frombuffer(b"".join(raw), np.uint8) if raw else np.zeros(0, np.uint8)
        chars = DeviceBuffer.from_numpy(blob if blob.size else np.zeros(1, np.uint8))
        return cls(chars, DeviceBuffer.from_numpy(idx if idx.size else np.zeros(1, np.int64)), len(raw),
                   int(blob.size))

    @staticmethod
    def decode(chars_ptr: int, idx_ptr: int, n: int) -> list:
        """Host list of bytes from device chars / indices of n strings."""
        if n == 0:
            return []
        idx = DeviceBuffer(n * 8, ptr=idx_ptr, owned=False).to_numpy(np.int64, n)
        total = int(idx[-1])
        blob = DeviceBuffer(max(total, 1), ptr=chars_ptr, owned=False).to_numpy(np.uint8, total).tobytes()
        starts = np.concatenate([[0], idx[:-1]])
        return [blob[int(a):int(b)] for a, b in zip(starts, idx)]


class GPUGroupByString:
    """GROUP BY over one string key column; `aggs` one aggregate or a list of them."""

    def __init__(self, aggs, value_type: int, max_hash_count: int, key_size_multiplier: int = 1,
                 out_types=None, flags: int = 0):
        ensure_device()
        self.aggs = [aggs] if isinstance(aggs, int) else list(aggs)
        self.value_type = value_type
        na = len(self.aggs)
        if out_types is None:
            out_types = [0] * na
        self.out_types = [
            (t if t else (_abi.LONG if a in (_abi.AGG_COUNT, _abi.AGG_COUNT_ALL) else value_type))
            for a, t in zip(self.aggs, out_types)]
        h = C.c_void_p()
        call("qk_gbs_create", C.byref(h), value_type, na, (C.c_int32 * na)(*self.aggs),
             (C.c_int32 * na)(*out_types), max_hash_count, key_size_multiplier, flags)
        self.h = h.value

    def close(self) -> None:
        if getattr(self, "h", None):
            _abi.load().qk_gbs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, stream=None) -> None:
        call("qk_gbs_reset", self.h, stream)

    def ProcessBlock(self, keys: DeviceStrings, values, n: int, key_nulls=None, value_nulls=None, stream=None):
        call("qk_gbs_process_block", self.h, keys.chars.ptr, keys.indices.ptr, _ptr(key_nulls), _ptr(values),
             _ptr(value_nulls), n, stream)

    def ProcessBlockWhere(self, keys: DeviceStrings, values, n: int, preds: Sequence[Predicate], key_nulls=None,
                          value_nulls=None, stream=None):
        arr = (Predicate * max(1, len(preds)))(*preds)
        call("qk_gbs_process_block_where", self.h, keys.chars.ptr, keys.indices.ptr, _ptr(key_nulls),
             _ptr(values), _ptr(value_nulls), n, len(preds), arr, stream)

    def group_count(self, stream=None) -> int:
        v = C.c_int64()
        call("qk_gbs_group_count", self.h, C.byref(v), stream)
        return v.value

    def results_host(self, stream=None) -> dict:
        """{'count', 'keys': [bytes or None], 'values': [np arrays], 'value_nulls': [bool arrays]}."""
        na = len(self.aggs)
        chars, idx, kn = C.c_void_p(), C.c_void_p(), C.c_void_p()
        vals, vnulls = (C.c_void_p * na)(), (C.c_void_p * na)()
        g = C.c_int32()
        call("qk_gbs_get_results", self.h, C.byref(chars), C.byref(idx), C.byref(kn), vals, vnulls, C.byref(g),
             stream)
        n = g.value
        owned = [p for p in (chars.value, idx.value, kn.value) if p] + [p for p in list(vals) + list(vnulls) if p]
        try:
            keys = DeviceStrings.decode(chars.value, idx.value, n)
            knull = (unpack_nulls(DeviceBuffer((n + 63) // 64 * 8, ptr=kn.value, owned=False).to_numpy(np.uint64), n)
                     if kn.value and n else np.zeros(n, bool))
            keys = [None if knull[i] else keys[i] for i in range(n)]
            out_vals, out_nulls = [], []
            for a in range(na):
                if not vals[a]:
                    out_vals.append(None)
                    out_nulls.append(None)
                    continue
                dt = _abi.NP_OF[self.out_types[a]]
                out_vals.append(DeviceBuffer(n * np.dtype(dt).itemsize, ptr=vals[a], owned=False).to_numpy(dt, n)
                                if n else np.empty(0, dt))
                out_nulls.append(unpack_nulls(DeviceBuffer((n + 63) // 64 * 8, ptr=vnulls[a], owned=False)
                                              .to_numpy(np.uint64), n) if (vnulls[a] and n) else np.zeros(n, bool))
        finally:
            for p in owned:
                call("qk_free", p)
        return {"count": n, "keys": keys, "values": out_vals, "value_nulls": out_nulls}


def _ptr(x):
    if x is None:
        return None
    return x.ptr if hasattr(x, "ptr") else int(x)


__all__ = ["DeviceStrings", "GPUGroupByString"]
