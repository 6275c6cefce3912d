"""Batched filtered GROUP BY queries over one device-resident table (qk_batch_* in
include/qikkdb_amd.h).

Each query is ``SELECT key, SUM(value), MIN(value), MAX(value), COUNT(value) FROM t
WHERE ts_lo <= ts < ts_hi GROUP BY key`` over a contiguous range of the table's blocks, with
``key = (a * b_range + b) mod key_mod``.  All queries of a run are executed by one
persistent kernel that takes (tile, query) work units from device queues, one per die.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _abi
from ._abi import BatchColumns, BatchQuery, BatchTuning, call
from .memory import ensure_device


@dataclass
class Query:
    first_block: int
    nblocks: int
    ts_lo: int
    ts_hi: int
    key_mod: int

    def as_c(self) -> BatchQuery:
        q = BatchQuery()
        q.first_block, q.nblocks = self.first_block, self.nblocks
        q.ts_lo, q.ts_hi, q.key_mod, q.reserved = self.ts_lo, self.ts_hi, self.key_mod, 0
        return q


@dataclass
class Tuning:
    tile_rows: int = 1 << 18
    lds_groups: int = 4096
    wgs_per_cu: int = 1
    use_graph: int = 1
    scratch_rows: int = 0


def _queries(qs: Sequence[Query]):
    arr = (BatchQuery * max(1, len(qs)))(*[q.as_c() for q in qs])
    return arr, len(qs)


class BatchEngine:
    """One table (device pointers to int32 a, int32 b, int64 ts, int64 value columns)."""

    def __init__(self, a, b, ts, value, rows: int, block_rows: int, b_range: int):
        ensure_device()
        cols = BatchColumns()
        cols.a, cols.b, cols.ts, cols.value = int(a), int(b), int(ts), int(value)
        cols.rows, cols.block_rows, cols.b_range = rows, block_rows, b_range
        self.cols = cols
        self.h = C.c_void_p()
        call("qk_batch_create", C.byref(self.h), C.byref(cols))
        self.last = 0

    def close(self) -> None:
        if self.h:
            call("qk_batch_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, t: Tuning) -> None:
        call("qk_batch_set_tuning", self.h,
             C.byref(BatchTuning(t.tile_rows, t.lds_groups, t.wgs_per_cu, t.use_graph, t.scratch_rows)))

    def tuning(self) -> Tuning:
        t = BatchTuning()
        call("qk_batch_get_tuning", self.h, C.byref(t))
        return Tuning(t.tile_rows, t.lds_groups, t.wgs_per_cu, t.use_graph, t.scratch_rows)

    def autotune(self, qs: Sequence[Query], stream=None) -> Tuning:
        arr, n = _queries(qs)
        t = BatchTuning()
        call("qk_batch_autotune", self.h, n, arr, C.byref(t), stream)
        return Tuning(t.tile_rows, t.lds_groups, t.wgs_per_cu, t.use_graph, t.scratch_rows)

    def run(self, qs: Sequence[Query], stream=None) -> None:
        arr, n = _queries(qs)
        call("qk_batch_run", self.h, n, arr, stream)
        self.last = n

    def estimate_rows(self, qs: Sequence[Query]) -> np.ndarray:
        """Estimated passing rows of each query, from the ts samples taken at creation."""
        arr, n = _queries(qs)
        out = np.zeros(max(1, n), np.int64)
        call("qk_batch_estimate_rows", self.h, n, arr, out.ctypes.data_as(C.POINTER(C.c_int64)))
        return out[:n]

    def completion_ms(self) -> np.ndarray:
        out = np.zeros(self.last, np.float64)
        call("qk_batch_completion_ms", self.h, out.ctypes.data_as(C.POINTER(C.c_double)), self.last)
        return out

    def results(self) -> tuple:
        """(count, sum, min, max) device pointers and the host offsets array of the last run."""
        p = [C.c_void_p() for _ in range(5)]
        call("qk_batch_results", self.h, *[C.byref(x) for x in p])
        offs = np.ctypeslib.as_array(C.cast(p[4], C.POINTER(C.c_int64)), shape=(self.last + 1,)).copy()
        return p[0].value, p[1].value, p[2].value, p[3].value, offs

    def result_arrays(self, i: int) -> list:
        """Dense result arrays of query i of the last run, [count, sum, min, max], each of
        key_mod int64 entries (group g exists when count[g] != 0)."""
        cnt, sm, mn, mx, offs = self.results()
        lo, hi = int(offs[i]), int(offs[i + 1])
        n = hi - lo
        arrs = []
        for p in (cnt, sm, mn, mx):
            a = np.empty(n, np.int64)
            if n:
                call("qk_memcpy_dtoh", a.ctypes.data, p + lo * 8, n * 8, None)
            arrs.append(a)
        return arrs

    def result_host(self, i: int) -> dict:
        """Groups of query i of the last run: key -> (sum, min, max, count)."""
        c, s, lo_v, hi_v = self.result_arrays(i)
        keys = np.nonzero(c)[0]
        return {int(k): (int(s[k]), int(lo_v[k]), int(hi_v[k]), int(c[k])) for k in keys}


def key_column(a, b, b_range: int, key_mod: int, n: int, out, stream=None) -> None:
    """out[i] = (a[i] * b_range + b[i]) mod key_mod as int64 (device pointers)."""
    call("qk_batch_key_column", int(a), int(b), b_range, key_mod, n, int(out), stream)


__all__ = ["BatchEngine", "Query", "Tuning", "key_column"]
_ = _abi
