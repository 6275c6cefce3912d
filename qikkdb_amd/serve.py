"""Continuous batching of batched-form queries over one device-resident table (qk_serve_* in
include/qikkdb_amd.h).

Requests are submitted one at a time from any thread; the library's dispatcher thread runs
the queued ones as batches on a small set of batch engines (one stream each), admitting them
by projected table bytes under a memory cap, and starts a new batch while the last one is
still on the device.  A done request's groups stay in device memory until it is released.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from ._abi import BatchColumns, BatchTuning, ServeConfig, ServeStats, call
from .batch import Query, Tuning
from .memory import ensure_device


@dataclass
class Config:
    engines: int = 0          # 0 = 2
    max_batch: int = 0        # 0 = 4096
    memory_cap: int = 0       # bytes, 0 = 200 GB
    wait_ms: float = 0.0      # 0 = half the last batch's time
    tuning: Tuning | None = None  # None = engine defaults without graph replay

    def as_c(self) -> ServeConfig:
        c = ServeConfig()
        c.engines, c.max_batch, c.memory_cap, c.wait_ms = self.engines, self.max_batch, self.memory_cap, self.wait_ms
        t = self.tuning
        if t is not None:
            c.tuning = BatchTuning(t.tile_rows, t.lds_groups, t.wgs_per_cu, t.use_graph, t.scratch_rows)
        return c


class Server:
    """Serving loop over one table (device pointers to int32 a, int32 b, int64 ts, int64 value)."""

    def __init__(self, a, b, ts, value, rows: int, block_rows: int, b_range: int, config: Config | None = None):
        ensure_device()
        cols = BatchColumns()
        cols.a, cols.b, cols.ts, cols.value = int(a), int(b), int(ts), int(value)
        cols.rows, cols.block_rows, cols.b_range = rows, block_rows, b_range
        self.h = C.c_void_p()
        cfg = (config or Config()).as_c()
        call("qk_serve_create", C.byref(self.h), C.byref(cols), C.byref(cfg))

    def close(self) -> None:
        if self.h:
            call("qk_serve_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, q: Query) -> int:
        t = C.c_int64()
        call("qk_serve_submit", self.h, C.byref(q.as_c()), C.byref(t))
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> None:
        call("qk_serve_wait", self.h, ticket, timeout_ms)

    def latency_ms(self, ticket: int) -> float:
        x = C.c_double()
        call("qk_serve_latency_ms", self.h, ticket, C.byref(x))
        return x.value

    def result_pointers(self, ticket: int) -> tuple:
        p = [C.c_void_p() for _ in range(4)]
        call("qk_serve_result", self.h, ticket, *[C.byref(x) for x in p])
        return tuple(x.value for x in p)

    def result_host(self, ticket: int, key_mod: int) -> dict:
        """Groups of a done request: key -> (sum, min, max, count)."""
        arrs = []
        for ptr in self.result_pointers(ticket):
            a = np.empty(key_mod, np.int64)
            call("qk_memcpy_dtoh", a.ctypes.data, ptr, key_mod * 8, None)
            arrs.append(a)
        c, s, lo, hi = arrs
        keys = np.nonzero(c)[0]
        return {int(k): (int(s[k]), int(lo[k]), int(hi[k]), int(c[k])) for k in keys}

    def release(self, ticket: int) -> None:
        call("qk_serve_release", self.h, ticket)

    def replay(self, qs: Sequence[Query], arrival_ms: np.ndarray) -> np.ndarray:
        """Open-loop replay; returns each request's latency (scheduled arrival to done) in ms."""
        from ._abi import BatchQuery

        n = len(qs)
        arr = (BatchQuery * max(1, n))(*[q.as_c() for q in qs])
        arrival = np.ascontiguousarray(arrival_ms, np.float64)
        lat = np.zeros(n, np.float64)
        call("qk_serve_replay", self.h, n, arr, arrival.ctypes.data_as(C.POINTER(C.c_double)),
             lat.ctypes.data_as(C.POINTER(C.c_double)))
        return lat

    def stats(self) -> dict:
        s = ServeStats()
        call("qk_serve_get_stats", self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServeStats._fields_}


def replay_plain(a, b, ts, value, rows: int, block_rows: int, b_range: int, qs: Sequence[Query],
                 arrival_ms: np.ndarray, stream=None) -> np.ndarray:
    """The plain one-request-at-a-time path under the same open-loop arrivals
    (qk_serve_replay_plain): each request's latency, arrival to groups in device memory, in ms."""
    from ._abi import BatchQuery

    ensure_device()
    cols = BatchColumns()
    cols.a, cols.b, cols.ts, cols.value = int(a), int(b), int(ts), int(value)
    cols.rows, cols.block_rows, cols.b_range = rows, block_rows, b_range
    n = len(qs)
    arr = (BatchQuery * max(1, n))(*[q.as_c() for q in qs])
    arrival = np.ascontiguousarray(arrival_ms, np.float64)
    lat = np.zeros(n, np.float64)
    call("qk_serve_replay_plain", C.byref(cols), n, arr, arrival.ctypes.data_as(C.POINTER(C.c_double)),
         lat.ctypes.data_as(C.POINTER(C.c_double)), stream)
    return lat


__all__ = ["Config", "Server", "replay_plain"]
