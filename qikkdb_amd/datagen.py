"""Seeded synthetic workloads (host generation through libqk_datagen.so).

Every column is a pure function of (seed, row), so slices regenerate independently.  The
workloads are the ones the benchmark and parity tests run:

* ``small``      - 1M rows, key int32 U[0,1000), value int32 U[-1000,1000]; WHERE value > 0;
                   GROUP BY key: SUM (int64 accumulator) / MIN / MAX / COUNT.
* ``zipf``       - key int64 Zipf(s=1.1) over 1e6 distinct values, value float32 lognormal(0,1),
                   5% NULL values; WHERE value BETWEEN 0.5 AND 4.0 (= value >= 0.5f AND
                   value <= 4.0f); SUM / AVG / COUNT.  256M rows in 2^26-row blocks.
* ``composite``  - key (a, b) int32 U[0,4096) each, ts int64 epoch-ms uniform over one year with
                   WHERE on a 30-day window, value int64 U[-1e9, 1e9]; SUM/MIN/MAX/COUNT.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
DATAGEN_PATH = os.path.join(_HERE, "libqk_datagen.so")
_lib = None


def threads() -> int:
    n = os.cpu_count() or 1
    return max(1, min(16, n))


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(DATAGEN_PATH):
            raise RuntimeError(f"{DATAGEN_PATH} is missing: run python -m qikkdb_amd.build")
        lib = C.CDLL(DATAGEN_PATH)
        u64, i32, i64, dbl, vp = C.c_uint64, C.c_int32, C.c_int64, C.c_double, C.c_void_p
        lib.qkgen_uniform_i32.argtypes = [u64, i32, i32, i64, vp, C.c_int]
        lib.qkgen_uniform_i64.argtypes = [u64, i64, i64, i64, vp, C.c_int]
        lib.qkgen_zipf_keys.argtypes = [u64, dbl, i64, i64, vp, vp, C.c_int]
        lib.qkgen_lognormal_f32.argtypes = [u64, dbl, dbl, i64, vp, C.c_int]
        lib.qkgen_null_mask.argtypes = [u64, dbl, i64, vp, C.c_int]
        for f in ("qkgen_uniform_i32", "qkgen_uniform_i64", "qkgen_zipf_keys",
                  "qkgen_lognormal_f32", "qkgen_null_mask"):
            getattr(lib, f).restype = C.c_int
        _lib = lib
    return _lib


def _chk(rc: int) -> None:
    if rc != 0:
        raise ValueError("invalid generator arguments")


def uniform_i32(seed: int, lo: int, hi: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    _chk(_load().qkgen_uniform_i32(seed, lo, hi, n, out.ctypes.data, threads()))
    return out


def uniform_i64(seed: int, lo: int, hi: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    _chk(_load().qkgen_uniform_i64(seed, lo, hi, n, out.ctypes.data, threads()))
    return out


def zipf_keys(seed: int, s: float, ndistinct: int, n: int, with_ranks: bool = False):
    out = np.empty(n, np.int64)
    ranks = np.empty(n, np.int64) if with_ranks else None
    _chk(_load().qkgen_zipf_keys(seed, s, ndistinct, n, out.ctypes.data,
                                 ranks.ctypes.data if with_ranks else None, threads()))
    return (out, ranks) if with_ranks else out


def lognormal_f32(seed: int, mu: float, sigma: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    _chk(_load().qkgen_lognormal_f32(seed, mu, sigma, n, out.ctypes.data, threads()))
    return out


def null_mask(seed: int, p: float, n: int) -> np.ndarray:
    """uint64 words, LSB-first, bit set = NULL."""
    out = np.empty((n + 63) // 64, np.uint64)
    _chk(_load().qkgen_null_mask(seed, p, n, out.ctypes.data, threads()))
    return out


# ---- workloads -------------------------------------------------------------------------------

YEAR_START_MS = 1735689600000          # 2025-01-01T00:00:00Z
DAY_MS = 86_400_000
YEAR_MS = 365 * DAY_MS
WINDOW_START_MS = YEAR_START_MS + 100 * DAY_MS
WINDOW_END_MS = WINDOW_START_MS + 30 * DAY_MS


def small(n: int = 1 << 20, seed: int = 1) -> dict:
    return {
        "key": uniform_i32(seed * 16 + 1, 0, 1000, n),
        "value": uniform_i32(seed * 16 + 2, -1000, 1001, n),
    }


def zipf(n: int, seed: int = 2, ndistinct: int = 1_000_000, s: float = 1.1,
         null_p: float = 0.05) -> dict:
    return {
        "key": zipf_keys(seed * 16 + 1, s, ndistinct, n),
        "value": lognormal_f32(seed * 16 + 2, 0.0, 1.0, n),
        "value_nulls": null_mask(seed * 16 + 3, null_p, n),
    }


def composite(n: int, seed: int = 3) -> dict:
    return {
        "a": uniform_i32(seed * 16 + 1, 0, 4096, n),
        "b": uniform_i32(seed * 16 + 2, 0, 4096, n),
        "ts": uniform_i64(seed * 16 + 3, YEAR_START_MS, YEAR_START_MS + YEAR_MS, n),
        "value": uniform_i64(seed * 16 + 4, -1_000_000_000, 1_000_000_001, n),
    }


ZIPF_LO, ZIPF_HI = np.float32(0.5), np.float32(4.0)
