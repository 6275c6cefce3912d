"""ctypes binding of the C ABI in include/qikkdb_amd.h.

The product path is the HIP library built in-tree (qikkdb_amd/libqikkdb_amd.so).  There is
no CPU fallback: loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqikkdb_amd.so")

# ---- constants (include/qikkdb_amd.h) -------------------------------------------------
QK_SUCCESS, QK_ERROR, QK_DIVISION_BY_ZERO, QK_INTEGER_OVERFLOW = 0, 1, 2, 3
QK_HASH_TABLE_FULL, QK_UNKNOWN_AGG_FUN = 4, 5
QK_RUNTIME_ERROR, QK_OUT_OF_MEMORY = 64, 65

INT, LONG, FLOAT, DOUBLE, INT8 = 8, 9, 10, 11, 15
GREATER, GREATER_EQUAL, LESS, LESS_EQUAL, EQUAL, NOT_EQUAL, LOGICAL_AND, LOGICAL_OR = range(8)
COLUMN, CONST = 0, 1
AGG_MIN, AGG_MAX, AGG_SUM, AGG_AVG, AGG_COUNT, AGG_NONE, AGG_COUNT_ALL = range(7)
GB_KEYS_NOT_NULL, GB_ASYNC, GB_AUTO_GROW = 1, 2, 4

NP_OF = {INT: np.int32, LONG: np.int64, FLOAT: np.float32, DOUBLE: np.float64, INT8: np.int8}
TYPE_OF = {np.dtype(v): k for k, v in NP_OF.items()}
SIZE_OF = {INT: 4, LONG: 8, FLOAT: 4, DOUBLE: 8, INT8: 1}

ERROR_NAMES = {
    QK_ERROR: "GPU_EXTENSION_ERROR",
    QK_DIVISION_BY_ZERO: "GPU_DIVISION_BY_ZERO_ERROR",
    QK_INTEGER_OVERFLOW: "GPU_INTEGER_OVERFLOW_ERROR",
    QK_HASH_TABLE_FULL: "GPU_HASH_TABLE_FULL",
    QK_UNKNOWN_AGG_FUN: "GPU_UNKNOWN_AGG_FUN",
    QK_RUNTIME_ERROR: "HIP_RUNTIME_ERROR",
    QK_OUT_OF_MEMORY: "OUT_OF_MEMORY",
}


class QueryEngineError(RuntimeError):
    """Mirror of the reference's query_engine_error (QueryEngine/GPUError.h:43-65)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code

    def GetQueryEngineError(self) -> int:  # noqa: N802 - reference method name
        return self.code


class Predicate(C.Structure):
    """qk_predicate: one comparison of a conjunctive WHERE."""

    _fields_ = [
        ("op", C.c_int32),
        ("type", C.c_int32),
        ("column", C.c_void_p),
        ("null_mask", C.c_void_p),
        ("rhs_kind", C.c_int32),
        ("rhs_type", C.c_int32),
        ("rhs_column", C.c_void_p),
        ("rhs_null_mask", C.c_void_p),
        ("rhs_int", C.c_int64),
        ("rhs_real", C.c_double),
    ]


class BatchColumns(C.Structure):
    """qk_batch_columns: the device-resident table of a batch engine."""

    _fields_ = [
        ("a", C.c_void_p),
        ("b", C.c_void_p),
        ("ts", C.c_void_p),
        ("value", C.c_void_p),
        ("rows", C.c_int64),
        ("block_rows", C.c_int32),
        ("b_range", C.c_int32),
    ]


class BatchQuery(C.Structure):
    """qk_batch_query: one filtered GROUP BY of a batch."""

    _fields_ = [
        ("first_block", C.c_int32),
        ("nblocks", C.c_int32),
        ("ts_lo", C.c_int64),
        ("ts_hi", C.c_int64),
        ("key_mod", C.c_int32),
        ("reserved", C.c_int32),
    ]


class BatchTuning(C.Structure):
    """qk_batch_tuning."""

    _fields_ = [
        ("tile_rows", C.c_int32),
        ("lds_groups", C.c_int32),
        ("wgs_per_cu", C.c_int32),
        ("use_graph", C.c_int32),
        ("scratch_rows", C.c_int64),
    ]


class ServeConfig(C.Structure):
    """qk_serve_config."""

    _fields_ = [
        ("engines", C.c_int32),
        ("max_batch", C.c_int32),
        ("memory_cap", C.c_int64),
        ("wait_ms", C.c_double),
        ("tuning", BatchTuning),
    ]


class ServeStats(C.Structure):
    """qk_serve_stats."""

    _fields_ = [
        ("batches", C.c_int64),
        ("served", C.c_int64),
        ("failed", C.c_int64),
        ("pending", C.c_int64),
        ("max_batch", C.c_int64),
        ("admitted_bytes", C.c_int64),
        ("peak_admitted_bytes", C.c_int64),
        ("busy_ms", C.c_double),
    ]


_vp, _i32, _i64, _u32, _sz = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_size_t
_P = C.POINTER

# name -> argtypes (every function returns int status, except qk_last_error_message)
PROTOTYPES = {
    "qk_init": [C.c_int],
    "qk_device_count": [_P(C.c_int)],
    "qk_malloc": [_P(_vp), _sz],
    "qk_free": [_vp],
    "qk_mem_stats": [_P(_sz), _P(_sz)],
    "qk_mem_release_cached": [],
    "qk_memcpy_htod": [_vp, _vp, _sz, _vp],
    "qk_memcpy_dtoh": [_vp, _vp, _sz, _vp],
    "qk_memcpy_dtod": [_vp, _vp, _sz, _vp],
    "qk_memset": [_vp, C.c_int, _sz, _vp],
    "qk_stream_create": [_P(_vp)],
    "qk_stream_destroy": [_vp],
    "qk_stream_synchronize": [_vp],
    "qk_device_synchronize": [],
    "qk_filter": [_i32, _i32, _i32, _vp, _i32, _i32, _vp, _i32, _vp, _vp, _vp],
    "qk_filter_bits": [_i32, _i32, _i32, _vp, _i32, _i32, _vp, _i32, _vp, _vp, _vp],
    "qk_logical_not": [_vp, _i32, _vp, _vp, _vp],
    "qk_prefix_sum": [_vp, _i32, _vp, _i32, _vp],
    "qk_prefix_sum_noalloc": [_vp, _i32, _vp, _i32, _vp],
    "qk_mask_sum": [_vp, _i32, _P(_i64), _vp],
    "qk_reconstruct_col_keep": [_i32, _vp, _vp, _i32, _vp, _P(_vp), _P(_i32), _P(_vp), _vp],
    "qk_generate_indexes_keep": [_vp, _i32, _i32, _P(_vp), _P(_i32), _vp],
    "qk_compress_null_mask": [_vp, _i32, _P(_vp), _vp],
    "qk_where_compact": [_i32, _P(Predicate), _i32, _P(_i32), _P(_vp), _P(_vp), _i32, _P(_vp),
                         _P(_vp), _P(_i32), _vp],
    "qk_gb_create": [_P(_vp), _i32, _P(_i32), _i32, _i32, _P(_i32), _P(_i32), _i32, _i32, _u32],
    "qk_gb_destroy": [_vp],
    "qk_gb_reset": [_vp, _vp],
    "qk_gb_get_max_hash_count": [_vp, _P(_i32)],
    "qk_gb_process_block": [_vp, _P(_vp), _P(_vp), _vp, _vp, _i32, _vp],
    "qk_gb_process_block_masked": [_vp, _P(_vp), _P(_vp), _vp, _vp, _vp, _i32, _vp],
    "qk_gb_process_block_where": [_vp, _P(_vp), _P(_vp), _vp, _vp, _i32, _i32, _P(Predicate), _vp],
    "qk_gb_check_error": [_vp, _vp],
    "qk_gb_get_results": [_vp, _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_i32), _vp],
    "qk_gb_group_count": [_vp, _P(_i64), _vp],
    "qk_gb_set_profiling": [_vp, _i32],
    "qk_gb_get_profile": [_vp, _P(C.c_double), _P(_i64)],
    "qk_gb_describe": [_vp, _P(_i32), _P(_i64), _P(_i32)],
    "qk_gb_set_tuning": [_vp, _i32, _i32, _i32],
    "qk_gb_set_path": [_vp, _i32, _i64],
    "qk_gb_set_defer_rows": [_vp, _i64],
    "qk_gb_stats": [_vp, _P(_i64), _i32, _vp],
    "qk_gb_seed_keys": [_vp, _vp, _i32, _i32, _vp],
    "qk_gb_merge": [_vp, _vp, _vp],
    "qk_gb_get_results_aggs": [_vp, _i32, _P(_i32), _P(_i32), _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_i32), _vp],
    "qk_batch_create": [_P(_vp), _P(BatchColumns)],
    "qk_batch_destroy": [_vp],
    "qk_batch_set_tuning": [_vp, _P(BatchTuning)],
    "qk_batch_get_tuning": [_vp, _P(BatchTuning)],
    "qk_batch_autotune": [_vp, _i32, _P(BatchQuery), _P(BatchTuning), _vp],
    "qk_batch_run": [_vp, _i32, _P(BatchQuery), _vp],
    "qk_batch_run_into": [_vp, _i32, _P(BatchQuery), _vp, _vp, _vp, _vp, _vp],
    "qk_batch_results": [_vp, _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_vp)],
    "qk_batch_completion_ms": [_vp, _P(C.c_double), _i32],
    "qk_batch_key_column": [_vp, _vp, _i32, _i32, _i32, _vp, _vp],
    "qk_batch_resample": [_vp],
    "qk_batch_estimate_rows": [_vp, _i32, _P(BatchQuery), _P(_i64)],
    "qk_gbs_create": [_P(_vp), _i32, _i32, _P(_i32), _P(_i32), _i32, _i32, _u32],
    "qk_gbs_destroy": [_vp],
    "qk_gbs_reset": [_vp, _vp],
    "qk_gbs_process_block": [_vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp],
    "qk_gbs_process_block_where": [_vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _P(Predicate), _vp],
    "qk_gbs_get_results": [_vp, _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_i32), _vp],
    "qk_gbs_group_count": [_vp, _P(_i64), _vp],
    "qk_mem_set_limit": [_sz],
    "qk_mem_set_limit_percent": [_i32],
    "qk_mem_get_limit": [_P(_sz)],
    "qk_mem_can_allocate": [_sz, _P(_i32)],
    "qk_cache_create": [_P(_vp), _sz],
    "qk_cache_destroy": [_vp],
    "qk_cache_get_column": [_vp, C.c_char_p, C.c_char_p, _i32, _sz, _i64, _i64, _P(_vp), _P(_sz), _P(_i32)],
    "qk_cache_contains": [_vp, C.c_char_p, C.c_char_p, _i32, _i64, _i64, _P(_i32)],
    "qk_cache_evict": [_vp, _P(_i32)],
    "qk_cache_clear_block": [_vp, C.c_char_p, C.c_char_p, _i32],
    "qk_cache_set_current_block": [_vp, _i64],
    "qk_cache_set_lock_list": [_i32, _P(C.c_char_p)],
    "qk_cache_stats": [_vp, _P(_sz), _P(_i64), _P(_i64), _P(_i64), _P(_i64)],
    "qk_serve_create": [_P(_vp), _P(BatchColumns), _P(ServeConfig)],
    "qk_serve_destroy": [_vp],
    "qk_serve_submit": [_vp, _P(BatchQuery), _P(_i64)],
    "qk_serve_wait": [_vp, _i64, C.c_double],
    "qk_serve_result": [_vp, _i64, _P(_vp), _P(_vp), _P(_vp), _P(_vp)],
    "qk_serve_latency_ms": [_vp, _i64, _P(C.c_double)],
    "qk_serve_release": [_vp, _i64],
    "qk_serve_replay": [_vp, _i32, _P(BatchQuery), _P(C.c_double), _P(C.c_double)],
    "qk_serve_get_stats": [_vp, _P(ServeStats)],
    "qk_serve_replay_plain": [_P(BatchColumns), _i32, _P(BatchQuery), _P(C.c_double), _P(C.c_double), _vp],
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP extension first "
            "(python -m qikkdb_amd.build or __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, args in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.qk_last_error_message.argtypes = []
    lib.qk_last_error_message.restype = C.c_char_p
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != QK_SUCCESS:
        msg = _lib.qk_last_error_message().decode(errors="replace") if _lib else ""
        raise QueryEngineError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
