"""qikkdb_amd - MI355X-native (gfx950) WHERE -> reconstruct -> GROUP BY for qikkDB.

The compute path is the HIP library libqikkdb_amd.so (C ABI in include/qikkdb_amd.h).  This
package is its Python host side: a ctypes binding plus mirrors of the reference's operator
classes (GPUBinary filter, GPUReconstruct, GPUGroupBy).  There is no CPU fallback.
"""
from ._abi import (  # noqa: F401
    AGG_AVG, AGG_COUNT, AGG_COUNT_ALL, AGG_MAX, AGG_MIN, AGG_NONE, AGG_SUM,
    COLUMN, CONST, DOUBLE, EQUAL, FLOAT, GB_ASYNC, GB_AUTO_GROW, GB_KEYS_NOT_NULL, GREATER,
    GREATER_EQUAL, INT, INT8, LESS, LESS_EQUAL, LOGICAL_AND, LOGICAL_OR, LONG, NOT_EQUAL,
    NP_OF, QK_ERROR, QK_HASH_TABLE_FULL, QK_SUCCESS, Predicate, QueryEngineError, load,
)
from .memory import (  # noqa: F401
    BlockCache, DeviceBuffer, DeviceColumn, can_allocate, ensure_device, memory_limit, memory_stats, pack_nulls,
    set_memory_limit, set_memory_limit_percent, unpack_nulls,
)
from .ops import Const, GPUBinary, GPUGroupBy, GPUReconstruct, GroupByResult, where  # noqa: F401

__all__ = [
    "DeviceBuffer", "DeviceColumn", "GPUBinary", "GPUGroupBy", "GPUReconstruct", "Const",
    "where", "QueryEngineError", "load", "ensure_device", "BlockCache", "set_memory_limit",
    "set_memory_limit_percent", "memory_limit", "can_allocate", "memory_stats",
]
