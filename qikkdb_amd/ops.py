"""Python mirror of the reference's operator interfaces for the WHERE -> reconstruct ->
GROUP BY path, calling the HIP library through its C ABI.

Names follow the reference classes so that tests read like its own:
  * ``GPUBinary.Filter``        - GPUBinary<OP,int8_t,L,R>::Binary (GPUBinary.cuh:554-579)
  * ``GPUReconstruct``          - GPUReconstruct statics (GPUReconstruct.cuh:196-628)
  * ``GPUGroupBy``              - GPUGroupBy<AGG,O,K,V> / multi-key (GPUGroupBy.cuh:276-791,
                                  GPUGroupByMultiKey.cuh:282-1325)
Errors surface as QueryEngineError with the reference's error codes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _abi
from ._abi import call, Predicate
from .memory import DeviceBuffer, DeviceColumn, ensure_device, ptr_of, unpack_nulls


def _vp_array(items) -> C.Array:
    arr = (C.c_void_p * max(1, len(items)))()
    for i, x in enumerate(items):
        arr[i] = ptr_of(x)
    return arr


def _scalar_bytes(value, qtype: int) -> C.Array:
    v = np.array([value], dtype=_abi.NP_OF[qtype])
    buf = (C.c_char * v.nbytes).from_buffer_copy(v.tobytes())
    return buf


# ---- WHERE ----------------------------------------------------------------------------

class Const:
    """A scalar operand with its literal type (S2: the constant keeps its type)."""

    def __init__(self, value, qtype: int):
        self.value, self.qtype = value, qtype


def where(op: int, column: DeviceColumn, rhs, rhs_type: int | None = None) -> Predicate:
    """One comparison `column OP rhs`; rhs is a Const, a DeviceColumn or a Python number
    (then rhs_type must be given)."""
    p = Predicate()
    p.op = op
    p.type = column.qtype
    p.column = column.ptr
    p.null_mask = column.null_ptr
    if isinstance(rhs, DeviceColumn):
        p.rhs_kind = _abi.COLUMN
        p.rhs_type = rhs.qtype
        p.rhs_column = rhs.ptr
        p.rhs_null_mask = rhs.null_ptr
    else:
        if isinstance(rhs, Const):
            value, qtype = rhs.value, rhs.qtype
        else:
            if rhs_type is None:
                raise ValueError("rhs_type is required for a plain constant")
            value, qtype = rhs, rhs_type
        p.rhs_kind = _abi.CONST
        p.rhs_type = qtype
        if qtype in (_abi.FLOAT, _abi.DOUBLE):
            p.rhs_real = float(np.array(value, dtype=_abi.NP_OF[qtype]))
        else:
            p.rhs_int = int(value)
    return p


def _pred_array(preds: Sequence[Predicate]):
    arr = (Predicate * max(1, len(preds)))()
    for i, p in enumerate(preds):
        arr[i] = p
    return arr


class GPUBinary:
    @staticmethod
    def Filter(op: int, A, B, n: int, null_mask: DeviceBuffer | None = None,
               out: DeviceBuffer | None = None, bits: bool = False, stream=None) -> DeviceBuffer:
        """int8 mask (or bit mask when bits=True) of `A OP B`; A/B are DeviceColumn or Const."""
        ensure_device()
        def operand(x):
            if isinstance(x, DeviceColumn):
                return x.qtype, _abi.COLUMN, C.c_void_p(x.ptr), None
            if isinstance(x, Const):
                b = _scalar_bytes(x.value, x.qtype)
                return x.qtype, _abi.CONST, C.cast(b, C.c_void_p), b
            raise TypeError("operand must be DeviceColumn or Const")
        lt, lk, lp, _lkeep = operand(A)
        rt, rk, rp, _rkeep = operand(B)
        if out is None:
            out = DeviceBuffer(((n + 63) // 64) * 8 if bits else n)
        fn = "qk_filter_bits" if bits else "qk_filter"
        call(fn, op, lt, lk, lp, rt, rk, rp, n, ptr_of(null_mask), out.ptr, stream)
        return out

    @staticmethod
    def LogicalNot(mask: DeviceBuffer, n: int, null_mask: DeviceBuffer | None = None,
                   stream=None) -> DeviceBuffer:
        out = DeviceBuffer(n)
        call("qk_logical_not", mask.ptr, n, ptr_of(null_mask), out.ptr, stream)
        return out


# ---- reconstruct ------------------------------------------------------------------------

class GPUReconstruct:
    @staticmethod
    def PrefixSum(mask: DeviceBuffer, n: int, inclusive: bool = True, stream=None) -> DeviceBuffer:
        ensure_device()
        out = DeviceBuffer(max(1, n) * 4)
        call("qk_prefix_sum", mask.ptr, n, out.ptr, 1 if inclusive else 0, stream)
        return out

    @staticmethod
    def PrefixSumExclusive(mask: DeviceBuffer, n: int, stream=None) -> DeviceBuffer:
        return GPUReconstruct.PrefixSum(mask, n, inclusive=False, stream=stream)

    @staticmethod
    def Sum(mask: DeviceBuffer, n: int, stream=None) -> int:
        ensure_device()
        v = C.c_int64()
        call("qk_mask_sum", mask.ptr, n, C.byref(v), stream)
        return v.value

    @staticmethod
    def reconstructColKeep(col: DeviceBuffer, elem_size: int, mask: DeviceBuffer | None, n: int,
                           in_null: DeviceBuffer | None = None, want_null: bool = True,
                           stream=None):
        """-> (out buffer or None, count, out null mask or None).  With mask=None the input
        buffer itself is returned (aliasing, GPUReconstruct.cuh:320-328)."""
        ensure_device()
        out, cnt, onull = C.c_void_p(), C.c_int32(), C.c_void_p()
        call("qk_reconstruct_col_keep", elem_size, ptr_of(col), ptr_of(mask), n, ptr_of(in_null),
             C.byref(out), C.byref(cnt), C.byref(onull) if want_null else None, stream)
        if mask is None:
            return col, cnt.value, in_null
        ob = DeviceBuffer(cnt.value * elem_size, ptr=out.value) if out.value else None
        nb = (DeviceBuffer(((cnt.value + 63) // 64) * 8, ptr=onull.value) if onull.value else None)
        return ob, cnt.value, nb

    @staticmethod
    def reconstructCol(col: DeviceColumn, mask: DeviceBuffer | None, stream=None):
        """Host result: (numpy values, numpy bool nulls or None)."""
        ob, cnt, nb = GPUReconstruct.reconstructColKeep(
            col.data, _abi.SIZE_OF[col.qtype], mask, col.n, col.nulls, True, stream)
        vals = ob.to_numpy(_abi.NP_OF[col.qtype], cnt) if (ob is not None and cnt) else \
            np.empty(0, _abi.NP_OF[col.qtype])
        nulls = unpack_nulls(nb.to_numpy(np.uint64), cnt) if nb is not None else None
        return vals, nulls

    @staticmethod
    def GenerateIndexesKeep(mask: DeviceBuffer, n: int, index_size: int = 4, stream=None):
        ensure_device()
        out, cnt = C.c_void_p(), C.c_int32()
        call("qk_generate_indexes_keep", mask.ptr, n, index_size, C.byref(out), C.byref(cnt), stream)
        ob = DeviceBuffer(cnt.value * index_size, ptr=out.value) if out.value else None
        return ob, cnt.value

    @staticmethod
    def GenerateIndexes(mask: DeviceBuffer, n: int, index_size: int = 4, stream=None) -> np.ndarray:
        ob, cnt = GPUReconstruct.GenerateIndexesKeep(mask, n, index_size, stream)
        dt = np.int32 if index_size == 4 else np.int64
        return ob.to_numpy(dt, cnt) if ob is not None else np.empty(0, dt)

    @staticmethod
    def CompressNullMask(null_array: DeviceBuffer, n: int, stream=None) -> DeviceBuffer | None:
        ensure_device()
        out = C.c_void_p()
        call("qk_compress_null_mask", null_array.ptr, n, C.byref(out), stream)
        return DeviceBuffer(((n + 63) // 64) * 8, ptr=out.value) if out.value else None

    @staticmethod
    def WhereCompact(preds: Sequence[Predicate], cols: Sequence[DeviceColumn], n: int,
                     stream=None):
        """Fused WHERE + reconstruct: -> (list of DeviceColumn of the kept rows, count)."""
        ensure_device()
        outs, onulls = [], []
        for c in cols:
            outs.append(DeviceBuffer(max(1, n) * _abi.SIZE_OF[c.qtype]))
            onulls.append(DeviceBuffer(((n + 63) // 64) * 8) if c.nulls is not None else None)
        sizes = (C.c_int32 * max(1, len(cols)))(*[_abi.SIZE_OF[c.qtype] for c in cols])
        cnt = C.c_int32()
        call("qk_where_compact", len(preds), _pred_array(preds), len(cols), sizes,
             _vp_array([c.data for c in cols]), _vp_array([c.nulls for c in cols]), n,
             _vp_array(outs), _vp_array(onulls), C.byref(cnt), stream)
        res = [DeviceColumn(o, c.qtype, cnt.value, on) for o, c, on in zip(outs, cols, onulls)]
        return res, cnt.value


# ---- GROUP BY ----------------------------------------------------------------------------

@dataclass
class GroupByResult:
    count: int
    keys: list = field(default_factory=list)        # DeviceBuffer per key column
    key_nulls: list = field(default_factory=list)   # DeviceBuffer or None per key column
    values: list = field(default_factory=list)      # DeviceBuffer or None per aggregate
    value_nulls: list = field(default_factory=list)


class GPUGroupBy:
    """Hash GROUP BY table.  `aggs` is one aggregate (the reference's AGG) or a list of them,
    which are then computed in the same pass."""

    def __init__(self, aggs, key_types, value_type: int, max_hash_count: int,
                 key_size_multiplier: int = 1, out_types=None, flags: int = 0):
        ensure_device()
        self.aggs = [aggs] if isinstance(aggs, int) else list(aggs)
        self.key_types = [key_types] if isinstance(key_types, int) else list(key_types)
        self.value_type = value_type
        na, nk = len(self.aggs), len(self.key_types)
        if out_types is None:
            out_types = [0] * na
        self.out_types = [
            (t if t else (_abi.LONG if a in (_abi.AGG_COUNT, _abi.AGG_COUNT_ALL) else value_type))
            for a, t in zip(self.aggs, out_types)]
        h = C.c_void_p()
        call("qk_gb_create", C.byref(h), nk, (C.c_int32 * nk)(*self.key_types), value_type, na,
             (C.c_int32 * na)(*self.aggs), (C.c_int32 * na)(*out_types), max_hash_count,
             key_size_multiplier, flags)
        self.h = h.value

    # -- life cycle
    def close(self) -> None:
        if getattr(self, "h", None):
            _abi.load().qk_gb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, stream=None) -> None:
        call("qk_gb_reset", self.h, stream)

    def GetMaxHashCount(self) -> int:
        v = C.c_int32()
        call("qk_gb_get_max_hash_count", self.h, C.byref(v))
        return v.value

    def set_tuning(self, block: int = -1, lds_bytes: int = -1, wgs_per_cu: int = -1) -> None:
        call("qk_gb_set_tuning", self.h, block, lds_bytes, wgs_per_cu)

    def set_path(self, path: int = -1, min_rows: int = -1) -> None:
        """-1 automatic, 0 one-pass kernel, 1 two-phase partitioned kernels."""
        call("qk_gb_set_path", self.h, path, min_rows)

    def set_defer_rows(self, rows: int = -1) -> None:
        """Rows whose partition merge an asynchronous table may defer (-1 default, 0 none)."""
        call("qk_gb_set_defer_rows", self.h, rows)

    def stats(self, stream=None) -> dict:
        names = ["path", "heavy_keys", "heavy_threshold", "region_rows", "p1_workgroups",
                 "partitions", "region_cap", "p1_lds_slots", "p2_lds_slots", "heavy_limit", "sub_tables",
                 "grows", "p1_form"]
        v = (C.c_int64 * len(names))()
        call("qk_gb_stats", self.h, v, len(names), stream)
        return dict(zip(names, list(v)))

    def set_profiling(self, on: bool = True) -> None:
        call("qk_gb_set_profiling", self.h, 1 if on else 0)

    def profile(self):
        ms, n = C.c_double(), C.c_int64()
        call("qk_gb_get_profile", self.h, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def describe(self):
        sb, slots, lds = C.c_int32(), C.c_int64(), C.c_int32()
        call("qk_gb_describe", self.h, C.byref(sb), C.byref(slots), C.byref(lds))
        return {"slot_bytes": sb.value, "slots": slots.value, "lds_slots": lds.value}

    # -- blocks
    @staticmethod
    def _as_list(x):
        if x is None:
            return None
        return list(x) if isinstance(x, (list, tuple)) else [x]

    def _key_args(self, keys, key_nulls):
        keys = self._as_list(keys)
        if len(keys) != len(self.key_types):
            raise _abi.QueryEngineError(_abi.QK_ERROR, "Incorrect number of key columns in ProcessBlock")
        kn = self._as_list(key_nulls)
        return _vp_array(keys), (_vp_array(kn) if kn is not None else None)

    def ProcessBlock(self, keys, values, n: int, key_nulls=None, value_nulls=None, stream=None):
        k, kn = self._key_args(keys, key_nulls)
        call("qk_gb_process_block", self.h, k, kn, ptr_of(values), ptr_of(value_nulls), n, stream)

    def ProcessBlockMasked(self, keys, values, mask, n: int, key_nulls=None, value_nulls=None,
                           stream=None):
        k, kn = self._key_args(keys, key_nulls)
        call("qk_gb_process_block_masked", self.h, k, kn, ptr_of(values), ptr_of(value_nulls),
             ptr_of(mask), n, stream)

    def ProcessBlockWhere(self, keys, values, n: int, preds: Sequence[Predicate],
                          key_nulls=None, value_nulls=None, stream=None):
        k, kn = self._key_args(keys, key_nulls)
        call("qk_gb_process_block_where", self.h, k, kn, ptr_of(values), ptr_of(value_nulls), n,
             len(preds), _pred_array(preds), stream)

    def check_error(self, stream=None) -> None:
        call("qk_gb_check_error", self.h, stream)

    def merge(self, src: "GPUGroupBy", stream=None) -> None:
        """Fold every group of `src` into this table (qk_gb_merge)."""
        call("qk_gb_merge", self.h, src.h, stream)

    def group_count(self, stream=None) -> int:
        v = C.c_int64()
        call("qk_gb_group_count", self.h, C.byref(v), stream)
        return v.value

    # -- results
    def GetResults(self, want_nulls: bool = True, stream=None) -> GroupByResult:
        nk, na = len(self.key_types), len(self.aggs)
        ok = (C.c_void_p * nk)()
        okn = (C.c_void_p * nk)()
        ov = (C.c_void_p * na)()
        ovn = (C.c_void_p * na)()
        cnt = C.c_int32()
        call("qk_gb_get_results", self.h, ok, okn if want_nulls else None, ov,
             ovn if want_nulls else None, C.byref(cnt), stream)
        g = cnt.value
        nbytes = ((g + 63) // 64) * 8
        wrap = lambda p, nb: DeviceBuffer(nb, ptr=p) if p else None  # noqa: E731
        return GroupByResult(
            count=g,
            keys=[wrap(ok[i], g * _abi.SIZE_OF[self.key_types[i]]) for i in range(nk)],
            key_nulls=[wrap(okn[i], nbytes) if want_nulls else None for i in range(nk)],
            values=[wrap(ov[a], g * _abi.SIZE_OF[self.out_types[a]]) for a in range(na)],
            value_nulls=[wrap(ovn[a], nbytes) if want_nulls else None for a in range(na)],
        )

    def results_host(self, stream=None) -> dict:
        """GetResults copied to the host: numpy keys / values and bool null flags."""
        r = self.GetResults(True, stream)
        g = r.count
        out = {"count": g, "keys": [], "key_nulls": [], "values": [], "value_nulls": []}
        for i, t in enumerate(self.key_types):
            out["keys"].append(r.keys[i].to_numpy(_abi.NP_OF[t], g) if g else np.empty(0, _abi.NP_OF[t]))
            kn = r.key_nulls[i]
            out["key_nulls"].append(unpack_nulls(kn.to_numpy(np.uint64), g) if (kn is not None and g)
                                    else np.zeros(g, bool))
        for a, t in enumerate(self.out_types):
            if r.values[a] is None:
                out["values"].append(None)
                out["value_nulls"].append(None)
                continue
            out["values"].append(r.values[a].to_numpy(_abi.NP_OF[t], g))
            vn = r.value_nulls[a]
            out["value_nulls"].append(unpack_nulls(vn.to_numpy(np.uint64), g) if vn is not None
                                      else np.zeros(g, bool))
        return out
