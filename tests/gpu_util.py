"""GPU-side helpers for the parity tests: run the HIP path and the oracle on the same
inputs."""
from __future__ import annotations

import numpy as np

import oracle as O
import qikkdb_amd as qk
from qikkdb_amd import DeviceBuffer, DeviceColumn
from helpers import pack_bits, unpack_bits


def col(arr, isnull=None) -> DeviceColumn:
    return DeviceColumn.from_numpy(np.ascontiguousarray(arr), isnull)


class Where:
    """A predicate described once and built for both sides: `name OP const` or
    `name OP other`."""

    def __init__(self, op, column, const=None, const_type=None, rhs=None):
        self.op, self.column, self.const, self.const_type, self.rhs = op, column, const, const_type, rhs

    def gpu(self, cols: dict, nulls: dict):
        c = cols[self.column]
        if self.rhs is not None:
            return qk.where(self.op, c, cols[self.rhs])
        return qk.where(self.op, c, qk.Const(self.const, self.const_type))

    def cpu(self, host: dict, hnulls: dict):
        a = host[self.column]
        if self.rhs is not None:
            return O.pred(self.op, a, host[self.rhs], None, hnulls.get(self.column), hnulls.get(self.rhs))
        return O.pred(self.op, a, self.const, self.const_type, hnulls.get(self.column))


def gpu_group_by(aggs, keys, value=None, key_nulls=None, value_nulls=None, where=(), mask=None,
                 block=None, capacity=1 << 16, flags=0, tuning=None, out_types=None,
                 extra_cols=None, value_type=None, path=None, defer_rows=None, stats=None):
    """keys: list of numpy arrays; key_nulls / value_nulls: bool arrays or None.
    Rows are processed in blocks of `block` rows (each block uploaded with its own masks).
    A dict passed as `stats` receives the table's tuning state after the results."""
    keys = [np.ascontiguousarray(k) for k in keys]
    n = keys[0].size
    ktypes = [qk._abi.TYPE_OF[k.dtype] for k in keys]
    vtype = value_type if value_type is not None else (
        qk._abi.TYPE_OF[value.dtype] if value is not None else qk.INT)
    gb = qk.GPUGroupBy(aggs, ktypes, vtype, capacity, 1, out_types=out_types, flags=flags)
    if tuning:
        gb.set_tuning(**tuning)
    if path is not None:
        gb.set_path(path, 0)
    if defer_rows is not None:
        gb.set_defer_rows(defer_rows)
    block = block or n
    extra_cols = extra_cols or {}
    for s in range(0, n, block):
        e = min(n, s + block)
        kcols = [col(k[s:e], None if key_nulls is None or key_nulls[i] is None else key_nulls[i][s:e])
                 for i, k in enumerate(keys)]
        vcol = col(value[s:e], None if value_nulls is None else value_nulls[s:e]) if value is not None else None
        cols = {"value": vcol} if vcol is not None else {}
        for i, kc in enumerate(kcols):
            cols[f"key{i}"] = kc
        for name, (arr, isnull) in extra_cols.items():
            cols[name] = col(arr[s:e], None if isnull is None else isnull[s:e])
        knm = [kc.nulls for kc in kcols] if key_nulls is not None else None
        vptr = vcol.data if vcol is not None else None
        vnm = vcol.nulls if vcol is not None else None
        if where:
            preds = [w.gpu(cols, {}) for w in where]
            gb.ProcessBlockWhere([kc.data for kc in kcols], vptr, e - s, preds, knm, vnm)
        elif mask is not None:
            m = DeviceBuffer.from_numpy(np.ascontiguousarray(mask[s:e], np.int8))
            gb.ProcessBlockMasked([kc.data for kc in kcols], vptr, m, e - s, knm, vnm)
        else:
            gb.ProcessBlock([kc.data for kc in kcols], vptr, e - s, knm, vnm)
    res = gb.results_host()
    if stats is not None:
        stats.update(gb.stats())
    gb.close()
    return res


def cpu_group_by(aggs, keys, value=None, key_nulls=None, value_nulls=None, where=(), mask=None,
                 out_types=None, extra_cols=None, value_type=None):
    keys = [np.ascontiguousarray(k) for k in keys]
    ktypes = [O.TYPE_OF[k.dtype] for k in keys]
    vtype = value_type if value_type is not None else (
        O.TYPE_OF[value.dtype] if value is not None else O.INT)
    g = O.GroupBy(aggs, ktypes, vtype, out_types)
    host = {"value": value} if value is not None else {}
    hnulls = {"value": pack_bits(value_nulls)} if value_nulls is not None else {}
    for i, k in enumerate(keys):
        host[f"key{i}"] = k
        if key_nulls is not None and key_nulls[i] is not None:
            hnulls[f"key{i}"] = pack_bits(key_nulls[i])
    for name, (arr, isnull) in (extra_cols or {}).items():
        host[name] = arr
        if isnull is not None:
            hnulls[name] = pack_bits(isnull)
    kn = None
    if key_nulls is not None:
        kn = [pack_bits(x) if x is not None else None for x in key_nulls]
    preds = [w.cpu(host, hnulls) for w in where]
    g.process(keys, value, hnulls.get("value"), kn, mask=None if mask is None else np.ascontiguousarray(mask, np.int8),
              preds=preds)
    return g.results()
