"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

INT, LONG, FLOAT, DOUBLE, INT8 = 8, 9, 10, 11, 15
NP_OF = {INT: np.int32, LONG: np.int64, FLOAT: np.float32, DOUBLE: np.float64, INT8: np.int8}
TYPE_OF = {np.dtype(v): k for k, v in NP_OF.items()}
AGG_MIN, AGG_MAX, AGG_SUM, AGG_AVG, AGG_COUNT, AGG_NONE, AGG_COUNT_ALL = range(7)
COLUMN, CONST = 0, 1


class OrPred(C.Structure):
    _fields_ = [
        ("op", C.c_int32), ("type", C.c_int32), ("column", C.c_void_p), ("null_mask", C.c_void_p),
        ("rhs_kind", C.c_int32), ("rhs_type", C.c_int32), ("rhs_column", C.c_void_p),
        ("rhs_null_mask", C.c_void_p), ("rhs_int", C.c_int64), ("rhs_real", C.c_double),
    ]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.or_filter.argtypes = [i32, i32, i32, vp, i32, i32, vp, i32, vp, vp]
        L.or_logical_not.argtypes = [vp, i32, vp, vp]
        L.or_logical_not.restype = None
        L.or_prefix_sum.argtypes = [vp, i32, vp, i32]
        L.or_prefix_sum.restype = None
        L.or_mask_sum.argtypes = [vp, i32]
        L.or_mask_sum.restype = i64
        L.or_compact.argtypes = [i32, vp, vp, i32, vp, vp, vp]
        L.or_compact.restype = i32
        L.or_generate_indexes.argtypes = [vp, i32, vp]
        L.or_generate_indexes.restype = i32
        L.or_compress_null_mask.argtypes = [vp, i32, vp]
        L.or_compress_null_mask.restype = None
        L.or_eval_where.argtypes = [i32, vp, i64]
        L.or_gb_create.argtypes = [i32, vp, i32, i32, vp, vp]
        L.or_gb_create.restype = vp
        L.or_gb_destroy.argtypes = [vp]
        L.or_gb_destroy.restype = None
        L.or_gb_process.argtypes = [vp, vp, vp, vp, vp, vp, i32, vp, i64, i64]
        L.or_gb_process.restype = None
        L.or_gb_merge.argtypes = [vp, vp]
        L.or_gb_count.argtypes = [vp]
        L.or_gb_count.restype = i64
        L.or_gb_results.argtypes = [vp, vp, vp, vp, vp, vp]
        L.or_gb_results.restype = None
        L.or_gb_result_is_real.argtypes = [vp, i32]
        L.or_gb_query_mt.argtypes = [i32, i32, vp, i32, i32, vp, vp, vp, vp, vp, vp, i32, vp, i64]
        L.or_gb_query_mt.restype = vp
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _vp_arr(items):
    arr = (C.c_void_p * max(1, len(items)))()
    for i, x in enumerate(items):
        arr[i] = None if x is None else x.ctypes.data
    return arr


def filter_mask(op, lhs, rhs, n, null_words=None, ltype=None, rtype=None):
    """lhs/rhs: numpy column or (value, type) tuple for a constant."""
    def operand(x, t):
        if isinstance(x, np.ndarray):
            return TYPE_OF[x.dtype], COLUMN, x
        v, t = x
        return t, CONST, np.array([v], dtype=NP_OF[t])
    lt, lk, la = operand(lhs, ltype)
    rt, rk, ra = operand(rhs, rtype)
    out = np.empty(n, np.int8)
    rc = lib().or_filter(op, lt, lk, _p(la), rt, rk, _p(ra), n, _p(null_words), _p(out))
    if rc != 0:
        raise ValueError("oracle: unsupported filter")
    return out


def logical_not(mask, null_words=None):
    out = np.empty(mask.size, np.int8)
    lib().or_logical_not(_p(mask), mask.size, _p(null_words), _p(out))
    return out


def prefix_sum(mask, inclusive=True):
    out = np.empty(mask.size, np.int32)
    lib().or_prefix_sum(_p(mask), mask.size, _p(out), 1 if inclusive else 0)
    return out


def mask_sum(mask):
    return lib().or_mask_sum(_p(mask), mask.size)


def compact(values, mask, in_null_words=None):
    n = values.size
    out = np.empty(n, values.dtype)
    out_null = np.empty((n + 63) // 64, np.uint64) if in_null_words is not None else None
    k = lib().or_compact(values.dtype.itemsize, _p(values), _p(mask), n, _p(in_null_words),
                         _p(out), _p(out_null))
    return out[:k], (out_null[: (k + 63) // 64] if out_null is not None else None), k


def generate_indexes(mask):
    out = np.empty(mask.size, np.int32)
    k = lib().or_generate_indexes(_p(mask), mask.size, _p(out))
    return out[:k]


def compress_null_mask(arr):
    arr = np.ascontiguousarray(arr, np.uint8)
    out = np.empty((arr.size + 63) // 64, np.uint64)
    lib().or_compress_null_mask(_p(arr), arr.size, _p(out))
    return out


def pred(op, column, rhs, rhs_type=None, null_words=None, rhs_null_words=None):
    """Oracle predicate over host columns; rhs = numpy column or constant (needs rhs_type)."""
    p = OrPred()
    p.op = op
    p.type = TYPE_OF[column.dtype]
    p.column = column.ctypes.data
    p.null_mask = _p(null_words)
    if isinstance(rhs, np.ndarray):
        p.rhs_kind, p.rhs_type, p.rhs_column = COLUMN, TYPE_OF[rhs.dtype], rhs.ctypes.data
        p.rhs_null_mask = _p(rhs_null_words)
    else:
        p.rhs_kind, p.rhs_type = CONST, rhs_type
        if rhs_type in (FLOAT, DOUBLE):
            p.rhs_real = float(np.array(rhs, dtype=NP_OF[rhs_type]))
        else:
            p.rhs_int = int(rhs)
    return p


class GroupBy:
    """Oracle GROUP BY table; results() returns keys/values in unspecified order."""

    def __init__(self, aggs, key_types, value_type, out_types=None, handle=None):
        self.aggs = [aggs] if isinstance(aggs, int) else list(aggs)
        self.key_types = [key_types] if isinstance(key_types, int) else list(key_types)
        self.value_type = value_type
        na, nk = len(self.aggs), len(self.key_types)
        self._ot = (C.c_int32 * na)(*(out_types or [0] * na))
        if handle is None:
            handle = lib().or_gb_create(nk, (C.c_int32 * nk)(*self.key_types), value_type, na,
                                        (C.c_int32 * na)(*self.aggs), self._ot)
            if not handle:
                raise ValueError("oracle: bad group by arguments")
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_gb_destroy(self.h)
            self.h = None

    def process(self, keys, values=None, value_nulls=None, key_nulls=None, mask=None,
                preds=(), begin=0, end=None):
        keys = list(keys) if isinstance(keys, (list, tuple)) else [keys]
        n = keys[0].size if end is None else end
        ka = _vp_arr(keys)
        kn = _vp_arr(key_nulls) if key_nulls is not None else None
        pa = (OrPred * max(1, len(preds)))(*preds)
        lib().or_gb_process(self.h, ka, kn, _p(values), _p(value_nulls), _p(mask), len(preds), pa,
                            begin, n)

    def count(self):
        return lib().or_gb_count(self.h)

    def results(self):
        g = self.count()
        L = lib()
        keys = [np.empty(g, NP_OF[t]) for t in self.key_types]
        knull = [np.empty(g, np.uint8) for _ in self.key_types]
        na = len(self.aggs)
        ints = [np.empty(g, np.int64) for _ in range(na)]
        reals = [np.empty(g, np.float64) for _ in range(na)]
        vnull = [np.empty(g, np.uint8) for _ in range(na)]
        L.or_gb_results(self.h, _vp_arr(keys), _vp_arr(knull), _vp_arr(ints), _vp_arr(reals),
                        _vp_arr(vnull))
        values = []
        for a in range(na):
            if self.aggs[a] == AGG_NONE:
                values.append(None)
            elif L.or_gb_result_is_real(self.h, a):
                values.append(reals[a])
            else:
                values.append(ints[a])
        return {"count": g, "keys": keys, "key_nulls": [k.astype(bool) for k in knull],
                "values": values, "value_nulls": [v.astype(bool) for v in vnull]}


class GroupByStrings:
    """Oracle GROUP BY over one string key column (GPUGroupByString.cuh:39-785 semantics): each
    distinct byte string is numbered in order of first appearance with a Python dict (exact
    equality), and the numbers are grouped by the oracle table above as int64 keys; results()
    maps them back.  A NULL key (key_nulls) is the NULL group."""

    def __init__(self, aggs, value_type, out_types=None):
        self.ids = {}
        self.names = []
        self.g = GroupBy(aggs, [LONG], value_type, out_types)

    def process(self, strings, values=None, value_nulls=None, key_nulls=None, preds=()):
        ids = np.empty(len(strings), np.int64)
        for i, s in enumerate(strings):
            b = s.encode() if isinstance(s, str) else bytes(s)
            k = self.ids.get(b)
            if k is None:
                k = self.ids[b] = len(self.names)
                self.names.append(b)
            ids[i] = k
        self.g.process([ids], values, value_nulls, [key_nulls] if key_nulls is not None else None, preds=preds)

    def results(self):
        r = self.g.results()
        keys = [None if r["key_nulls"][0][i] else self.names[int(k)] for i, k in enumerate(r["keys"][0])]
        return {"count": r["count"], "keys": keys, "values": r["values"], "value_nulls": r["value_nulls"]}


def query_mt(nthreads, aggs, key_types, value_type, keys, values=None, value_nulls=None,
             key_nulls=None, preds=(), out_types=None) -> GroupBy:
    """Threaded oracle run (CPU baseline): one table per thread, merged."""
    aggs = [aggs] if isinstance(aggs, int) else list(aggs)
    key_types = [key_types] if isinstance(key_types, int) else list(key_types)
    keys = list(keys) if isinstance(keys, (list, tuple)) else [keys]
    na, nk = len(aggs), len(key_types)
    ot = (C.c_int32 * na)(*(out_types or [0] * na))
    pa = (OrPred * max(1, len(preds)))(*preds)
    h = lib().or_gb_query_mt(nthreads, nk, (C.c_int32 * nk)(*key_types), value_type, na,
                             (C.c_int32 * na)(*aggs), ot, _vp_arr(keys),
                             _vp_arr(key_nulls) if key_nulls is not None else None,
                             _p(values), _p(value_nulls), len(preds), pa, keys[0].size)
    if not h:
        raise RuntimeError("oracle threaded run failed")
    return GroupBy(aggs, key_types, value_type, out_types, handle=h)


def _key_bits(k: np.ndarray) -> np.ndarray:
    """Integer view of a key column for sorting; float keys with -0.0 folded into +0.0."""
    k = np.asarray(k)
    if np.issubdtype(k.dtype, np.floating):
        k = k + k.dtype.type(0)  # -0.0 + 0.0 == +0.0; NaN payloads are kept
        return k.view(np.int32 if k.dtype.itemsize == 4 else np.int64)
    return k


def sort_groups(res: dict) -> dict:
    """The groups of a results dict (oracle or GPU) ordered by key, NULL keys first within each
    column, so that two results compare row by row whatever their slot order (S10)."""
    g = res["count"]
    nk = len(res["keys"])
    cols = []
    for c in range(nk):
        kn = np.asarray(res["key_nulls"][c], bool)[:g]
        kb = np.where(kn, 0, _key_bits(res["keys"][c][:g]))
        cols += [kb, ~kn]
    order = np.lexsort(cols[::-1]) if g else np.zeros(0, np.int64)
    out = {"count": g,
           "keys": [np.where(np.asarray(res["key_nulls"][c], bool)[:g], 0, _key_bits(res["keys"][c][:g]))[order]
                    for c in range(nk)],
           "key_nulls": [np.asarray(res["key_nulls"][c], bool)[:g][order] for c in range(nk)],
           "values": [], "value_nulls": []}
    for a in range(len(res["values"])):
        v = res["values"][a]
        vn = res["value_nulls"][a]
        out["values"].append(None if v is None else np.asarray(v)[:g][order])
        out["value_nulls"].append(np.zeros(g, bool) if vn is None else np.asarray(vn, bool)[:g][order])
    return out


def compare_groups(got: dict, want: dict, aggs=None, rel_aggs=(), rel: float = 0.0) -> None:
    """Every group of `got` equals the group of `want` with the same key: the same key set,
    the same NULL flags, values bit-exact (integers) or within `rel` relative (aggregates in
    rel_aggs).  Vectorised, for results of millions of groups."""
    assert got["count"] == want["count"], f"group counts differ: {got['count']} vs {want['count']}"
    a, b = sort_groups(got), sort_groups(want)
    for c in range(len(a["keys"])):
        bad = np.flatnonzero((a["keys"][c] != b["keys"][c]) | (a["key_nulls"][c] != b["key_nulls"][c]))
        assert bad.size == 0, f"key column {c}: {bad.size} groups differ, first at sorted row {bad[0]}"
    naggs = len(b["values"]) if aggs is None else len(aggs)
    for i in range(naggs):
        if b["values"][i] is None:
            continue
        vn_a, vn_b = a["value_nulls"][i], b["value_nulls"][i]
        bad = np.flatnonzero(vn_a != vn_b)
        assert bad.size == 0, f"aggregate {i}: NULL flags differ in {bad.size} groups"
        live = ~vn_b
        x = np.asarray(a["values"][i])[live]
        y = np.asarray(b["values"][i])[live]
        if i in rel_aggs and rel > 0:
            x = x.astype(np.float64)
            y = y.astype(np.float64)
            err = np.abs(x - y) > rel * np.maximum(np.abs(y), 1e-30)
            bad = np.flatnonzero(err)
            assert bad.size == 0, (f"aggregate {i}: {bad.size} groups outside rel {rel}, "
                                   f"e.g. {x[bad[0]]} vs {y[bad[0]]}")
        elif np.issubdtype(x.dtype, np.floating) or np.issubdtype(y.dtype, np.floating):
            x = x.astype(np.float64)
            y = y.astype(np.float64)
            bad = np.flatnonzero((x != y) & ~(np.isnan(x) & np.isnan(y)))
            assert bad.size == 0, f"aggregate {i}: {bad.size} groups differ, e.g. {x[bad[0]]} vs {y[bad[0]]}"
        else:
            bad = np.flatnonzero(x.astype(np.int64) != y.astype(np.int64))
            assert bad.size == 0, f"aggregate {i}: {bad.size} groups differ, e.g. {x[bad[0]]} vs {y[bad[0]]}"
