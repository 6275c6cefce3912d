"""Pins the CPU oracle (no GPU needed): reference known answers, closed forms, invariants and
brute force on tiny inputs.  Each check is chosen so that a plausible mistake in the oracle
(wrong operand conversion, dropped NULL rule, wrong AVG rounding, lost int32 wrap, wrong
index offset) fails it."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import AGG_NAMES, OP_NAMES, as_map, expect_map, golden, maps_close, pack_bits, unpack_bits

G = golden()


# ---- reconstruct ----------------------------------------------------------------------

@pytest.mark.parametrize("case", G["generate_indexes"], ids=lambda c: c["source"].split()[-1])
def test_generate_indexes_known_answers(case):
    mask = np.array(case["mask"], np.int8)
    assert O.generate_indexes(mask).tolist() == case["expect"]


def test_generate_indexes_alternating():
    n = G["generate_indexes_alternating"]["n"]
    mask = (np.arange(n) % 2).astype(np.int8)
    np.testing.assert_array_equal(O.generate_indexes(mask), np.arange(1, n, 2))


@pytest.mark.parametrize("kind", ["ones", "zeros", "alternating", "random"])
@pytest.mark.parametrize("inclusive", [True, False])
def test_prefix_sum_closed_form(kind, inclusive):
    n = 1 << 18
    mask = {"ones": np.ones(n, np.int8), "zeros": np.zeros(n, np.int8),
            "alternating": (np.arange(n) % 2).astype(np.int8),
            "random": np.random.default_rng(5).integers(0, 2, n).astype(np.int8)}[kind]
    got = O.prefix_sum(mask, inclusive)
    c = np.cumsum(mask.astype(np.int64))
    want = c if inclusive else c - mask
    np.testing.assert_array_equal(got, want)
    if kind == "ones":  # closed form: i + 1 / i
        np.testing.assert_array_equal(got, np.arange(n) + (1 if inclusive else 0))
    assert O.mask_sum(mask) == int(mask.sum())


def test_compact_with_nulls():
    rng = np.random.default_rng(7)
    n = 1000
    vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
    mask = rng.integers(0, 2, n).astype(np.int8)
    isnull = rng.random(n) < 0.3
    out, out_null, k = O.compact(vals, mask, pack_bits(isnull))
    sel = mask != 0
    np.testing.assert_array_equal(out, vals[sel])
    np.testing.assert_array_equal(unpack_bits(out_null, k), isnull[sel])


def test_compress_null_mask():
    rng = np.random.default_rng(3)
    for n in (1, 63, 64, 65, 1000):
        arr = rng.integers(0, 2, n).astype(np.uint8)
        np.testing.assert_array_equal(unpack_bits(O.compress_null_mask(arr), n), arr.astype(bool))


# ---- filter -----------------------------------------------------------------------------

@pytest.mark.parametrize("op", list(OP_NAMES)[:6])
@pytest.mark.parametrize("dt", [np.int32, np.int64, np.float32, np.float64])
def test_filter_same_type_closed_form(op, dt):
    # the reference's GPUFilterTests style: 2^18 random values, col-col and col-const
    rng = np.random.default_rng(11)
    n = 1 << 18
    if np.issubdtype(dt, np.integer):
        a = rng.integers(np.iinfo(dt).min, np.iinfo(dt).max, n, dtype=dt)
        b = rng.integers(np.iinfo(dt).min, np.iinfo(dt).max, n, dtype=dt)
        b[::7] = a[::7]
    else:
        a = rng.uniform(-32000, 32000, n).astype(dt)
        b = rng.uniform(-32000, 32000, n).astype(dt)
        b[::7] = a[::7]
    npop = {"GREATER": np.greater, "GREATER_EQUAL": np.greater_equal, "LESS": np.less,
            "LESS_EQUAL": np.less_equal, "EQUAL": np.equal, "NOT_EQUAL": np.not_equal}[op]
    t = O.TYPE_OF[np.dtype(dt)]
    got = O.filter_mask(OP_NAMES[op], a, b, n)
    np.testing.assert_array_equal(got, npop(a, b).astype(np.int8))
    c = a[12345]
    got = O.filter_mask(OP_NAMES[op], a, (c, t), n)
    np.testing.assert_array_equal(got, npop(a, c).astype(np.int8))
    got = O.filter_mask(OP_NAMES[op], (c, t), b, n)
    np.testing.assert_array_equal(got, npop(c, b).astype(np.int8))


def test_filter_usual_arithmetic_conversions():
    # int32 vs float compares in float (S2): 16777217 becomes 16777216.0f
    a = np.array([16777217, 16777216, 3], np.int32)
    got = O.filter_mask(OP_NAMES["EQUAL"], a, (16777216.0, O.FLOAT), 3)
    assert got.tolist() == [1, 1, 0]
    # ... but int32 vs double compares exactly
    got = O.filter_mask(OP_NAMES["EQUAL"], a, (16777216.0, O.DOUBLE), 3)
    assert got.tolist() == [0, 1, 0]
    # int64 vs int32 compares in int64 (no truncation of the wide operand)
    w = np.array([(1 << 40) + 5, 5], np.int64)
    got = O.filter_mask(OP_NAMES["EQUAL"], w, (5, O.INT), 2)
    assert got.tolist() == [0, 1]


def test_filter_null_rows_fail_and_logic_ops():
    a = np.array([5, 6, 7, 8], np.int32)
    nulls = pack_bits([False, True, False, True])
    got = O.filter_mask(OP_NAMES["GREATER"], a, (0, O.INT), 4, nulls)
    assert got.tolist() == [1, 0, 1, 0]
    m1 = np.array([1, 1, 0, 0], np.int8)
    m2 = np.array([1, 0, 1, 0], np.int8)
    assert O.filter_mask(OP_NAMES["LOGICAL_AND"], m1, m2, 4).tolist() == [1, 0, 0, 0]
    assert O.filter_mask(OP_NAMES["LOGICAL_OR"], m1, m2, 4).tolist() == [1, 1, 1, 0]
    assert O.logical_not(m1, pack_bits([False, False, True, False])).tolist() == [0, 0, 0, 1]
    # const-const fills the whole block
    assert O.filter_mask(OP_NAMES["LESS"], (1, O.INT), (2, O.LONG), 5).tolist() == [1] * 5


# ---- GROUP BY: reference known answers ------------------------------------------------------

def _run_single(case):
    keys = np.array(case["keys"], np.int32)
    vals = np.array(case["values"], np.int32)
    agg = AGG_NAMES[case["agg"]]
    g = O.GroupBy(agg, O.INT, O.INT)
    kn = [pack_bits(case["key_nulls"])] if "key_nulls" in case else None
    vn = pack_bits(case["value_nulls"]) if "value_nulls" in case else None
    preds = []
    if "where" in case:
        w = case["where"]
        col = keys if w["column"] == "keys" else np.array(case["filter"], np.int32)
        preds = [O.pred(OP_NAMES[w["op"]], col, w["const"], O.INT)]
    n, b = keys.size, case["block"]
    for s in range(0, n, b):
        g.process([keys], None if case.get("count_star") else vals, vn, kn, preds=preds, begin=s,
                  end=min(n, s + b))
    return as_map(g.results())


@pytest.mark.parametrize("case", G["group_by_single"], ids=lambda c: c["source"].split()[1])
def test_group_by_single_known_answers(case):
    maps_close(_run_single(case), expect_map(case["expect"], False))


def _run_multi(case):
    a = np.array(case["a"], np.int32)
    b = np.array(case["b"], np.int32)
    vals = np.array(case["values"], np.int32)
    g = O.GroupBy(AGG_NAMES[case["agg"]], [O.INT, O.INT], O.INT)
    kn = None
    if "a_nulls" in case:
        kn = [pack_bits(case["a_nulls"]), pack_bits(case["b_nulls"])]
    vn = pack_bits(case["value_nulls"]) if "value_nulls" in case else None
    n, blk = a.size, case["block"]
    for s in range(0, n, blk):
        g.process([a, b], None if case.get("count_star") else vals, vn, kn, begin=s, end=min(n, s + blk))
    return as_map(g.results())


@pytest.mark.parametrize("case", G["group_by_multi"], ids=lambda c: c["source"].split()[1])
def test_group_by_multi_known_answers(case):
    maps_close(_run_multi(case), expect_map(case["expect"], True))


def test_whole_column_float_sum_known_answer():
    case = G["whole_column"][0]
    v = np.full(case["n"], case["value"], np.float32)
    g = O.GroupBy(AGG_NAMES["SUM"], O.INT, O.FLOAT)
    g.process([np.zeros(case["n"], np.int32)], v)
    assert as_map(g.results()) == {(0,): case["sum"]}


# ---- GROUP BY: typing rules and invariants ---------------------------------------------------

def test_int32_sum_wraps_and_widens():
    k = np.zeros(2, np.int32)
    v = np.array([2**31 - 1, 2**31 - 1], np.int32)
    g = O.GroupBy(AGG_NAMES["SUM"], O.INT, O.INT)
    g.process([k], v)
    assert as_map(g.results()) == {(0,): -2}          # int32 wraps modulo 2^32 (S7)
    g = O.GroupBy(AGG_NAMES["SUM"], O.INT, O.INT, out_types=[O.LONG])
    g.process([k], v)
    assert as_map(g.results()) == {(0,): 2 * (2**31 - 1)}  # SUM(int64 acc)
    g = O.GroupBy(AGG_NAMES["AVG"], O.INT, O.INT)
    g.process([k], v)
    assert as_map(g.results()) == {(0,): -1}          # AVG divides the wrapped int32 sum


def test_avg_truncates_toward_zero():
    k = np.zeros(2, np.int32)
    g = O.GroupBy(AGG_NAMES["AVG"], O.INT, O.INT)
    g.process([k], np.array([-3, -4], np.int32))
    assert as_map(g.results()) == {(0,): -3}
    g = O.GroupBy(AGG_NAMES["AVG"], O.INT, O.FLOAT)
    g.process([k], np.array([-3, -4], np.float32))
    assert as_map(g.results()) == {(0,): -3.5}


def test_lowest_key_is_a_normal_group():
    # INT64_MIN collides with the reference's empty marker (quirk Q1); here it is a normal key
    k = np.array([-(2**63), -(2**63), 0], np.int64)
    g = O.GroupBy(AGG_NAMES["COUNT"], O.LONG, O.LONG)
    g.process([k], k)
    assert as_map(g.results()) == {(-(2**63),): 2, (0,): 1}


def _brute(keys, vals, vnull, sel):
    d = {}
    for i in range(keys.size):
        if not sel[i]:
            continue
        e = d.setdefault(int(keys[i]), [0, 0, None, None])
        if vnull[i]:
            continue
        v = int(vals[i])
        e[0] += v
        e[1] += 1
        e[2] = v if e[2] is None else min(e[2], v)
        e[3] = v if e[3] is None else max(e[3], v)
    return d


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_group_by_brute_force_tiny(seed):
    rng = np.random.default_rng(seed)
    n = 300
    keys = rng.integers(-5, 5, n).astype(np.int64)
    vals = rng.integers(-100, 100, n).astype(np.int64)
    vnull = rng.random(n) < 0.2
    filt = rng.integers(0, 10, n).astype(np.int32)
    sel = filt > 3
    aggs = [AGG_NAMES[a] for a in ("SUM", "COUNT", "MIN", "MAX", "AVG", "COUNT_ALL")]
    g = O.GroupBy(aggs, O.LONG, O.LONG)
    g.process([keys], vals, pack_bits(vnull), preds=[O.pred(OP_NAMES["GREATER"], filt, 3, O.INT)])
    res = g.results()
    want = _brute(keys, vals, vnull, sel)
    m = [as_map(res, a) for a in range(len(aggs))]
    assert set(m[0]) == {(k,) for k in want}
    for k, (s, c, mn, mx) in want.items():
        assert m[1][(k,)] == c
        assert m[5][(k,)] == int(np.sum(sel & (keys == k)))
        if c == 0:
            assert m[0][(k,)] is None and m[2][(k,)] is None and m[4][(k,)] is None
        else:
            assert m[0][(k,)] == s and m[2][(k,)] == mn and m[3][(k,)] == mx
            assert m[4][(k,)] == int(s / c)  # int() truncates toward zero, like C++ division


def test_threaded_equals_single():
    rng = np.random.default_rng(9)
    n = 200_000
    keys = rng.integers(0, 5000, n).astype(np.int64)
    vals = rng.lognormal(0, 1, n).astype(np.float32)
    vnull = pack_bits(rng.random(n) < 0.05)
    preds = [O.pred(OP_NAMES["GREATER_EQUAL"], vals, 0.5, O.FLOAT, vnull),
             O.pred(OP_NAMES["LESS_EQUAL"], vals, 4.0, O.FLOAT, vnull)]
    aggs = [AGG_NAMES["SUM"], AGG_NAMES["COUNT"]]
    g1 = O.GroupBy(aggs, O.LONG, O.FLOAT)
    g1.process([keys], vals, vnull, preds=preds)
    g8 = O.query_mt(8, aggs, O.LONG, O.FLOAT, [keys], vals, vnull, preds=preds)
    for a in range(2):
        maps_close(as_map(g8.results(), a), as_map(g1.results(), a), rel=1e-12)
    # invariant: the counts add up to the rows that pass the filter
    sel = (vals >= np.float32(0.5)) & (vals <= np.float32(4.0)) & ~unpack_bits(vnull, n)
    assert sum(as_map(g1.results(), 1).values()) == int(sel.sum())
    assert abs(sum(as_map(g1.results(), 0).values()) - float(vals[sel].astype(np.float64).sum())) < 1e-6 * float(vals[sel].sum())


def test_compare_groups_catches_planted_differences():
    """The all-groups checker used at full benchmark size: equal results pass whatever their
    order; a changed value, a changed key, a flipped NULL flag or an FP error above the
    tolerance each fail."""
    import copy
    rng = np.random.default_rng(3)
    n = 50_000
    keys = rng.integers(0, 3000, n).astype(np.int64)
    vals = rng.lognormal(0, 1, n).astype(np.float32)
    vn = rng.random(n) < 0.3
    g = O.GroupBy([O.AGG_SUM, O.AGG_COUNT, O.AGG_MAX], O.LONG, O.FLOAT)
    g.process([keys], vals, pack_bits(vn))
    a = g.results()
    perm = rng.permutation(a["count"])
    b = {"count": a["count"], "keys": [a["keys"][0][perm]], "key_nulls": [a["key_nulls"][0][perm]],
         "values": [v[perm] for v in a["values"]], "value_nulls": [v[perm] for v in a["value_nulls"]]}
    O.compare_groups(b, a, rel_aggs=(0,), rel=1e-5)
    for plant in ("value", "key", "null", "fp"):
        c = copy.deepcopy(b)
        if plant == "value":
            c["values"][1][7] += 1
        elif plant == "key":
            c["keys"][0][11] = 10_000
        elif plant == "null":
            c["value_nulls"][2][5] = ~c["value_nulls"][2][5]
        else:
            c["values"][0][3] *= 1 + 3e-5
        with pytest.raises(AssertionError):
            O.compare_groups(c, a, rel_aggs=(0,), rel=1e-5)


# ---- string keys ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", G["group_by_string"], ids=lambda c: c["source"].split()[1])
def test_group_by_string_known_answers(case):
    g = O.GroupByStrings(AGG_NAMES[case["agg"]], O.INT)
    for keys, vals in zip(case["blocks"], case["values"]):
        g.process(keys, np.array(vals, np.int32))
    r = g.results()
    got = {k.decode(): int(v) for k, v in zip(r["keys"], r["values"][0])}
    assert got == case["expect"]


def test_group_by_string_exact_bytes_nulls_and_where():
    # byte strings that differ only in a suffix, a prefix or a trailing NUL byte stay apart; the
    # empty string is a key of its own, distinct from the NULL key
    keys = [b"ab", b"ab\x00", b"a", b"", b"ab", b"", b"b", b"ab\x00"]
    vals = np.array([1, 2, 4, 8, 16, 32, 64, 128], np.int32)
    knull = pack_bits(np.array([0, 0, 0, 0, 0, 0, 1, 0], bool))
    g = O.GroupByStrings([O.AGG_SUM, O.AGG_COUNT], O.INT)
    g.process(keys, vals, key_nulls=knull, preds=[O.pred(1, vals, 2, O.INT)])  # value >= 2
    r = g.results()
    got = {k: (int(s), int(c)) for k, s, c in zip(r["keys"], r["values"][0], r["values"][1])}
    assert got == {b"ab\x00": (130, 2), b"a": (4, 1), b"": (40, 2), b"ab": (16, 1), None: (64, 1)}


def _sign_class(x: float):
    """(is NaN, sign bit, value) of a float: equal for two floats that are the same value,
    telling -0.0 from +0.0 and NaNs apart by sign only."""
    x = float(x)
    return (math.isnan(x), math.copysign(1.0, x) < 0, None if math.isnan(x) else x)


def _total_min_max(vals):
    """MIN and MAX of a group's float values under IEEE totalOrder, stated through numpy's
    nanmin/nanmax and the signs of the NaNs and zeros present: a positive NaN is the largest
    value, a negative NaN the smallest, and -0.0 lies below +0.0."""
    v = np.asarray(vals, np.float64)
    nan, neg = np.isnan(v), np.signbit(v)
    num = v[~nan]
    if (nan & neg).any():
        mn = -math.nan
    elif num.size:
        mn = float(np.nanmin(num))
        if mn == 0.0:
            mn = -0.0 if (neg & (v == 0)).any() else 0.0
    else:
        mn = math.nan
    if (nan & ~neg).any():
        mx = math.nan
    elif num.size:
        mx = float(np.nanmax(num))
        if mx == 0.0:
            mx = 0.0 if (~neg & (v == 0)).any() else -0.0
    else:
        mx = -math.nan
    return mn, mx


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_nan_and_signed_zero_rules(dt, seed):
    """N1: -0.0 and +0.0 keys form one group and every NaN key (any sign or payload) one group,
    grouped here by Python's own float equality with NaN mapped to one marker; MIN/MAX follow
    totalOrder (closed form above); SUM is NaN exactly when numpy's FP64 sum is."""
    rng = np.random.default_rng(seed)
    qnan_payload = np.array([0x7FC12345], np.uint32).view(np.float32)[0]
    key_pool = np.array([0.0, -0.0, 1.5, -1.5, np.nan, -np.nan, qnan_payload, np.inf, -np.inf], dt)
    val_pool = np.array([1.0, -1.0, 0.0, -0.0, np.nan, -np.nan, np.inf, -np.inf, 2.5, -7.25], dt)
    n = 400
    keys = key_pool[rng.integers(0, key_pool.size, n)]
    vals = val_pool[rng.integers(0, val_pool.size, n)]
    vnull = rng.random(n) < 0.1
    ft = O.FLOAT if dt == np.float32 else O.DOUBLE
    aggs = [AGG_NAMES[a] for a in ("MIN", "MAX", "SUM", "COUNT")]
    g = O.GroupBy(aggs, ft, ft)
    g.process([keys], vals, pack_bits(vnull))
    r = g.results()
    want = {}
    for k, v, isnull in zip(keys.tolist(), vals.tolist(), vnull.tolist()):
        e = want.setdefault("nan" if math.isnan(k) else k, [])
        if not isnull:
            e.append(v)
    assert r["count"] == len(want)
    seen = set()
    for i in range(r["count"]):
        k = float(r["keys"][0][i])
        gk = "nan" if math.isnan(k) else k
        assert gk not in seen
        seen.add(gk)
        assert not math.copysign(1.0, k) < 0 or k != 0.0, "a zero key is +0.0"
        if math.isnan(k):
            assert not math.copysign(1.0, k) < 0, "the NaN key is the positive quiet NaN"
        gv = want[gk]
        assert int(r["values"][3][i]) == len(gv)
        if not gv:
            assert r["value_nulls"][0][i] and r["value_nulls"][1][i]
            continue
        mn, mx = _total_min_max(gv)
        assert _sign_class(r["values"][0][i]) == _sign_class(mn), (gk, gv)
        assert _sign_class(r["values"][1][i]) == _sign_class(mx), (gk, gv)
        s = float(np.sum(np.asarray(gv, np.float64)))
        assert math.isnan(float(r["values"][2][i])) == math.isnan(s)
        if not math.isnan(s):
            assert float(r["values"][2][i]) == pytest.approx(s, rel=1e-6)
