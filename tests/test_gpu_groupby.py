"""GPU parity of GROUP BY against the oracle: the reference's known answers, random inputs over
every key/value type and aggregate, NULL handling, composite keys, table-full behaviour,
and the benchmark workloads (bit-exact integers; FP32 SUM/AVG within 1e-5 relative of FP64)."""
import numpy as np
import pytest

import oracle as O
import qikkdb_amd as qk
from qikkdb_amd import datagen
from gpu_util import Where, col, cpu_group_by, gpu_group_by
from helpers import AGG_NAMES, as_map, expect_map, golden, maps_close

pytestmark = pytest.mark.gpu
G = golden()
FP_REL = 1e-5  # FP32 SUM/AVG tolerance against the FP64 oracle


def compare(gpu, cpu, naggs, real_aggs=()):
    for a in range(naggs):
        maps_close(as_map(gpu, a), as_map(cpu, a), rel=FP_REL if a in real_aggs else 0.0)


# ---- reference known answers ----------------------------------------------------------------

@pytest.mark.parametrize("case", G["group_by_single"], ids=lambda c: c["source"].split()[1])
def test_known_answers_single(case):
    keys = np.array(case["keys"], np.int32)
    vals = np.array(case["values"], np.int32)
    kn = [np.array(case["key_nulls"])] if "key_nulls" in case else None
    vn = np.array(case["value_nulls"]) if "value_nulls" in case else None
    where, extra = (), None
    if "where" in case:
        w = case["where"]
        if w["column"] == "keys":
            where = [Where(getattr(qk, w["op"]), "key0", w["const"], qk.INT)]
        else:
            extra = {"filter": (np.array(case["filter"], np.int32), None)}
            where = [Where(getattr(qk, w["op"]), "filter", w["const"], qk.INT)]
    agg = AGG_NAMES[case["agg"]]
    res = gpu_group_by(agg, [keys], None if case.get("count_star") else vals, kn,
                       None if case.get("count_star") else vn, where=where, block=case["block"],
                       capacity=262144, extra_cols=extra, value_type=qk.INT)
    maps_close(as_map(res), expect_map(case["expect"], False))


@pytest.mark.parametrize("case", G["group_by_multi"], ids=lambda c: c["source"].split()[1])
def test_known_answers_multi(case):
    a = np.array(case["a"], np.int32)
    b = np.array(case["b"], np.int32)
    vals = np.array(case["values"], np.int32)
    kn = [np.array(case["a_nulls"]), np.array(case["b_nulls"])] if "a_nulls" in case else None
    vn = np.array(case["value_nulls"]) if "value_nulls" in case else None
    res = gpu_group_by(AGG_NAMES[case["agg"]], [a, b], None if case.get("count_star") else vals, kn,
                       None if case.get("count_star") else vn, block=case["block"],
                       capacity=case.get("max_hash_count", 262144), value_type=qk.INT)
    maps_close(as_map(res), expect_map(case["expect"], True))


def test_whole_column_float_sum_known_answer():
    case = G["whole_column"][0]
    v = np.full(case["n"], case["value"], np.float32)
    res = gpu_group_by(qk.AGG_SUM, [np.zeros(case["n"], np.int32)], v)
    assert as_map(res) == {(0,): case["sum"]}


# ---- random parity over types ----------------------------------------------------------------

ALL_AGGS = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_AVG, qk.AGG_COUNT, qk.AGG_COUNT_ALL]


def make_keys(dt, n, card, rng):
    if np.issubdtype(dt, np.integer):
        base = rng.integers(0, card, n)
        return (base * 7919 - card * 3000).astype(dt)
    return (rng.integers(0, card, n) * 0.25 - card / 8).astype(dt)


def make_vals(dt, n, rng):
    if np.issubdtype(dt, np.integer):
        return rng.integers(-10**6, 10**6, n).astype(dt)
    return rng.lognormal(0, 1, n).astype(dt)


@pytest.mark.parametrize("kt", [np.int32, np.int64, np.float32, np.float64])
@pytest.mark.parametrize("vt", [np.int32, np.int64, np.float32, np.float64])
@pytest.mark.parametrize("card", [7, 3000, 200_000])
def test_random_types(kt, vt, card):
    rng = np.random.default_rng(card + 17 * np.dtype(kt).itemsize + np.dtype(vt).itemsize)
    n = 400_000
    keys = make_keys(kt, n, card, rng)
    vals = make_vals(vt, n, rng)
    kn = [rng.random(n) < 0.02]
    vn = rng.random(n) < 0.1
    aggs = [qk.AGG_COUNT, qk.AGG_SUM, qk.AGG_AVG, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT_ALL]
    real = (1, 2) if np.issubdtype(vt, np.floating) else ()
    gpu = gpu_group_by(aggs, [keys], vals, kn, vn, block=1 << 17, capacity=card * 2 + 16)
    cpu = cpu_group_by(aggs, [keys], vals, kn, vn)
    compare(gpu, cpu, len(aggs), real)


@pytest.mark.parametrize("lds", [0, 4096, -1])
def test_lds_stage_does_not_change_results(lds):
    rng = np.random.default_rng(lds + 5)
    n = 1_000_000
    keys = datagen.zipf_keys(11, 1.1, 50_000, n)
    vals = rng.integers(-1000, 1000, n).astype(np.int64)
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]
    tuning = None if lds < 0 else {"lds_bytes": lds}
    gpu = gpu_group_by(aggs, [keys], vals, capacity=100_000, tuning=tuning)
    cpu = cpu_group_by(aggs, [keys], vals)
    compare(gpu, cpu, len(aggs))


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_block_shapes(block):
    rng = np.random.default_rng(block)
    n = 300_001
    keys = rng.integers(0, 5000, n).astype(np.int32)
    vals = rng.integers(-1000, 1000, n).astype(np.int32)
    aggs = [qk.AGG_SUM, qk.AGG_COUNT]
    gpu = gpu_group_by(aggs, [keys], vals, capacity=10_000, tuning={"block": block})
    compare(gpu, cpu_group_by(aggs, [keys], vals), 2)


def test_masked_and_where_forms_agree_with_oracle():
    rng = np.random.default_rng(21)
    n = 500_000
    keys = rng.integers(0, 1000, n).astype(np.int32)
    vals = rng.integers(-1000, 1001, n).astype(np.int32)
    other = rng.integers(-1000, 1001, n).astype(np.int32)
    mask = (vals > 0).astype(np.int8)
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]
    cpu = cpu_group_by(aggs, [keys], vals, mask=mask)
    compare(gpu_group_by(aggs, [keys], vals, mask=mask, capacity=2000), cpu, 4)
    w = [Where(qk.GREATER, "value", 0, qk.INT)]
    compare(gpu_group_by(aggs, [keys], vals, where=w, capacity=2000), cpu, 4)
    # column-vs-column predicate on a column that is not the value
    w2 = [Where(qk.LESS, "other", None, None, rhs="value")]
    extra = {"other": (other, None)}
    compare(gpu_group_by(aggs, [keys], vals, where=w2, extra_cols=extra, capacity=2000),
            cpu_group_by(aggs, [keys], vals, where=w2, extra_cols=extra), 4)


def test_edge_keys_sentinel_and_signed_zero():
    keys = np.array([np.iinfo(np.int64).min] * 3 + [0, -1, np.iinfo(np.int64).max], np.int64)
    vals = np.arange(6, dtype=np.int64)
    gpu = gpu_group_by([qk.AGG_SUM, qk.AGG_COUNT], [keys], vals, capacity=4)
    assert as_map(gpu, 0) == {(np.iinfo(np.int64).min,): 3, (0,): 3, (-1,): 4, (np.iinfo(np.int64).max,): 5}
    fk = np.array([0.0, -0.0, 1.5, -1.5], np.float32)
    gpu = gpu_group_by(qk.AGG_COUNT, [fk], fk, capacity=4)
    assert as_map(gpu) == {(0.0,): 2, (1.5,): 1, (-1.5,): 1}


def test_table_full_is_reported_and_capacity_independent():
    rng = np.random.default_rng(4)
    n = 100_000
    keys = rng.integers(0, 5000, n).astype(np.int32)
    vals = rng.integers(0, 10, n).astype(np.int32)
    with pytest.raises(qk.QueryEngineError) as e:
        gpu_group_by(qk.AGG_SUM, [keys], vals, capacity=1000)
    assert e.value.code == qk.QK_HASH_TABLE_FULL
    ref = as_map(cpu_group_by(qk.AGG_SUM, [keys], vals))
    for cap in (5000, 6000, 1 << 20):
        maps_close(as_map(gpu_group_by(qk.AGG_SUM, [keys], vals, capacity=cap)), ref)


def test_max_hash_count_and_empty_results():
    gb = qk.GPUGroupBy(qk.AGG_SUM, qk.INT, qk.INT, 262144, 2)
    assert gb.GetMaxHashCount() == 262144 * 2 + 1
    r = gb.GetResults()
    assert r.count == 0 and r.keys[0] is None and r.values[0] is None
    gb.close()
    gb = qk.GPUGroupBy(qk.AGG_SUM, [qk.INT, qk.INT], qk.INT, 8, 1)
    assert gb.GetMaxHashCount() == 8
    gb.close()


def test_type_rules_rejected_like_reference():
    with pytest.raises(qk.QueryEngineError):
        qk.GPUGroupBy(qk.AGG_COUNT, qk.INT, qk.INT, 16, 1, out_types=[qk.INT])
    with pytest.raises(qk.QueryEngineError):
        qk.GPUGroupBy(qk.AGG_MIN, qk.INT, qk.INT, 16, 1, out_types=[qk.LONG])


# ---- composite keys ----------------------------------------------------------------------------

@pytest.mark.parametrize("path", [None, 1])
@pytest.mark.parametrize("flags", [0, qk.GB_KEYS_NOT_NULL])
def test_composite_int_int(flags, path):
    # non-NULL (int32, int32) keys pack into one word and, forced to the two-phase path, are
    # read by its two-column key loads; negative second fields fill the word's high half
    rng = np.random.default_rng(flags + 3)
    n = 600_000
    a = rng.integers(0, 300, n).astype(np.int32)
    b = rng.integers(-300, 300, n).astype(np.int32)
    vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
    kn = None if flags else [rng.random(n) < 0.05, rng.random(n) < 0.05]
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT, qk.AGG_AVG]
    gpu = gpu_group_by(aggs, [a, b], vals, kn, capacity=200_000, flags=flags, path=path)
    cpu = cpu_group_by(aggs, [a, b], vals, kn)
    compare(gpu, cpu, len(aggs))


@pytest.mark.parametrize("mode", ["mask", "value_range", "mixed_where"])
def test_composite_pair_key_two_phase_filters(mode):
    rng = np.random.default_rng(17)
    n = 900_000
    a = rng.integers(-2**31, 2**31, n).astype(np.int32) % 5000
    b = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    b[::7] = rng.integers(-50, 50, b[::7].size)  # some keys repeat
    vals = rng.integers(-10**6, 10**6, n).astype(np.int32)
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]
    kw = dict(capacity=1_000_000, flags=qk.GB_KEYS_NOT_NULL, block=1 << 18)
    if mode == "mask":
        m = (rng.random(n) < 0.4).astype(np.int8)
        gpu = gpu_group_by(aggs, [a, b], vals, mask=m, path=1, **kw)
        cpu = cpu_group_by(aggs, [a, b], vals, mask=m)
    else:
        w = [Where(qk.GREATER, "value", -200_000, qk.INT), Where(qk.LESS_EQUAL, "value", 500_000, qk.INT)]
        if mode == "mixed_where":
            w.append(Where(qk.NOT_EQUAL, "key0", 3, qk.INT))
        gpu = gpu_group_by(aggs, [a, b], vals, where=w, path=1, **kw)
        cpu = cpu_group_by(aggs, [a, b], vals, where=w)
    compare(gpu, cpu, len(aggs))


def test_composite_wide_keys():
    rng = np.random.default_rng(8)
    n = 300_000
    k1 = (rng.integers(0, 50, n) * (1 << 40)).astype(np.int64)
    k2 = rng.integers(0, 50, n).astype(np.float64)
    k3 = rng.integers(-5, 5, n).astype(np.int32)
    vals = rng.lognormal(0, 1, n).astype(np.float64)
    kn = [rng.random(n) < 0.02, None, rng.random(n) < 0.02]
    aggs = [qk.AGG_SUM, qk.AGG_COUNT, qk.AGG_MAX]
    gpu = gpu_group_by(aggs, [k1, k2, k3], vals, kn, capacity=50_000)
    cpu = cpu_group_by(aggs, [k1, k2, k3], vals, kn)
    compare(gpu, cpu, len(aggs), real_aggs=(0,))


# ---- benchmark workloads -----------------------------------------------------------------------

def test_workload_small_full_size_bit_exact():
    d = datagen.small(1 << 20, seed=1)
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]
    w = [Where(qk.GREATER, "value", 0, qk.INT)]
    out_types = [qk.LONG, 0, 0, 0]
    gpu = gpu_group_by(aggs, [d["key"]], d["value"], where=w, capacity=1000, out_types=out_types)
    cpu = cpu_group_by(aggs, [d["key"]], d["value"], where=w, out_types=out_types)
    assert gpu["count"] == 1000
    compare(gpu, cpu, 4)


def _unpack(words, n):
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)


@pytest.mark.parametrize("n", [1 << 22])
def test_workload_zipf_fp_tolerance(n):
    d = datagen.zipf(n, seed=2)
    vn = _unpack(d["value_nulls"], n)
    aggs = [qk.AGG_SUM, qk.AGG_AVG, qk.AGG_COUNT]
    w = [Where(qk.GREATER_EQUAL, "value", 0.5, qk.FLOAT), Where(qk.LESS_EQUAL, "value", 4.0, qk.FLOAT)]
    gpu = gpu_group_by(aggs, [d["key"]], d["value"], None, vn, where=w, block=1 << 20,
                       capacity=1_000_000, flags=qk.GB_KEYS_NOT_NULL)
    cpu = cpu_group_by(aggs, [d["key"]], d["value"], None, vn, where=w)
    compare(gpu, cpu, 3, real_aggs=(0, 1))


def test_workload_composite_bit_exact():
    n = 1 << 23
    d = datagen.composite(n, seed=3)
    ts = d["ts"]
    w = [Where(qk.GREATER_EQUAL, "ts", datagen.WINDOW_START_MS, qk.LONG),
         Where(qk.LESS, "ts", datagen.WINDOW_END_MS, qk.LONG)]
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]
    extra = {"ts": (ts, None)}
    gpu = gpu_group_by(aggs, [d["a"], d["b"]], d["value"], where=w, block=1 << 22, capacity=1 << 24,
                       flags=qk.GB_KEYS_NOT_NULL, extra_cols=extra)
    cpu = cpu_group_by(aggs, [d["a"], d["b"]], d["value"], where=w, extra_cols=extra)
    compare(gpu, cpu, 4)


# ---- both row paths: two-phase partitioned and one-pass ------------------------------------

@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("kt,vt", [(np.int64, np.float32), (np.int32, np.int32), (np.float64, np.int64),
                                   (np.int64, np.float64), (np.float32, np.float32)])
def test_paths_agree_with_oracle(path, kt, vt):
    rng = np.random.default_rng(np.dtype(kt).itemsize * 10 + np.dtype(vt).itemsize + path)
    n = 700_001
    if np.issubdtype(kt, np.integer):
        keys = datagen.zipf_keys(5, 1.1, 40_000, n).astype(kt)
    else:
        keys = (datagen.zipf_keys(5, 1.1, 40_000, n) % 100_000).astype(kt) * 0.5
    vals = make_vals(vt, n, rng)
    kn = [rng.random(n) < 0.01]
    vn = rng.random(n) < 0.05
    aggs = [qk.AGG_SUM, qk.AGG_COUNT, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_AVG, qk.AGG_COUNT_ALL]
    real = (0, 4) if np.issubdtype(vt, np.floating) else ()
    gpu = gpu_group_by(aggs, [keys], vals, kn, vn, block=1 << 18, capacity=100_000, path=path)
    cpu = cpu_group_by(aggs, [keys], vals, kn, vn)
    compare(gpu, cpu, len(aggs), real)


@pytest.mark.parametrize("path", [0, 1])
def test_paths_where_mask_and_count_only(path):
    rng = np.random.default_rng(31 + path)
    n = 1_500_000
    keys = datagen.zipf_keys(9, 1.1, 200_000, n)
    vals = rng.lognormal(0, 1, n).astype(np.float32)
    vn = rng.random(n) < 0.05
    w = [Where(qk.GREATER_EQUAL, "value", 0.5, qk.FLOAT), Where(qk.LESS_EQUAL, "value", 4.0, qk.FLOAT)]
    aggs = [qk.AGG_SUM, qk.AGG_AVG, qk.AGG_COUNT]
    compare(gpu_group_by(aggs, [keys], vals, None, vn, where=w, capacity=300_000, path=path,
                         flags=qk.GB_KEYS_NOT_NULL),
            cpu_group_by(aggs, [keys], vals, None, vn, where=w), 3, real_aggs=(0, 1))
    mask = (rng.random(n) < 0.3).astype(np.int8)
    compare(gpu_group_by([qk.AGG_COUNT, qk.AGG_COUNT_ALL], [keys], None, None, None, mask=mask,
                         capacity=300_000, path=path, value_type=qk.INT),
            cpu_group_by([qk.AGG_COUNT, qk.AGG_COUNT_ALL], [keys], None, None, None, mask=mask,
                         value_type=qk.INT), 2)


def test_two_phase_uniform_high_cardinality_and_sentinel():
    # no heavy hitters: every row goes through the partitions; the empty-marker key and NULL
    # keys go through the side slots
    rng = np.random.default_rng(77)
    n = 2_000_000
    keys = rng.integers(-(2**62), 2**62, n // 4).astype(np.int64)[rng.integers(0, n // 4, n)]
    keys[::1000] = np.iinfo(np.int64).min
    kn = [rng.random(n) < 0.001]
    vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]
    gpu = gpu_group_by(aggs, [keys], vals, kn, None, block=1 << 20, capacity=n, path=1)
    cpu = cpu_group_by(aggs, [keys], vals, kn, None)
    compare(gpu, cpu, 4)


def expected_layout(capacity, sum_only):
    """(sub-tables, cold-row partitions) of a two-phase table: sub-tables of at most 160 KB of
    owned LDS state (20 B per slot for SUM/AVG/COUNT, 36 B with MIN/MAX), at most 1024
    partitions of up to 16 sub-tables each; otherwise 256 sub-tables, one per partition."""
    cap = 64
    while cap < 2 * capacity:
        cap *= 2
    per_slot = 20 if sum_only else 36
    s = 1
    while 2 * s * per_slot <= 160 * 1024:
        s *= 2
    f = max(256, cap // s)
    return (f, min(f, 1024)) if f <= 16384 else (256, 256)


@pytest.mark.parametrize("flags", [0, qk.GB_ASYNC])
@pytest.mark.parametrize("sum_only,capacity", [(False, 5_000_000), (True, 5_000_000), (True, 40_000_000),
                                               (False, 40_000_000), (True, 100_000_000)])
def test_two_phase_large_tables(flags, sum_only, capacity):
    """Tables whose 256 sub-tables are too large for the LDS.  Up to 16 sub-tables per
    cold-row partition, phase 2 runs one workgroup per sub-table that keeps its own rows of the
    partition (2^24 slots: 2 or 4 per partition; 2^27 slots of SUM/AVG/COUNT: 16); beyond that
    it aggregates in an LDS hash table and merges into the partition's sub-table (2^27 slots
    with MIN/MAX, 2^28 slots).  Heavy hitters, a long tail and NULL keys."""
    rng = np.random.default_rng(91)
    n = 2_000_000
    keys = np.where(rng.random(n) < 0.5, datagen.zipf_keys(13, 1.1, 5000, n),
                    rng.integers(-(2**40), 2**40, n)).astype(np.int64)
    kn = [rng.random(n) < 0.001]
    if sum_only:
        vals = rng.lognormal(0, 1, n).astype(np.float32)
        aggs, real = [qk.AGG_SUM, qk.AGG_AVG, qk.AGG_COUNT], (0, 1)
    else:
        vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
        aggs, real = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT], ()
    st = {}
    gpu = gpu_group_by(aggs, [keys], vals, kn, None, block=1 << 19, capacity=capacity, path=1, flags=flags,
                       stats=st)
    assert (st["sub_tables"], st["partitions"]) == expected_layout(capacity, sum_only)
    cpu = cpu_group_by(aggs, [keys], vals, kn, None)
    compare(gpu, cpu, len(aggs), real_aggs=real)


def test_two_phase_small_table_all_heavy_over_blocks():
    d = datagen.small(1 << 21, seed=4)
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]
    w = [Where(qk.GREATER, "value", 0, qk.INT)]
    gpu = gpu_group_by(aggs, [d["key"]], d["value"], where=w, block=1 << 19, capacity=1000,
                       out_types=[qk.LONG, 0, 0, 0], path=1)
    cpu = cpu_group_by(aggs, [d["key"]], d["value"], where=w, out_types=[qk.LONG, 0, 0, 0])
    compare(gpu, cpu, 4)


def test_two_phase_table_full_reported():
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 50_000, 2_000_000).astype(np.int32)
    with pytest.raises(qk.QueryEngineError) as e:
        gpu_group_by(qk.AGG_COUNT, [keys], None, capacity=10_000, path=1, value_type=qk.INT)
    assert e.value.code == qk.QK_HASH_TABLE_FULL


@pytest.mark.parametrize("defer_rows", [-1, 0, 3 << 18, 1 << 18])
def test_two_phase_deferred_merge(defer_rows):
    # an asynchronous table keeps the partitioned rows of several blocks and merges them
    # when read; a budget that fills mid-query forces merges between blocks
    rng = np.random.default_rng(41)
    n = 3_000_000
    keys = datagen.zipf_keys(12, 1.1, 150_000, n)
    vals = rng.lognormal(0, 1, n).astype(np.float32)
    vn = rng.random(n) < 0.05
    kn = [rng.random(n) < 0.002]
    w = [Where(qk.GREATER_EQUAL, "value", 0.5, qk.FLOAT), Where(qk.LESS_EQUAL, "value", 4.0, qk.FLOAT)]
    aggs = [qk.AGG_SUM, qk.AGG_AVG, qk.AGG_COUNT, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT_ALL]
    gpu = gpu_group_by(aggs, [keys], vals, kn, vn, where=w, block=1 << 18, capacity=200_000,
                       path=1, flags=qk.GB_ASYNC, defer_rows=defer_rows)
    cpu = cpu_group_by(aggs, [keys], vals, kn, vn, where=w)
    compare(gpu, cpu, len(aggs), real_aggs=(0, 1, 3, 4))


def test_two_phase_reset_drops_pending_rows():
    rng = np.random.default_rng(43)
    n = 1 << 20
    ka = rng.integers(0, 5000, n).astype(np.int64)
    kb = rng.integers(10_000, 12_000, n).astype(np.int64)
    va = rng.integers(-1000, 1000, n).astype(np.int64)
    vb = rng.integers(-1000, 1000, n).astype(np.int64)
    aggs = [qk.AGG_SUM, qk.AGG_COUNT]
    gb = qk.GPUGroupBy(aggs, qk.LONG, qk.LONG, 20_000, 1, flags=qk.GB_ASYNC)
    gb.set_path(1, 0)
    bufs = []
    for k, v in ((ka, va), (kb, vb)):
        kc, vc = col(k), col(v)
        bufs.append((kc, vc))
        gb.ProcessBlock([kc.data], vc.data, n)
        if k is ka:
            gb.reset()
    got = gb.results_host()
    gb.close()
    compare(got, cpu_group_by(aggs, [kb], vb), 2)


# WHERE clauses on the value column alone are folded into one interval per comparison
# domain; each case crosses a conversion boundary of the usual arithmetic conversions
FOLD_CASES = [
    # (value dtype, [(op, const, const type), ...])
    (np.int32, [(qk.GREATER, 100, qk.INT), (qk.LESS_EQUAL, 5000, qk.LONG)]),
    (np.int32, [(qk.GREATER_EQUAL, 16777217, qk.FLOAT), (qk.LESS, 2.5e7, qk.DOUBLE)]),
    (np.int32, [(qk.EQUAL, 777, qk.INT)]),
    (np.int64, [(qk.GREATER, 2**53 + 1, qk.DOUBLE), (qk.LESS, 2**62, qk.LONG)]),
    (np.int64, [(qk.GREATER, 3.0e15, qk.FLOAT)]),
    (np.int64, [(qk.GREATER, 10, qk.INT), (qk.LESS, 5, qk.INT)]),          # empty
    (np.float32, [(qk.GREATER, 0.1, qk.DOUBLE), (qk.LESS_EQUAL, 0.1, qk.FLOAT)]),
    (np.float32, [(qk.GREATER_EQUAL, 0.5, qk.FLOAT), (qk.LESS_EQUAL, 4.0, qk.FLOAT)]),
    (np.float32, [(qk.GREATER, 0, qk.INT), (qk.GREATER, -1, qk.LONG)]),
    (np.float32, [(qk.EQUAL, 0.0, qk.FLOAT)]),
    (np.float64, [(qk.GREATER, 2**60 + 1, qk.LONG), (qk.LESS_EQUAL, 1e300, qk.DOUBLE)]),
    (np.float64, [(qk.GREATER_EQUAL, 1.0, qk.FLOAT), (qk.GREATER, 1.0, qk.DOUBLE)]),
]


def fold_values(vt, n, rng):
    if vt == np.int32:
        v = rng.integers(-3000, 30_000_000, n).astype(np.int64)
        v[::7] = 16777216 + rng.integers(-3, 4, v[::7].size)
        v[::11] = 777
        return v.astype(np.int32)
    if vt == np.int64:
        v = rng.integers(-2**62, 2**62, n)
        v[::5] = 2**53 + rng.integers(-3, 4, v[::5].size)
        v[::9] = 3_000_000_000_000_000 + rng.integers(-2**20, 2**20, v[::9].size)
        v[::13] = rng.integers(0, 20, v[::13].size)
        return v.astype(np.int64)
    if vt == np.float32:
        v = (rng.random(n) * 8 - 2).astype(np.float32)
        v[::5] = np.float32(0.1)
        v[::7] = np.nextafter(np.float32(0.1), np.float32(1))
        v[::11] = 0.0
        v[::13] = -0.0
        v[::17] = np.nan
        v[::19] = np.float32(4.0)
        return v
    v = rng.random(n) * 4 - 1
    v[::5] = float(2**60)
    v[::7] = np.nextafter(float(2**60), 2.0**61)
    v[::11] = 1.0
    v[::13] = np.nan
    v[::17] = np.inf
    return v.astype(np.float64)


COLUMN_FOLD_CASES = [c for c in FOLD_CASES if c[0] in (np.int32, np.int64)]


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("case", range(len(COLUMN_FOLD_CASES)))
def test_column_range_fold_matches_predicates(case, path):
    """WHERE on another integer column than the aggregated one (the folded column interval):
    same groups as the predicates evaluated one by one, NULLs of that column failing them."""
    ct, preds = COLUMN_FOLD_CASES[case]
    rng = np.random.default_rng(700 + case)
    n = 300_000
    keys = rng.integers(0, 5000, n).astype(np.int64)
    vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
    filt = fold_values(ct, n, rng)
    fnull = rng.random(n) < 0.05
    w = [Where(op, "filter", c, t) for op, c, t in preds]
    aggs = [qk.AGG_SUM, qk.AGG_COUNT, qk.AGG_MIN, qk.AGG_MAX]
    extra = {"filter": (filt, fnull)}
    gpu = gpu_group_by(aggs, [keys], vals, None, None, where=w, capacity=10_000, path=path, block=100_000,
                       extra_cols=extra)
    cpu = cpu_group_by(aggs, [keys], vals, None, None, where=w, extra_cols=extra)
    compare(gpu, cpu, len(aggs))


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("case", range(len(FOLD_CASES)))
def test_value_range_fold_matches_predicates(case, path):
    vt, preds = FOLD_CASES[case]
    rng = np.random.default_rng(500 + case)
    n = 300_000
    keys = rng.integers(0, 5000, n).astype(np.int64)
    vals = fold_values(vt, n, rng)
    vn = rng.random(n) < 0.05
    w = [Where(op, "value", c, ct) for op, c, ct in preds]
    aggs = [qk.AGG_COUNT, qk.AGG_COUNT_ALL, qk.AGG_MIN, qk.AGG_MAX]
    gpu = gpu_group_by(aggs, [keys], vals, None, vn, where=w, capacity=10_000, path=path)
    cpu = cpu_group_by(aggs, [keys], vals, None, vn, where=w)
    compare(gpu, cpu, len(aggs))


# ---- COUNT of a nullable value with the WHERE on another column --------------------------------

@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("form", ["where", "mask"])
def test_count_nullable_value_where_on_other_column(path, form):
    """COUNT(value) over a nullable value column with the WHERE on a different column: the
    NULL bits that count are those of the rows that pass the filter (correct SQL; the
    reference's dispatcher skips compacting the value's null mask for COUNT,
    GpuSqlDispatcherAggregationFunctions.h:392-400 vs :422-424, quirk Q3).  COUNT(*) counts
    every passing row."""
    rng = np.random.default_rng(300 + path)
    n = 1_200_000
    keys = datagen.zipf_keys(21, 1.1, 20_000, n)
    vals = rng.integers(-10**6, 10**6, n).astype(np.int64)
    vn = rng.random(n) < 0.2
    filt = rng.integers(-100, 100, n).astype(np.int32)
    fnull = rng.random(n) < 0.05
    aggs = [qk.AGG_COUNT, qk.AGG_COUNT_ALL, qk.AGG_SUM, qk.AGG_MIN]
    extra = {"filter": (filt, fnull)}
    if form == "where":
        w = [Where(qk.GREATER, "filter", 10, qk.INT)]
        gpu = gpu_group_by(aggs, [keys], vals, None, vn, where=w, extra_cols=extra, capacity=40_000,
                           path=path, block=1 << 19)
        cpu = cpu_group_by(aggs, [keys], vals, None, vn, where=w, extra_cols=extra)
    else:
        m = ((filt > 10) & ~fnull).astype(np.int8)
        gpu = gpu_group_by(aggs, [keys], vals, None, vn, mask=m, capacity=40_000, path=path, block=1 << 19)
        cpu = cpu_group_by(aggs, [keys], vals, None, vn, mask=m)
    compare(gpu, cpu, len(aggs))
    # the property that distinguishes Q3: COUNT(value) < COUNT(*) in groups with NULL values
    cnt, cntall = as_map(gpu, 0), as_map(gpu, 1)
    assert sum(cnt.values()) < sum(cntall.values())
    passing = (filt > 10) & ~fnull
    assert sum(cntall.values()) == int(passing.sum())
    assert sum(cnt.values()) == int((passing & ~vn).sum())


# ---- the packed (int32, int32) key equal to the empty marker ------------------------------------

@pytest.mark.parametrize("path", [None, 1])
def test_pair_key_equal_to_empty_marker(path):
    """Two non-NULL int32 key columns pack into one word, a in the low half and b in the high
    half; (0, INT32_MIN) packs to exactly the empty-slot marker 0x8000000000000000.  That key
    takes its side slot in phase 1 and in the table and comes back out as (0, INT32_MIN)."""
    rng = np.random.default_rng(61)
    n = 1 << 21
    a = rng.integers(0, 2000, n).astype(np.int32)
    b = rng.integers(-2000, 2000, n).astype(np.int32)
    edge = rng.choice(n, 3000, replace=False)
    a[edge] = 0
    b[edge] = np.iinfo(np.int32).min
    near = rng.choice(n, 500, replace=False)
    a[near] = 0
    b[near] = np.iinfo(np.int32).min + 1
    vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]
    gpu = gpu_group_by(aggs, [a, b], vals, capacity=5_000_000, flags=qk.GB_KEYS_NOT_NULL, path=path,
                       block=1 << 20)
    cpu = cpu_group_by(aggs, [a, b], vals)
    compare(gpu, cpu, len(aggs))
    m = as_map(gpu, 3)
    assert m[(0, int(np.iinfo(np.int32).min))] == int(np.count_nonzero((a == 0) & (b == np.iinfo(np.int32).min)))


def _same_floats(x, y):
    """Element-wise: both NaN with the same sign, or equal with the same sign (so -0.0 differs
    from +0.0)."""
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    nx, ny = np.isnan(x), np.isnan(y)
    return (nx == ny) & (np.signbit(x) == np.signbit(y)) & (nx | (x == y))


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_nan_and_signed_zero_keys_and_values(dt, path):
    """NaN and signed zeros as keys and values (the oracle's rule N1): -0.0 groups with +0.0,
    all NaN keys form one group keyed by the positive quiet NaN, MIN/MAX follow IEEE totalOrder
    (sign of NaNs and zeros included), SUM/AVG are NaN exactly where the oracle's are. Hot keys
    from a small pool (aggregated in LDS on the two-phase path) mixed with 100K cold ones."""
    rng = np.random.default_rng(31 if dt == np.float32 else 32)
    n = 1 << 20
    qnan_payload = np.array([0x7FC12345], np.uint32).view(np.float32)[0]
    key_pool = np.array([0.0, -0.0, 1.5, np.nan, -np.nan, qnan_payload, np.inf, -np.inf], dt)
    val_pool = np.array([1.0, -1.0, 0.0, -0.0, np.nan, -np.nan, np.inf, -np.inf, 2.5], dt)
    hot = rng.random(n) < 0.6
    keys = np.where(hot, key_pool[rng.integers(0, key_pool.size, n)],
                    (rng.integers(1, 100_000, n) * 0.25).astype(dt)).astype(dt)
    vals = np.where(rng.random(n) < 0.3, val_pool[rng.integers(0, val_pool.size, n)],
                    rng.standard_normal(n).astype(dt)).astype(dt)
    vn = rng.random(n) < 0.05
    aggs = [qk.AGG_MIN, qk.AGG_MAX, qk.AGG_SUM, qk.AGG_AVG, qk.AGG_COUNT]
    gpu = gpu_group_by(aggs, [keys], vals, None, vn, block=1 << 19, capacity=200_000, path=path)
    cpu = cpu_group_by(aggs, [keys], vals, None, vn)
    assert gpu["count"] == cpu["count"]
    a, b = O.sort_groups(gpu), O.sort_groups(cpu)
    assert np.array_equal(a["keys"][0], b["keys"][0])
    nan_key = np.flatnonzero(np.isnan(np.asarray(gpu["keys"][0][:gpu["count"]], np.float64)))
    assert nan_key.size == 1, "one NaN group"
    assert not np.signbit(np.float64(gpu["keys"][0][nan_key[0]]))
    zero = np.asarray(gpu["keys"][0][:gpu["count"]], np.float64) == 0
    assert zero.sum() == 1 and not np.signbit(np.asarray(gpu["keys"][0][:gpu["count"]], np.float64)[zero]).any()
    for i in range(len(aggs)):
        assert np.array_equal(a["value_nulls"][i], b["value_nulls"][i])
    live = ~b["value_nulls"][0]
    for i in (0, 1):
        ok = _same_floats(a["values"][i][live], b["values"][i][live])
        assert ok.all(), f"aggregate {i}: {np.count_nonzero(~ok)} groups differ"
    for i in (2, 3):
        x, y = (np.asarray(v[live], np.float64) for v in (a["values"][i], b["values"][i]))
        assert np.array_equal(np.isnan(x), np.isnan(y))
        fin = ~np.isnan(y) & np.isfinite(y)
        assert np.allclose(x[fin], y[fin], rtol=FP_REL, atol=1e-3)
        assert np.array_equal(x[np.isinf(y)], y[np.isinf(y)])
    assert np.array_equal(a["values"][4], b["values"][4])


# ---- the streaming phase 1 at the edges of its blocks ---------------------------------------

STREAM_BLOCKS = [1, 63, 255, 256, 257, 4095, 4097, 65_537, 300_001]


def run_blocks(aggs, keys, vals, kn, vn, where, extra, flags):
    """One table fed blocks of the sizes in STREAM_BLOCKS (two-phase path forced), its groups
    and stats; the same rows through the oracle in one go."""
    n = keys[0].size
    ktypes = [qk._abi.TYPE_OF[k.dtype] for k in keys]
    gb = qk.GPUGroupBy(aggs, ktypes, qk._abi.TYPE_OF[vals.dtype], 200_000, 1, flags=flags)
    gb.set_path(1, 0)
    s = 0
    for b in STREAM_BLOCKS:
        e = min(n, s + b)
        kcols = [col(k[s:e], None if kn is None or kn[i] is None else kn[i][s:e]) for i, k in enumerate(keys)]
        vcol = col(vals[s:e], None if vn is None else vn[s:e])
        cols = {"value": vcol}
        for name, (arr, isnull) in extra.items():
            cols[name] = col(arr[s:e], None if isnull is None else isnull[s:e])
        knm = [c.nulls for c in kcols] if kn is not None else None
        preds = [w.gpu(cols, {}) for w in where]
        gb.ProcessBlockWhere([c.data for c in kcols], vcol.data, e - s, preds, knm, vcol.nulls)
        s = e
    assert s == n
    got = gb.results_host()
    st = gb.stats()
    gb.close()
    want = cpu_group_by(aggs, keys, vals, kn, vn, where=where, extra_cols=extra)
    return got, want, st


@pytest.mark.parametrize("flags", [0, qk.GB_ASYNC])
@pytest.mark.parametrize("form", ["long_value_range", "int_no_where", "pair_column_range"])
def test_streaming_phase1_ragged_blocks(form, flags):
    """Blocks of 1 row up to 300,001 rows, none a multiple of a chunk except 256, through the
    streaming phase 1 (rows past a block's end read as zeros and must drop out): nullable int64
    keys with the empty-marker key and a WHERE folded on the value; nullable int32 keys without
    a WHERE; two int32 keys declared non-NULL (packed into one word) with a WHERE folded on
    another, nullable column."""
    rng = np.random.default_rng({"long_value_range": 1, "int_no_where": 2, "pair_column_range": 3}[form])
    n = sum(STREAM_BLOCKS)
    vn = rng.random(n) < 0.05
    extra, where, kn = {}, [], None
    if form == "long_value_range":
        keys = [np.where(rng.random(n) < 0.6, datagen.zipf_keys(3, 1.1, 3000, n),
                         rng.integers(-(2**40), 2**40, n)).astype(np.int64)]
        keys[0][rng.random(n) < 0.01] = np.iinfo(np.int64).min
        kn = [rng.random(n) < 0.01]
        vals = rng.lognormal(0, 1, n).astype(np.float32)
        where = [Where(qk.GREATER_EQUAL, "value", 0.5, qk.FLOAT), Where(qk.LESS_EQUAL, "value", 4.0, qk.FLOAT)]
        aggs, real = [qk.AGG_SUM, qk.AGG_AVG, qk.AGG_COUNT, qk.AGG_COUNT_ALL], (0, 1)
    elif form == "int_no_where":
        keys = [datagen.zipf_keys(4, 1.1, 50_000, n).astype(np.int32)]
        kn = [rng.random(n) < 0.02]
        vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
        aggs, real = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT, qk.AGG_COUNT_ALL], ()
    else:
        keys = [rng.integers(0, 300, n).astype(np.int32), rng.integers(-(2**31), 2**31, n).astype(np.int32)]
        keys[1][rng.random(n) < 0.5] = 7
        vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
        ts = rng.integers(0, 10**6, n).astype(np.int64)
        extra = {"ts": (ts, rng.random(n) < 0.03)}
        where = [Where(qk.GREATER_EQUAL, "ts", 100_000, qk.LONG), Where(qk.LESS, "ts", 700_000, qk.LONG)]
        aggs, real = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT], ()
        flags |= qk.GB_KEYS_NOT_NULL  # the two int32 fields then pack into one key word
    got, want, st = run_blocks(aggs, keys, vals, kn, vn, where, extra, flags)
    assert st["path"] == 1 and st["p1_form"] == 1
    compare(got, want, len(aggs), real)
