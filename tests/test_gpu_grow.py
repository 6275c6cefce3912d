"""GPU parity of growing GROUP BY tables (QK_GB_AUTO_GROW), in the pattern of the reference's
UnlimitedNumberOfKeys tests (DispatcherGroupByTests.cpp:1652-1802): 1.5x to 2.5x more distinct
keys than the table's initial capacity, blocks of half the capacity (growth between blocks) or
one block holding every key (sub-tables fill mid-block, the rows whose keys found no slot are
aggregated again after the growth).  Every result equals the oracle's and a large table's;
both row-kernel paths, single and composite keys, WHERE, masks, NULL keys and merges."""
import numpy as np
import pytest

import qikkdb_amd as qk
from gpu_util import Where, col, cpu_group_by, gpu_group_by
from helpers import as_map, maps_close

pytestmark = pytest.mark.gpu
B = 4096  # initial capacity (the reference's GroupByBuckets)
AGGS = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT, qk.AGG_AVG, qk.AGG_COUNT_ALL]


def compare(got, want):
    assert got["count"] == want["count"]
    for a in range(len(AGGS)):
        maps_close(as_map(got, a), as_map(want, a))


@pytest.mark.parametrize("path", [0, 1], ids=["one_pass", "two_phase"])
@pytest.mark.parametrize("shape", ["keys_1.5x_blocks", "keys_2.5x_blocks", "keys_2.5x_one_block"])
def test_unlimited_keys_single(path, shape):
    rng = np.random.default_rng(1652)
    if shape == "keys_1.5x_blocks":
        unique, n, block = B * 3 // 2, B * 3 // 2 * 9 // 8, B // 2
    elif shape == "keys_2.5x_blocks":
        unique, n, block = B * 5 // 2, B * 5 // 2, B // 2
    else:
        unique, n, block = B * 5 // 2, B * 5 // 2 * 3, None
    keys = (np.arange(n) % unique).astype(np.int32)
    rng.shuffle(keys)
    vals = rng.integers(-1000, 1000, n).astype(np.int32)
    print(f"case: grow {shape} path {path}: {unique} keys, {n} rows, block {block}", flush=True)
    st = {}
    got = gpu_group_by(AGGS, [keys], vals, block=block, capacity=B, flags=qk.GB_AUTO_GROW, path=path,
                       stats=st)
    assert st["grows"] >= 1
    compare(got, cpu_group_by(AGGS, [keys], vals))
    compare(got, gpu_group_by(AGGS, [keys], vals, block=block, capacity=4 * unique, path=path))


@pytest.mark.parametrize("path", [0, 1], ids=["one_pass", "two_phase"])
def test_grow_with_where_nulls_and_int64_keys(path):
    # NULL keys, NULL values and a WHERE on another column; one block far beyond the capacity
    rng = np.random.default_rng(1700)
    n = 6 * B
    keys = (rng.integers(0, 3 * B, n) * 1_000_003 - 7).astype(np.int64)
    knull = rng.random(n) < 0.03
    vals = rng.standard_normal(n).astype(np.float32)
    vnull = rng.random(n) < 0.05
    flt = rng.integers(0, 10, n).astype(np.int32)
    where = [Where(qk.GREATER, "filter", 2, qk.INT)]
    extra = {"filter": (flt, None)}
    aggs = [qk.AGG_SUM, qk.AGG_COUNT, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT_ALL]
    st = {}
    got = gpu_group_by(aggs, [keys], vals, [knull], vnull, where=where, capacity=B // 4,
                       flags=qk.GB_AUTO_GROW, path=path, extra_cols=extra, stats=st)
    assert st["grows"] >= 2
    want = cpu_group_by(aggs, [keys], vals, [knull], vnull, where=where, extra_cols=extra)
    assert got["count"] == want["count"]
    for a in range(len(aggs)):
        maps_close(as_map(got, a), as_map(want, a), rel=1e-5 if a == 0 else 0.0)


@pytest.mark.parametrize("flags", [0, qk.GB_KEYS_NOT_NULL], ids=["nullable", "not_null"])
def test_grow_composite_keys_masked(flags):
    rng = np.random.default_rng(1760)
    n = 5 * B
    a = rng.integers(0, 96, n).astype(np.int32)
    b = rng.integers(0, 96, n).astype(np.int32)
    kn = None if flags else [rng.random(n) < 0.02, rng.random(n) < 0.02]
    vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
    mask = (rng.random(n) < 0.7).astype(np.int8)
    st = {}
    got = gpu_group_by(AGGS, [a, b], vals, kn, mask=mask, capacity=B // 8,
                       flags=flags | qk.GB_AUTO_GROW, block=B, stats=st)
    assert st["grows"] >= 1
    compare(got, cpu_group_by(AGGS, [a, b], vals, kn, mask=mask))


def test_full_without_growth_still_reported():
    keys = np.arange(3 * B, dtype=np.int32)
    with pytest.raises(qk.QueryEngineError) as e:
        gpu_group_by(qk.AGG_COUNT, [keys], None, capacity=B)
    assert e.value.code == qk.QK_HASH_TABLE_FULL


def test_merge_into_growing_table():
    rng = np.random.default_rng(1802)
    k1 = rng.integers(0, 3 * B, 4 * B).astype(np.int32)
    k2 = rng.integers(2 * B, 5 * B, 4 * B).astype(np.int32)
    v1 = rng.integers(-50, 50, k1.size).astype(np.int32)
    v2 = rng.integers(-50, 50, k2.size).astype(np.int32)
    dst = qk.GPUGroupBy(AGGS, qk.INT, qk.INT, B, 1, flags=qk.GB_AUTO_GROW)
    src = qk.GPUGroupBy(AGGS, qk.INT, qk.INT, 8 * B, 1)
    try:
        c1, c2 = col(k1), col(k2)
        d1, d2 = col(v1), col(v2)
        dst.ProcessBlock([c1.data], d1.data, k1.size)
        src.ProcessBlock([c2.data], d2.data, k2.size)
        dst.merge(src)
        got = dst.results_host()
        assert dst.stats()["grows"] >= 1
    finally:
        dst.close()
        src.close()
    compare(got, cpu_group_by(AGGS, [np.concatenate([k1, k2])], np.concatenate([v1, v2])))
