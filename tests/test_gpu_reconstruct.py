"""GPU parity of prefix sums, compaction, index generation and null-mask handling."""
import numpy as np
import pytest

import oracle as O
import qikkdb_amd as qk
from qikkdb_amd import DeviceBuffer, GPUReconstruct
from gpu_util import Where, col
from helpers import golden, pack_bits, unpack_bits

pytestmark = pytest.mark.gpu
G = golden()
SIZES = [1, 17, 4095, 4096, 4097, 65536 * 3 + 5, (1 << 20) + 3]


@pytest.mark.parametrize("n", SIZES + [1 << 24])
@pytest.mark.parametrize("inclusive", [True, False])
def test_prefix_sum(n, inclusive):
    rng = np.random.default_rng(n)
    mask = rng.integers(0, 2, n).astype(np.int8)
    d = DeviceBuffer.from_numpy(mask)
    got = GPUReconstruct.PrefixSum(d, n, inclusive).to_numpy(np.int32, n)
    np.testing.assert_array_equal(got, O.prefix_sum(mask, inclusive))
    assert GPUReconstruct.Sum(d, n) == O.mask_sum(mask)


def test_prefix_sum_of_values_and_repeat():
    # the scan sums the int8 values (CUB InclusiveSum semantics); repeated launches reuse
    # the look-back status words through epochs
    rng = np.random.default_rng(2)
    n = 300_001
    mask = rng.integers(-3, 4, n).astype(np.int8)
    d = DeviceBuffer.from_numpy(mask)
    want = O.prefix_sum(mask, True)
    for _ in range(5):
        np.testing.assert_array_equal(GPUReconstruct.PrefixSum(d, n).to_numpy(np.int32, n), want)


@pytest.mark.parametrize("dt", [np.int8, np.int16, np.int32, np.float32, np.int64, np.float64])
@pytest.mark.parametrize("sel", [0.0, 0.001, 0.5, 1.0])
def test_reconstruct_col_keep(dt, sel):
    rng = np.random.default_rng(int(sel * 1000) + np.dtype(dt).itemsize)
    n = 250_007
    vals = rng.integers(-100, 100, n).astype(dt)
    mask = (rng.random(n) < sel).astype(np.int8)
    isnull = rng.random(n) < 0.1
    out, cnt, onull = GPUReconstruct.reconstructColKeep(
        DeviceBuffer.from_numpy(vals), vals.itemsize, DeviceBuffer.from_numpy(mask), n,
        DeviceBuffer.from_numpy(pack_bits(isnull)))
    want, want_null, k = O.compact(vals, mask, pack_bits(isnull))
    assert cnt == k
    if k == 0:
        assert out is None and onull is None
        return
    np.testing.assert_array_equal(out.to_numpy(dt, cnt), want)
    np.testing.assert_array_equal(unpack_bits(onull.to_numpy(np.uint64), cnt), unpack_bits(want_null, k))


def test_reconstruct_aliases_without_mask_and_empty_block():
    v = DeviceBuffer.from_numpy(np.arange(10, dtype=np.int32))
    out, cnt, _ = GPUReconstruct.reconstructColKeep(v, 4, None, 10)
    assert out is v and cnt == 10
    out, cnt, on = GPUReconstruct.reconstructColKeep(v, 4, DeviceBuffer.from_numpy(np.ones(1, np.int8)), 0)
    assert out is None and cnt == 0


@pytest.mark.parametrize("case", G["generate_indexes"], ids=lambda c: c["source"].split()[-1])
def test_generate_indexes_known_answers(case):
    mask = np.array(case["mask"], np.int8)
    got = GPUReconstruct.GenerateIndexes(DeviceBuffer.from_numpy(mask), mask.size)
    assert got.tolist() == case["expect"]


@pytest.mark.parametrize("index_size", [4, 8])
def test_generate_indexes_alternating_and_random(index_size):
    n = G["generate_indexes_alternating"]["n"]
    mask = (np.arange(n) % 2).astype(np.int8)
    got = GPUReconstruct.GenerateIndexes(DeviceBuffer.from_numpy(mask), n, index_size)
    np.testing.assert_array_equal(got, np.arange(1, n, 2))
    rng = np.random.default_rng(4)
    m = (rng.random(1_000_003) < 0.3).astype(np.int8)
    got = GPUReconstruct.GenerateIndexes(DeviceBuffer.from_numpy(m), m.size, index_size)
    np.testing.assert_array_equal(got, O.generate_indexes(m))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 100_000])
def test_compress_null_mask(n):
    arr = np.random.default_rng(n).integers(0, 2, n).astype(np.uint8)
    got = GPUReconstruct.CompressNullMask(DeviceBuffer.from_numpy(arr), n)
    np.testing.assert_array_equal(got.to_numpy(np.uint64), O.compress_null_mask(arr))


def test_where_compact_matches_filter_then_reconstruct():
    rng = np.random.default_rng(6)
    n = 1_000_005
    key = rng.integers(0, 100, n).astype(np.int64)
    val = rng.lognormal(0, 1, n).astype(np.float32)
    vnull = rng.random(n) < 0.05
    cols = {"key": col(key), "value": col(val, vnull)}
    where = [Where(qk.GREATER_EQUAL, "value", 0.5, qk.FLOAT), Where(qk.LESS_EQUAL, "value", 4.0, qk.FLOAT)]
    preds = [w.gpu(cols, {}) for w in where]
    outs, cnt = GPUReconstruct.WhereCompact(preds, [cols["key"], cols["value"]], n)
    hp = [w.cpu({"value": val}, {"value": pack_bits(vnull)}) for w in where]
    sel = np.array([O.lib().or_eval_where(len(hp), (O.OrPred * 2)(*hp), i) for i in range(0, n, 1001)], bool)
    mask = ((val >= np.float32(0.5)) & (val <= np.float32(4.0)) & ~vnull).astype(np.int8)
    np.testing.assert_array_equal(sel, mask[::1001].astype(bool))  # the closed form agrees with the oracle
    assert cnt == int(mask.sum())
    np.testing.assert_array_equal(outs[0].to_numpy(), key[mask != 0])
    np.testing.assert_array_equal(outs[1].to_numpy(), val[mask != 0])
