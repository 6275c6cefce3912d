"""GPU parity of the filter kernels against the oracle (bit-exact masks)."""
import numpy as np
import pytest

import oracle as O
import qikkdb_amd as qk
from qikkdb_amd import DeviceBuffer
from gpu_util import col
from helpers import pack_bits, unpack_bits

pytestmark = pytest.mark.gpu

TYPES = [np.int32, np.int64, np.float32, np.float64]
PAIRS = [(t, t) for t in TYPES] + [(np.int32, np.float32), (np.int64, np.float32),
                                   (np.int32, np.int64), (np.int32, np.float64),
                                   (np.float32, np.float64), (np.int8, np.int8)]


def rand(dt, n, rng):
    if np.issubdtype(dt, np.integer):
        lo, hi = (-1000, 1000) if dt in (np.int8,) else (-50, 50)
        return rng.integers(max(lo, np.iinfo(dt).min), min(hi, np.iinfo(dt).max), n).astype(dt)
    return (rng.integers(-50, 50, n) / 2).astype(dt)


@pytest.mark.parametrize("lt,rt", PAIRS, ids=lambda t: np.dtype(t).name)
@pytest.mark.parametrize("op", range(6))
def test_filter_col_col_and_const(lt, rt, op):
    rng = np.random.default_rng(op)
    n = 100_003
    a, b = rand(lt, n, rng), rand(rt, n, rng)
    isnull = rng.random(n) < 0.1
    ca, cb = col(a), col(b)
    nm = DeviceBuffer.from_numpy(pack_bits(isnull))
    got = qk.GPUBinary.Filter(op, ca, cb, n, nm).to_numpy(np.int8, n)
    np.testing.assert_array_equal(got, O.filter_mask(op, a, b, n, pack_bits(isnull)))
    c = rt(b[17])
    tb = O.TYPE_OF[np.dtype(rt)]
    got = qk.GPUBinary.Filter(op, ca, qk.Const(c, tb), n).to_numpy(np.int8, n)
    np.testing.assert_array_equal(got, O.filter_mask(op, a, (c, tb), n))
    ta = O.TYPE_OF[np.dtype(lt)]
    c2 = lt(a[5])
    got = qk.GPUBinary.Filter(op, qk.Const(c2, ta), cb, n).to_numpy(np.int8, n)
    np.testing.assert_array_equal(got, O.filter_mask(op, (c2, ta), b, n))
    bits = qk.GPUBinary.Filter(op, ca, cb, n, nm, bits=True).to_numpy(np.uint64)
    np.testing.assert_array_equal(unpack_bits(bits, n), O.filter_mask(op, a, b, n, pack_bits(isnull)).astype(bool))


def test_filter_conversions_and_const_const():
    a = np.array([16777217, 16777216, 3, 7, 9], np.int32)
    got = qk.GPUBinary.Filter(qk.EQUAL, col(a), qk.Const(16777216.0, qk.FLOAT), 5).to_numpy(np.int8, 5)
    assert got.tolist() == [1, 1, 0, 0, 0]
    got = qk.GPUBinary.Filter(qk.EQUAL, col(a), qk.Const(16777216.0, qk.DOUBLE), 5).to_numpy(np.int8, 5)
    assert got.tolist() == [0, 1, 0, 0, 0]
    got = qk.GPUBinary.Filter(qk.LESS, qk.Const(1, qk.INT), qk.Const(2, qk.LONG), 70).to_numpy(np.int8, 70)
    assert got.tolist() == [1] * 70
    bits = qk.GPUBinary.Filter(qk.LESS, qk.Const(1, qk.INT), qk.Const(2, qk.LONG), 70, bits=True)
    assert unpack_bits(bits.to_numpy(np.uint64), 70).all()
    assert bits.to_numpy(np.uint64)[1] == (1 << 6) - 1


def test_logic_ops_and_not():
    rng = np.random.default_rng(1)
    n = 70_001
    m1 = rng.integers(0, 2, n).astype(np.int8)
    m2 = rng.integers(0, 2, n).astype(np.int8)
    for op in (qk.LOGICAL_AND, qk.LOGICAL_OR):
        got = qk.GPUBinary.Filter(op, col(m1), col(m2), n).to_numpy(np.int8, n)
        np.testing.assert_array_equal(got, O.filter_mask(op, m1, m2, n))
    isnull = rng.random(n) < 0.2
    got = qk.GPUBinary.LogicalNot(DeviceBuffer.from_numpy(m1), n,
                                  DeviceBuffer.from_numpy(pack_bits(isnull))).to_numpy(np.int8, n)
    np.testing.assert_array_equal(got, O.logical_not(m1, pack_bits(isnull)))


def test_filter_rejects_empty_block():
    with pytest.raises(qk.QueryEngineError) as e:
        qk.GPUBinary.Filter(qk.LESS, col(np.zeros(4, np.int32)), qk.Const(1, qk.INT), 0)
    assert e.value.code == qk.QK_ERROR
