"""qk_prefix_sum_noalloc (the prefix sum whose look-back state lives in the library's fixed
device array) against qk_prefix_sum, through the C ABI: two calls in a row, the second on a
smaller input, both scan forms, and the host-side refusal of inputs beyond the array."""

import numpy as np
import pytest

import oracle as O
import qikkdb_amd as qk
from qikkdb_amd import DeviceBuffer

pytestmark = pytest.mark.gpu


def scan(fn, d_mask, n, inclusive):
    out = DeviceBuffer(max(n, 1) * 4)
    qk._abi.call(fn, d_mask.ptr, n, out.ptr, 1 if inclusive else 0, None)
    qk._abi.call("qk_device_synchronize")
    return out.to_numpy(np.int32, n)


@pytest.mark.parametrize("inclusive", [True, False])
def test_noalloc_matches_prefix_sum_twice_second_smaller(inclusive):
    qk.ensure_device(0)
    rng = np.random.default_rng(5 + inclusive)
    for n in ((1 << 22) + 4097, 300_001):
        mask = rng.integers(-2, 3, n).astype(np.int8)
        d = DeviceBuffer.from_numpy(mask)
        print(f"case: prefix_sum_noalloc n={n} inclusive={inclusive}", flush=True)
        got = scan("qk_prefix_sum_noalloc", d, n, inclusive)
        want = scan("qk_prefix_sum", d, n, inclusive)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(got, O.prefix_sum(mask, inclusive))


def test_noalloc_refuses_more_rows_than_its_array_and_takes_empty_input():
    qk.ensure_device(0)
    d = DeviceBuffer(64)
    lib = qk._abi.load()
    rc = lib.qk_prefix_sum_noalloc(d.ptr, (1 << 30) + 1, d.ptr,
                                   1, None)
    assert rc == qk._abi.QK_ERROR
    assert b"exceed" in lib.qk_last_error_message()
    assert lib.qk_prefix_sum_noalloc(d.ptr, 0, d.ptr, 1, None) == 0
