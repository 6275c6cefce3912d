"""GPU GROUP BY on real data: the 10,000 taxi trips of the reference's sample database
(tests/golden/trips_sample.npz, read from its binary column files by qikkdb_amd.storage),
against the oracle: numeric keys, a composite key, an int64 key with thousands of groups, a
WHERE on a float column, and string keys (the pickup date)."""
import os

import numpy as np
import pytest

import oracle as O
import qikkdb_amd as qk
from qikkdb_amd.strings import DeviceStrings, GPUGroupByString
from gpu_util import Where, cpu_group_by, gpu_group_by
from helpers import compare_all_groups

pytestmark = pytest.mark.gpu
FX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trips_sample.npz"))
FP_REL = 1e-5
AGGS = [qk.AGG_SUM, qk.AGG_AVG, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]


@pytest.mark.parametrize("key,value", [("passenger_count", "total_amount"), ("cab_type", "trip_distance"),
                                       ("pickup_datetime", "total_amount")])
def test_single_key_on_sample_trips(key, value):
    k, v = FX[key], FX[value]
    print(f"case: trips GROUP BY {key}, aggregates of {value}", flush=True)
    got = gpu_group_by(AGGS, [k], v, block=4096, capacity=16384)
    want = cpu_group_by(AGGS, [k], v)
    assert got["count"] == want["count"]
    compare_all_groups(got, want, rel_aggs=(0, 1), rel=FP_REL)


def test_where_and_composite_key_on_sample_trips():
    k1, k2, v = FX["cab_type"], FX["passenger_count"], FX["total_amount"]
    w = [Where(qk.GREATER, "v", 2.0, qk.FLOAT)]
    got = gpu_group_by(AGGS, [k1, k2], v, where=w, block=3000, capacity=1024, extra_cols={"v": (v, None)})
    want = cpu_group_by(AGGS, [k1, k2], v, where=w, extra_cols={"v": (v, None)})
    assert got["count"] == want["count"]
    compare_all_groups(got, want, rel_aggs=(0, 1), rel=FP_REL)


def test_string_key_on_sample_trips():
    dates = [bytes(d) for d in FX["pickup_date"]]
    v = FX["total_amount"]
    gb = GPUGroupByString([qk.AGG_SUM, qk.AGG_COUNT, qk.AGG_MAX], qk.FLOAT, 1024)
    try:
        for s in range(0, len(dates), 2500):
            e = min(len(dates), s + 2500)
            gb.ProcessBlock(DeviceStrings.from_list(dates[s:e]), qk.DeviceBuffer.from_numpy(v[s:e]), e - s)
        got = gb.results_host()
    finally:
        gb.close()
    o = O.GroupByStrings([O.AGG_SUM, O.AGG_COUNT, O.AGG_MAX], O.FLOAT)
    o.process(dates, v)
    want = o.results()
    assert got["count"] == want["count"] > 1
    gm = {k: (float(a), int(b), float(c)) for k, a, b, c in zip(got["keys"], *got["values"])}
    wm = {k: (float(a), int(b), float(c)) for k, a, b, c in zip(want["keys"], *want["values"])}
    assert gm.keys() == wm.keys()
    for k in wm:
        assert gm[k][1:] == wm[k][1:]
        assert abs(gm[k][0] - wm[k][0]) <= FP_REL * abs(wm[k][0]) + 1e-9
