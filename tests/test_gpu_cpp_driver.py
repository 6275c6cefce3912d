"""The C++ drop-in headers (include/qikkdb/) driven the way the reference's dispatcher drives
its operators: filter -> reconstructColKeep(key), reconstructColKeep(value) -> ProcessBlock
per block -> GetResults, one table per aggregate.  The output must equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from qikkdb_amd import build as B
from qikkdb_amd import datagen

pytestmark = pytest.mark.gpu


def run_driver(tmp_path, keys, vals, block, threshold):
    assert os.path.exists(B.DRIVER), "tests/cpp/groupby_driver is built by qikkdb_amd.build"
    kp, vp = tmp_path / "k.i32", tmp_path / "v.i32"
    keys.astype(np.int32).tofile(kp)
    vals.astype(np.int32).tofile(vp)
    r = subprocess.run([B.DRIVER, str(keys.size), str(block), str(kp), str(vp), str(threshold)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, s, mn, mx, c = (int(x) for x in line.split())
        out[k] = (s, mn, mx, c)
    return out


def oracle(keys, vals, threshold):
    res = {}
    for agg in (O.AGG_SUM, O.AGG_MIN, O.AGG_MAX, O.AGG_COUNT):
        out_types = [O.LONG] if agg == O.AGG_SUM else None
        g = O.GroupBy([agg], [O.INT], O.INT, out_types)
        g.process([keys], vals, preds=[O.pred(0, vals, threshold, O.INT)])
        r = g.results()
        for k, v in zip(r["keys"][0], r["values"][0]):
            res.setdefault(int(k), []).append(int(v))
    return {k: tuple(v) for k, v in res.items()}


def test_driver_small_workload(tmp_path):
    d = datagen.small(1 << 20, seed=1)
    got = run_driver(tmp_path, d["key"], d["value"], 1 << 18, 0)
    assert got == oracle(d["key"], d["value"], 0)


def test_driver_ragged_blocks_and_empty_block(tmp_path):
    rng = np.random.default_rng(3)
    n = 100_003
    keys = rng.integers(-50, 50, n).astype(np.int32)
    vals = rng.integers(-1000, 1000, n).astype(np.int32)
    vals[:30_000] = -5  # the first block keeps no row
    got = run_driver(tmp_path, keys, vals, 30_000, 10)
    assert got == oracle(keys, vals, 10)
