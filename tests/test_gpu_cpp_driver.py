"""The C++ drop-in headers (include/qikkdb/) driven the way the reference's dispatcher drives
its operators: filter -> reconstructColKeep(key), reconstructColKeep(value) -> ProcessBlock
per block -> GetResults, one table per aggregate.  The output must equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import as_map
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


# ---- the dispatcher-shaped caller: GroupByHelper over groupByTables_ ---------------------------

AGG_CODES = {"min": O.AGG_MIN, "max": O.AGG_MAX, "sum": O.AGG_SUM, "avg": O.AGG_AVG, "count": O.AGG_COUNT}


def _pack(isnull):
    from helpers import pack_bits
    return pack_bits(isnull)


def test_dispatcher_shaped_caller(tmp_path):
    """GroupByHelper-shaped single- and composite-key queries over two dispatcher tables merged by
    GetResults(..., groupByTables_, ...), each aggregate against the oracle; the driver's own
    checks of reconstructCol, GenerateIndexes, PrefixSum, Sum/SumKeep, CompressNullMask, the
    pre-seeded constructor and ReconstructRawNumbers must all pass."""
    assert os.path.exists(B.DISPATCHER), "tests/cpp/dispatcher_driver is built by qikkdb_amd.build"
    rng = np.random.default_rng(12)
    n = 60_001
    a = rng.integers(-40, 40, n).astype(np.int32)
    b = rng.integers(0, 6, n).astype(np.int32)
    v = rng.integers(-1000, 1000, n).astype(np.int32)
    f = rng.integers(-100, 100, n).astype(np.int32)
    an, bn, vn = rng.random(n) < 0.03, rng.random(n) < 0.03, rng.random(n) < 0.1
    for name, arr in (("a", a), ("b", b), ("v", v), ("f", f)):
        arr.tofile(tmp_path / f"{name}.i32")
    for name, arr in (("an", an), ("bn", bn), ("vn", vn)):
        arr.astype(np.uint8).tofile(tmp_path / f"{name}.u8")
    threshold, block = -30, 7_000
    r = subprocess.run([B.DISPATCHER, str(tmp_path), str(n), str(block), str(threshold)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = {}
    checks = set()
    for line in r.stdout.splitlines():
        p = line.split()
        if p[0] == "CHECK":
            checks.add(p[1])
            continue
        agg = p[1]
        if p[0] == "S":
            key = (None if int(p[2]) else int(p[3]),)
            val = None if int(p[4]) else int(p[5])
        else:
            key = (None if int(p[2]) else int(p[3]), None if int(p[4]) else int(p[5]))
            val = None if int(p[6]) else int(p[7])
        m = got.setdefault((p[0], agg), {})
        assert key not in m, f"duplicate group {key} in {p[0]} {agg}"
        m[key] = val
    assert checks == {"reconstructCol", "GenerateIndexes", "PrefixSum", "Sum", "CompressNullMask", "seeded",
                      "ReconstructRawNumbers", "empty", "block"}
    preds = [O.pred(0, f, threshold, O.INT)]
    for form, keys, kn in (("S", [a], [_pack(an)]), ("M", [a, b], [_pack(an), _pack(bn)])):
        for agg, code in AGG_CODES.items():
            out_types = [O.LONG] if code == O.AGG_COUNT else None
            g = O.GroupBy([code], [O.INT] * len(keys), O.INT, out_types)
            g.process(keys, v, _pack(vn), kn, preds=preds)
            want = as_map(g.results())
            assert got[(form, agg)] == want, f"{form} {agg} differs from the oracle"


def test_string_caller_known_answers():
    # the drop-in GPUGroupBy<AGG,O,std::string,V> from C++ on the reference's six string cases
    r = subprocess.run([B.STRING_DRIVER], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    checks = {l.split()[1] for l in r.stdout.splitlines() if l.startswith("CHECK")}
    assert checks == {"StringUnique", "StringSimple", "StringMultiBlockSimple", "StringMultiBlockMediumSum",
                      "StringMultiBlockMediumMin", "StringMultiBlockMediumMax"}
