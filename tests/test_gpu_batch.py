"""Batched queries (qk_batch_*): every query of a batch must equal the same query run alone
through the oracle, whatever the tuning, and replays of the captured graph must agree with
eager launches."""
import numpy as np
import pytest

import oracle as O
import qikkdb_amd as qk
from qikkdb_amd import batch as B
from qikkdb_amd import datagen

pytestmark = pytest.mark.gpu

N = 1 << 22
BLOCK = 1 << 18
B_RANGE = 4096


@pytest.fixture(scope="module")
def table():
    d = datagen.composite(N, seed=11)
    cols = {k: qk.DeviceBuffer.from_numpy(v) for k, v in d.items()}
    eng = B.BatchEngine(cols["a"].ptr, cols["b"].ptr, cols["ts"].ptr, cols["value"].ptr, N, BLOCK, B_RANGE)
    yield d, cols, eng
    eng.close()


def make_queries(n, seed):
    rng = np.random.default_rng(seed)
    nblk = N // BLOCK
    mods = [1, 2, 7, 1000, 2047, 2048, 4096, 4097, 5120, 5121, 65_536, 100_003, 1 << 24]
    qs = []
    for i in range(n):
        first = int(rng.integers(0, nblk))
        count = int(rng.integers(1, nblk - first + 1))
        sel = float(rng.uniform(0.001, 0.5))
        lo = datagen.YEAR_START_MS + int(rng.uniform(0, 1 - sel) * datagen.YEAR_MS)
        hi = lo + int(sel * datagen.YEAR_MS)
        qs.append(B.Query(first, count, lo, hi, mods[i % len(mods)]))
    return qs


def oracle_result(d, q):
    r0, r1 = q.first_block * BLOCK, min((q.first_block + q.nblocks) * BLOCK, N)
    a, b, ts, v = d["a"][r0:r1], d["b"][r0:r1], d["ts"][r0:r1], d["value"][r0:r1]
    key = (a.astype(np.int64) * B_RANGE + b) % q.key_mod
    g = O.GroupBy([O.AGG_SUM, O.AGG_MIN, O.AGG_MAX, O.AGG_COUNT], [O.LONG], O.LONG)
    g.process([np.ascontiguousarray(key)], np.ascontiguousarray(v),
              preds=[O.pred(1, ts, q.ts_lo, O.LONG), O.pred(2, ts, q.ts_hi, O.LONG)])
    r = g.results()
    return {int(k): (int(r["values"][0][i]), int(r["values"][1][i]), int(r["values"][2][i]), int(r["values"][3][i]))
            for i, k in enumerate(r["keys"][0])}


def check_all(d, eng, qs):
    for i, q in enumerate(qs):
        got = eng.result_host(i)
        want = oracle_result(d, q)
        assert got == want, f"query {i} ({q}) differs"


@pytest.mark.parametrize("tuning", [B.Tuning(1 << 18, 4096, 1, 1), B.Tuning(1 << 16, 2048, 2, 0),
                                    B.Tuning(1 << 19, 5120, 1, 0), B.Tuning(4096, 0, 1, 1),
                                    B.Tuning(1 << 16, 2048, 1, 1, scratch_rows=200_000),
                                    B.Tuning(1 << 16, 2048, 1, 0, scratch_rows=5_000)],
                         ids=["default", "small-tiles-2wg", "big-tiles", "all-two-pass", "waves",
                              "pool-overflow"])
def test_batch_matches_oracle(table, tuning):
    d, _, eng = table
    qs = make_queries(40, seed=1)
    eng.set_tuning(tuning)
    eng.run(qs)
    qk._abi.call("qk_device_synchronize")
    check_all(d, eng, qs)


def test_graph_replays_match_eager(table):
    d, _, eng = table
    qs = make_queries(24, seed=2)
    eng.set_tuning(B.Tuning(1 << 18, 4096, 1, 0))
    eng.run(qs)
    qk._abi.call("qk_device_synchronize")
    eager = [eng.result_host(i) for i in range(len(qs))]
    eng.set_tuning(B.Tuning(1 << 18, 4096, 1, 1))
    for rep in range(3):
        eng.run(qs)
        qk._abi.call("qk_device_synchronize")
        assert [eng.result_host(i) for i in range(len(qs))] == eager, f"replay {rep}"
    # a different batch through the same graph, then the first again
    other = make_queries(7, seed=3)
    eng.run(other)
    qk._abi.call("qk_device_synchronize")
    check_all(d, eng, other)
    eng.run(qs)
    qk._abi.call("qk_device_synchronize")
    assert [eng.result_host(i) for i in range(len(qs))] == eager


def test_completion_times_and_autotune(table):
    d, _, eng = table
    qs = make_queries(16, seed=4)
    t = eng.autotune(qs)
    assert t.tile_rows % 4096 == 0 and t.wgs_per_cu in (1, 2)
    eng.run(qs)
    c = eng.completion_ms()
    assert c.shape == (16,) and (c > 0).all()
    check_all(d, eng, qs)


def test_key_column_matches_numpy(table):
    d, cols, _ = table
    n = 100_000
    out = qk.DeviceBuffer(n * 8)
    for mod in (1, 3, 4096, 999_983, 1 << 24):
        B.key_column(cols["a"].ptr, cols["b"].ptr, B_RANGE, mod, n, out.ptr)
        got = out.to_numpy(np.int64, n)
        want = (d["a"][:n].astype(np.int64) * B_RANGE + d["b"][:n]) % mod
        np.testing.assert_array_equal(got, want)


def test_bad_queries_rejected(table):
    _, _, eng = table
    with pytest.raises(qk.QueryEngineError):
        eng.run([B.Query(0, 1, 0, 1, 0)])
    with pytest.raises(qk.QueryEngineError):
        eng.run([B.Query(N // BLOCK - 1, 2, 0, 1, 5)])
    with pytest.raises(qk.QueryEngineError):
        eng.set_tuning(B.Tuning(1000, 4096, 1, 1))
