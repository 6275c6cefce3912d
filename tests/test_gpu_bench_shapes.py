"""GPU parity at the exact shapes the benchmark times, every group compared with the oracle:

* zipf-256M: 2^28 rows in four 2^26-row blocks of an asynchronous table (phase 2 deferred over
  the query), 1e6 Zipf keys, f32 values with 5% NULLs, the BETWEEN filter; SUM/AVG within
  1e-5 relative of the FP64 oracle, COUNT bit-exact.
* composite-512M: 2^29 rows in eight 2^26-row blocks, every block deferred into one phase 2
  over 8192 owned sub-tables, 15.6M groups; bit-exact.
* batched-4096: the 4096-query mix over the same 512M-row table with the autotuned tuning;
  every group of every query equals the plain one-query-at-a-time path, and the queries of a
  stratified sample (every key_mod decade) equal the oracle.

The workloads, table flags and block loop are bench.py's own (imported), so these tests run
exactly what the benchmark times."""
import numpy as np
import pytest

import bench
import oracle as O
import qikkdb_amd as qk
from qikkdb_amd import batch as B, datagen
from helpers import compare_all_groups

pytestmark = pytest.mark.gpu
FP_REL = 1e-5


def threads():
    return bench.host_cpu()["threads"]


def run_bench_query(w, blk=1 << 26, repeats=2):
    """The benchmark's step, `repeats` times on one table (so the second query also runs on a
    table that was reset after a full query), returning the last step's groups on the host."""
    gb = w.table()
    step = bench.make_step(w, gb, blk)
    r = None
    for _ in range(repeats):
        del r
        r = step()
    host = bench.gb_result_host(qk, gb, r)
    stats = gb.stats()
    del r
    gb.close()
    return host, stats


def test_zipf_256M_all_groups_equal_oracle():
    print("case: zipf-256M, 4 x 2^26-row blocks, GB_ASYNC deferred phase 2", flush=True)
    w = bench.Zipf(qk, bench.Zipf.rows)
    try:
        got, stats = run_bench_query(w)
        assert stats["path"] == 1 and stats["sub_tables"] == 256 and stats["p1_form"] == 1
        want = w.cpu_query(O, w.n, threads())().results()
        assert got["count"] == want["count"] == 999_401
        compare_all_groups(got, want, rel_aggs=bench.Zipf.rel_aggs, rel=FP_REL)
    finally:
        w.free()


@pytest.fixture(scope="module")
def composite_table():
    print("case: generating the composite-512M table", flush=True)
    w = bench.Composite(qk, bench.Composite.rows)
    yield w
    w.free()


def test_composite_512M_all_groups_equal_oracle(composite_table):
    print("case: composite-512M, 8 x 2^26-row blocks, one deferred phase 2 over 8192 sub-tables", flush=True)
    w = composite_table
    got, stats = run_bench_query(w)
    assert stats["path"] == 1 and stats["sub_tables"] == 8192 and stats["partitions"] == 1024 and stats["p1_form"] == 1
    want = w.cpu_query(O, w.n, threads())().results()
    assert got["count"] == want["count"]
    assert 15_000_000 < got["count"] < 16_777_216
    compare_all_groups(got, want)


def oracle_batch_query(d, q, blk):
    """One query of the batched mix on the CPU oracle: key (a*4096+b) mod key_mod over the
    query's blocks, WHERE ts_lo <= ts < ts_hi, SUM/MIN/MAX/COUNT of value."""
    lo, hi = q.first_block * blk, (q.first_block + q.nblocks) * blk
    a, b = d["a"][lo:hi].astype(np.int64), d["b"][lo:hi].astype(np.int64)
    key = np.ascontiguousarray((a * 4096 + b) % q.key_mod)
    ts, v = d["ts"][lo:hi], d["value"][lo:hi]
    preds = [O.pred(1, ts, q.ts_lo, O.LONG), O.pred(2, ts, q.ts_hi, O.LONG)]
    g = O.query_mt(threads(), [O.AGG_COUNT, O.AGG_SUM, O.AGG_MIN, O.AGG_MAX], O.LONG, O.LONG, [key], v,
                   preds=preds)
    return g.results()


def test_batched_4096_all_queries_equal_plain_path_and_oracle(composite_table):
    print("case: batched-4096 over the 512M-row table (autotuned)", flush=True)
    w = composite_table
    blk = 1 << 24
    n = w.n
    qs = bench.batched_queries(4096, n // blk, seed=7)
    eng = B.BatchEngine(w.cols["a"].ptr, w.cols["b"].ptr, w.cols["ts"].ptr, w.cols["value"].ptr, n, blk, 4096)
    try:
        eng.autotune(qs[:256])
        eng.run(qs)
        qk._abi.call("qk_device_synchronize")
        # every query against the plain path
        keybuf = qk.DeviceBuffer(blk * 8)
        bad = []
        for i, q in enumerate(qs):
            r, gb = bench.plain_query(qk, B, w.cols, q, blk, keybuf)
            if not bench.batch_matches_plain(qk, eng, i, r):
                bad.append(i)
            del r
            gb.close()
        assert not bad, f"{len(bad)} queries differ from the plain path, first {bad[:5]}"
        # a stratified sample against the oracle: in each key_mod decade, the queries that
        # scan the fewest passing rows
        est = eng.estimate_rows(qs)
        picked = []
        for dec in range(7):
            idx = [i for i, q in enumerate(qs) if 10 ** dec <= q.key_mod < 10 ** (dec + 1)]
            idx.sort(key=lambda i: est[i])
            picked += idx[:8]
        assert len(picked) >= 40
        for i in picked:
            q = qs[i]
            want = oracle_batch_query(w.d, q, blk)
            c, s, mn, mx = eng.result_arrays(i)
            live = np.flatnonzero(c)
            got = {"count": live.size, "keys": [live.astype(np.int64)], "key_nulls": [np.zeros(live.size, bool)],
                   "values": [c[live], s[live], mn[live], mx[live]],
                   "value_nulls": [np.zeros(live.size, bool)] * 4}
            compare_all_groups(got, want)
    finally:
        eng.close()
