"""Serving loop (qk_serve_*): every request's groups are the same whether it runs alone or in
the mix of concurrently submitted requests, admission respects the memory cap, and the
open-loop replay serves every request."""
import threading

import numpy as np
import pytest

import oracle as O
import qikkdb_amd as qk
from qikkdb_amd import batch as B
from qikkdb_amd import datagen
from qikkdb_amd import serve as S

pytestmark = pytest.mark.gpu

N = 1 << 22
BLOCK = 1 << 18
B_RANGE = 4096
MODS = [1, 3, 100, 2047, 4096, 4097, 9_999, 65_536, 250_000, 1_000_000]


@pytest.fixture(scope="module")
def table():
    d = datagen.composite(N, seed=21)
    cols = {k: qk.DeviceBuffer.from_numpy(v) for k, v in d.items()}
    yield d, cols


def ptrs(cols):
    return cols["a"].ptr, cols["b"].ptr, cols["ts"].ptr, cols["value"].ptr


def make_queries(n, seed):
    rng = np.random.default_rng(seed)
    nblk = N // BLOCK
    qs = []
    for i in range(n):
        first = int(rng.integers(0, nblk))
        count = int(rng.integers(1, nblk - first + 1))
        sel = float(rng.uniform(0.001, 0.5))
        lo = datagen.YEAR_START_MS + int(rng.uniform(0, 1 - sel) * datagen.YEAR_MS)
        hi = lo + int(sel * datagen.YEAR_MS)
        qs.append(B.Query(first, count, lo, hi, MODS[i % len(MODS)]))
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


def alone(cols, qs):
    """Each query run by itself as a batch of one."""
    eng = B.BatchEngine(*ptrs(cols), N, BLOCK, B_RANGE)
    out = []
    try:
        for q in qs:
            eng.run([q])
            qk._abi.call("qk_device_synchronize")
            out.append(eng.result_host(0))
    finally:
        eng.close()
    return out


def test_results_same_alone_and_in_the_mix(table):
    d, cols = table
    qs = make_queries(48, seed=5)
    ref = alone(cols, qs)
    for i in (0, 7, 19, 33):  # the single-query runs themselves agree with the oracle
        assert ref[i] == oracle_result(d, qs[i]), f"query {i}"
    with S.Server(*ptrs(cols), N, BLOCK, B_RANGE, S.Config(engines=2, max_batch=16)) as srv:
        tickets = [None] * len(qs)

        def client(lo, hi):
            for i in range(lo, hi):
                tickets[i] = srv.submit(qs[i])

        th = [threading.Thread(target=client, args=(k * 12, (k + 1) * 12)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for i, t in enumerate(tickets):
            srv.wait(t, timeout_ms=120_000)
            assert srv.result_host(t, qs[i].key_mod) == ref[i], f"query {i} differs in the mix"
            assert srv.latency_ms(t) >= 0
            srv.release(t)
        st = srv.stats()
        assert st["served"] == len(qs) and st["failed"] == 0 and st["pending"] == 0
        assert st["batches"] >= len(qs) // 16 and st["max_batch"] <= 16
        assert st["admitted_bytes"] == 0


def test_release_in_any_order(table):
    """Results live in per-batch blocks, not in the engines: waiting for and releasing requests
    against the batch order (the last submitted first) completes, with few engines and small
    batches, and every result stays intact until its own release."""
    d, cols = table
    qs = make_queries(24, seed=8)
    ref = alone(cols, qs)
    with S.Server(*ptrs(cols), N, BLOCK, B_RANGE, S.Config(engines=2, max_batch=3)) as srv:
        tickets = [srv.submit(q) for q in qs]
        for i in reversed(range(len(qs))):
            srv.wait(tickets[i], timeout_ms=120_000)
        for i in reversed(range(len(qs))):
            assert srv.result_host(tickets[i], qs[i].key_mod) == ref[i], f"query {i}"
            srv.release(tickets[i])
        st = srv.stats()
        assert st["served"] == len(qs) and st["admitted_bytes"] == 0 and st["batches"] >= 8


def test_memory_cap_limits_admission(table):
    d, cols = table
    qs = [B.Query(0, N // BLOCK, datagen.YEAR_START_MS, datagen.YEAR_START_MS + datagen.YEAR_MS // 4, m)
          for m in (1000, 2000, 3000, 1000, 500)]
    cap = 3000 * 32 + 2000 * 32  # two of the requests at a time at most
    ref = alone(cols, qs)
    with S.Server(*ptrs(cols), N, BLOCK, B_RANGE, S.Config(engines=3, memory_cap=cap)) as srv:
        ts = [srv.submit(q) for q in qs]
        for i, t in enumerate(ts):
            srv.wait(t, timeout_ms=120_000)
            assert srv.result_host(t, qs[i].key_mod) == ref[i]
            srv.release(t)
        st = srv.stats()
        assert st["peak_admitted_bytes"] <= cap and st["served"] == len(qs)
        with pytest.raises(qk.QueryEngineError):
            srv.submit(B.Query(0, 1, 0, 1, 10_000))  # 320 kB of table alone exceeds the cap


def test_release_wait_and_bad_requests(table):
    _, cols = table
    with S.Server(*ptrs(cols), N, BLOCK, B_RANGE) as srv:
        t = srv.submit(B.Query(0, 2, datagen.YEAR_START_MS, datagen.YEAR_START_MS + datagen.YEAR_MS, 7))
        srv.wait(t, timeout_ms=60_000)
        assert sum(v[3] for v in srv.result_host(t, 7).values()) == 2 * BLOCK
        srv.release(t)
        with pytest.raises(qk.QueryEngineError):
            srv.release(t)
        with pytest.raises(qk.QueryEngineError):
            srv.wait(12345)
        with pytest.raises(qk.QueryEngineError):
            srv.submit(B.Query(N // BLOCK, 1, 0, 1, 5))  # outside the table
        with pytest.raises(qk.QueryEngineError):
            srv.submit(B.Query(0, 1, 0, 1, 0))


def test_open_loop_replay(table):
    _, cols = table
    rng = np.random.default_rng(9)
    qs = make_queries(300, seed=6)
    arrival = np.cumsum(rng.exponential(1000.0 / 2000.0, len(qs)))  # Poisson, 2000 requests/s
    with S.Server(*ptrs(cols), N, BLOCK, B_RANGE) as srv:
        lat = srv.replay(qs, arrival)
        st = srv.stats()
    assert (lat >= 0).all() and st["served"] == len(qs) and st["failed"] == 0
    assert st["batches"] >= 2


def test_served_rate_follows_offered_below_capacity(table):
    # capacity from a closed-loop replay (every request queued at once), then Poisson arrivals
    # at 40% of it: the loop keeps up, so the served rate is the offered rate
    _, cols = table
    with S.Server(*ptrs(cols), N, BLOCK, B_RANGE) as srv:
        qs = make_queries(400, seed=12)
        srv.replay(qs, np.zeros(len(qs)))  # warm-up
        lat = srv.replay(qs, np.zeros(len(qs)))
        capacity = len(qs) / (lat.max() / 1e3)
        offered = min(0.4 * capacity, 3000.0)
        m = int(offered * 1.5)
        rng = np.random.default_rng(13)
        arrival = np.cumsum(rng.exponential(1000.0 / offered, m))
        qs = make_queries(m, seed=14)
        print(f"case: capacity {capacity:.0f} q/s, offered {offered:.0f} q/s, {m} requests", flush=True)
        lat = srv.replay(qs, arrival)
    served = m / ((arrival + lat).max() / 1e3)
    assert (lat >= 0).all()
    assert served >= 0.9 * offered, (served, offered, capacity)


def test_plain_replay_serves_in_arrival_order(table):
    _, cols = table
    qs = make_queries(24, seed=15)
    arrival = np.linspace(0.0, 20.0, len(qs))
    lat = S.replay_plain(*ptrs(cols), N, BLOCK, B_RANGE, qs, arrival)
    done = arrival + lat
    assert (lat >= 0).all()
    assert (np.diff(done) >= 0).all()  # one worker, first come first served
    with pytest.raises(qk.QueryEngineError):
        S.replay_plain(*ptrs(cols), N, BLOCK, B_RANGE, qs[:2], np.array([5.0, 1.0]))
