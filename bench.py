"""Benchmark: filtered GROUP BY rows/s on one MI355X with the roofline of the dominant kernel,
batched queries/s at p99, and the CPU oracle timed on the host as a baseline (which also checks
every group of the GPU's results).

The default run measures, in one process:

* zipf-256M (the line's `value`): 256M rows in four 2^26-row blocks; key int64 Zipf(s=1.1)
  over 1e6 distinct values; value float32 lognormal(0,1) with 5% NULL values; WHERE value
  BETWEEN 0.5 AND 4.0; GROUP BY key with SUM / AVG / COUNT.
* composite-512M (`composite`): 512M rows of (a int32, b int32) keys over 16.7M possible
  groups, WHERE ts in a 30-day window of one year, SUM / MIN / MAX / COUNT of an int64 value.
* batched-4096 (`batched`): 4096 independent filtered GROUP BYs over the same device-resident
  512M-row table in 2^24-row blocks, one persistent kernel, against the plain one-query-at-a-
  time path.

Columns are generated on the host and copied to HBM before the timed region.  One step of a
query workload = clear the table, one fused WHERE+GROUP BY launch sequence per block,
GetResults into device buffers.  After the timed steps the CPU oracle runs the same query on
every row (timed: the CPU baseline) and every group of the last GPU step is compared with it;
a mismatch makes the run exit with status 1.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|zipf|composite|batched|serving]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "filtered group-by rows/sec and % HBM peak at 256M rows; batched queries/sec at p99"
HBM_PEAK_GBS = 8000.0
FP_REL = 1e-5  # FP32 SUM/AVG against the FP64 oracle


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(path: str):
    """HBM bytes per launch of the dominant kernel from a committed rocprof counter summary
    (a line 'traffic_bytes_per_launch: N'), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def host_cpu() -> dict:
    """CPU model and the CPUs this process may run on (the box's share, not the host's count)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    share = int(omp) if omp.isdigit() and int(omp) > 0 else avail
    return {"model": model, "cpus_visible": os.cpu_count() or 1, "cpus_available": avail,
            "threads": max(1, min(avail, share))}


class Zipf:
    """int64 Zipf key, f32 lognormal value with NULLs, WHERE value BETWEEN 0.5 AND 4.0."""
    name = "zipf"
    rows = 1 << 28
    bytes_per_row = 8 + 4 + 1.0 / 8  # int64 key + float32 value + value null bit
    description = ("zipf-256M: int64 key Zipf(s=1.1) over 1e6 distinct, f32 lognormal(0,1) "
                   "value with 5% NULLs, WHERE value BETWEEN 0.5 AND 4.0, SUM/AVG/COUNT")
    dtype = "f32"
    counters = "k_tp_phase1s_zipf256M_counters.txt"
    rel_aggs = (0, 1)  # SUM and AVG of float32 within FP_REL of the FP64 oracle

    def __init__(self, qk, n):
        from qikkdb_amd import datagen
        self.qk, self.n = qk, n
        self.d = datagen.zipf(n, seed=2)
        self.key = qk.DeviceBuffer.from_numpy(self.d["key"])
        self.val = qk.DeviceBuffer.from_numpy(self.d["value"])
        self.vnull = qk.DeviceBuffer.from_numpy(self.d["value_nulls"])
        self.aggs = [qk.AGG_SUM, qk.AGG_AVG, qk.AGG_COUNT]

    def table(self):
        qk = self.qk
        return qk.GPUGroupBy(self.aggs, qk.LONG, qk.FLOAT, 1 << 20, 1,
                             flags=qk.GB_KEYS_NOT_NULL | qk.GB_ASYNC)

    def block(self, off, blk):
        qk = self.qk
        vp = self.val.ptr + off * 4
        np_ = self.vnull.ptr + (off // 64) * 8
        vcol = qk.DeviceColumn(qk.DeviceBuffer(blk * 4, ptr=vp, owned=False), qk.FLOAT, blk,
                               qk.DeviceBuffer(blk // 8, ptr=np_, owned=False))
        preds = [qk.where(qk.GREATER_EQUAL, vcol, qk.Const(0.5, qk.FLOAT)),
                 qk.where(qk.LESS_EQUAL, vcol, qk.Const(4.0, qk.FLOAT))]
        return [self.key.ptr + off * 8], vp, np_, preds

    def passing(self):
        v = self.d["value"]
        isnull = np.unpackbits(self.d["value_nulls"].view(np.uint8), bitorder="little")[:self.n].astype(bool)
        return int(np.count_nonzero((v >= np.float32(0.5)) & (v <= np.float32(4.0)) & ~isnull))

    def cpu_query(self, O, rows, nthreads):
        keys = self.d["key"][:rows]
        vals = self.d["value"][:rows]
        vnull = self.d["value_nulls"][: (rows + 63) // 64]
        preds = [O.pred(1, vals, 0.5, O.FLOAT, vnull), O.pred(3, vals, 4.0, O.FLOAT, vnull)]
        if nthreads == 1:
            def one():
                g = O.GroupBy([O.AGG_SUM, O.AGG_AVG, O.AGG_COUNT], O.LONG, O.FLOAT)
                g.process([keys], vals, vnull, preds=preds)
                return g
            return one
        return lambda: O.query_mt(nthreads, [O.AGG_SUM, O.AGG_AVG, O.AGG_COUNT], O.LONG, O.FLOAT,
                                  [keys], vals, vnull, preds=preds)

    def free(self):
        for b in (self.key, self.val, self.vnull):
            b.free()
        self.d = None


class Composite:
    """(a int32, b int32) key, WHERE ts in a 30-day window, SUM/MIN/MAX/COUNT of int64."""
    name = "composite"
    rows = 1 << 29
    bytes_per_row = 4 + 4 + 8 + 8  # a + b + ts + value
    description = ("composite-512M: keys (int32 a, int32 b) uniform [0,4096)^2, WHERE int64 ts in "
                   "a 30-day window of one year, SUM/MIN/MAX/COUNT of int64 value")
    dtype = "int64"
    counters = "k_tp_phase1s_composite512M_counters.txt"
    rel_aggs = ()

    def __init__(self, qk, n):
        from qikkdb_amd import datagen
        self.qk, self.n, self.dg = qk, n, datagen
        self.d = datagen.composite(n, seed=3)
        self.cols = {k: qk.DeviceBuffer.from_numpy(v) for k, v in self.d.items()}
        self.aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT]

    def table(self):
        qk = self.qk
        t = qk.GPUGroupBy(self.aggs, [qk.INT, qk.INT], qk.LONG, 1 << 24, 1,
                          flags=qk.GB_KEYS_NOT_NULL | qk.GB_ASYNC)
        t.set_defer_rows(self.n)  # the query's 8 blocks share one phase 2 (about 13 GB of regions)
        return t

    def block(self, off, blk):
        qk = self.qk
        ts = qk.DeviceColumn(qk.DeviceBuffer(blk * 8, ptr=self.cols["ts"].ptr + off * 8, owned=False),
                             qk.LONG, blk)
        preds = [qk.where(qk.GREATER_EQUAL, ts, qk.Const(self.dg.WINDOW_START_MS, qk.LONG)),
                 qk.where(qk.LESS, ts, qk.Const(self.dg.WINDOW_END_MS, qk.LONG))]
        return ([self.cols["a"].ptr + off * 4, self.cols["b"].ptr + off * 4],
                self.cols["value"].ptr + off * 8, None, preds)

    def passing(self):
        ts = self.d["ts"]
        return int(np.count_nonzero((ts >= self.dg.WINDOW_START_MS) & (ts < self.dg.WINDOW_END_MS)))

    def cpu_query(self, O, rows, nthreads):
        a, b = self.d["a"][:rows], self.d["b"][:rows]
        ts, v = self.d["ts"][:rows], self.d["value"][:rows]
        preds = [O.pred(1, ts, self.dg.WINDOW_START_MS, O.LONG), O.pred(2, ts, self.dg.WINDOW_END_MS, O.LONG)]
        aggs = [O.AGG_SUM, O.AGG_MIN, O.AGG_MAX, O.AGG_COUNT]
        if nthreads == 1:
            def one():
                g = O.GroupBy(aggs, [O.INT, O.INT], O.LONG)
                g.process([a, b], v, preds=preds)
                return g
            return one
        return lambda: O.query_mt(nthreads, aggs, [O.INT, O.INT], O.LONG, [a, b], v, preds=preds)

    def free(self):
        for b in self.cols.values():
            b.free()
        self.d = None


WORKLOADS = {"zipf": Zipf, "composite": Composite}


def cpu_baseline_and_parity(w, gpu_host: dict, single_rows: int, check: bool = True) -> tuple[dict, str]:
    """The CPU baseline leg: the multithreaded oracle over every row of the workload (timed),
    then the single-threaded oracle on a sample (timed).  The multithreaded run's groups are
    the checker of the GPU's last step: every group is compared (integers bit-exact, FP32
    SUM/AVG within FP_REL of FP64)."""
    import oracle as O

    cpu = host_cpu()
    nthreads = cpu["threads"]
    run = w.cpu_query(O, w.n, nthreads)
    t0 = time.perf_counter()
    g = run()
    mt_s = time.perf_counter() - t0
    parity = "not checked"
    if check:
        want = g.results()
        try:
            O.compare_groups(gpu_host, want, rel_aggs=w.rel_aggs, rel=FP_REL)
            parity = f"all groups equal ({want['count']} groups)"
        except AssertionError as e:
            parity = f"MISMATCH: {e}"
        del want
    del g
    rows1 = min(single_rows, w.n)
    run1 = w.cpu_query(O, rows1, 1)
    t0 = time.perf_counter()
    g1 = run1()
    st_s = time.perf_counter() - t0
    del g1
    base = {"value": w.n / mt_s, "unit": "rows/s", "cores": nthreads, "kind": "port",
            "sample": f"all {w.n} rows of the same workload, {nthreads} threads (one table per thread, "
                      f"merged); {mt_s:.2f} s wall",
            "cpu_model": cpu["model"], "cpus_visible": cpu["cpus_visible"],
            "cpus_available": cpu["cpus_available"],
            "single_thread": {"value": rows1 / st_s, "unit": "rows/s", "rows": rows1, "seconds": st_s}}
    return base, parity


def make_step(w, gb, blk: int):
    """One query of workload w on table gb: clear the table, one fused WHERE + GROUP BY per
    block of blk rows, GetResults into device buffers."""
    blocks = [w.block(b * blk, blk) for b in range(w.n // blk)]

    def step():
        gb.reset()
        for kp, vp, np_, preds in blocks:
            gb.ProcessBlockWhere(kp, vp, blk, preds, None, np_)
        return gb.GetResults(want_nulls=True)
    return step


def run_query(W, qk, args, blk: int, steps: int, warmup: int, keep: bool = False) -> dict:
    """Times `steps` queries of workload W (after `warmup`), then the CPU baseline and the check
    of every group.  Returns the result dict (with `_w` = the workload when keep is set)."""
    n = args.rows or W.rows
    assert n % blk == 0 and blk % 64 == 0
    nblocks = n // blk
    t0 = time.perf_counter()
    w = W(qk, n)
    log(f"{W.name}: generated {n} rows and copied {n * W.bytes_per_row / 1e9:.2f} GB to HBM in "
        f"{time.perf_counter() - t0:.1f} s")
    if args.copy_bytes > 0:
        src = qk.DeviceBuffer(args.copy_bytes)
        dst = qk.DeviceBuffer(args.copy_bytes)
        qk._abi.call("qk_memcpy_dtod", dst.ptr, src.ptr, args.copy_bytes, None)
        qk._abi.call("qk_device_synchronize")
        log(f"reference copy of {args.copy_bytes} bytes done")
        del src, dst

    gb = w.table()
    step = make_step(w, gb, blk)
    for i in range(warmup):
        r = step()
        log(f"{W.name} warmup {i + 1}/{warmup}: {r.count} groups")
        del r
    qk._abi.call("qk_device_synchronize")
    gb.set_profiling(True)
    gb.profile()  # clear
    t0 = time.perf_counter()
    r = None
    for _ in range(steps):
        del r
        r = step()
    qk._abi.call("qk_device_synchronize")
    elapsed = time.perf_counter() - t0
    kern_ms, launches = gb.profile()
    gb.set_profiling(False)
    stats = gb.stats()
    kernel = ("k_tp_phase1s" if stats.get("p1_form") == 1 else "k_tp_phase1") if stats["path"] == 1 else "k_gb_rows"
    ms_per_step = elapsed * 1e3 / steps
    rows_per_s = n * steps / elapsed
    avg_launch_ms = kern_ms / max(1, launches)
    achieved = blk * W.bytes_per_row / (avg_launch_ms * 1e-3) / 1e9
    query_gbs = n * W.bytes_per_row / (ms_per_step * 1e-3) / 1e9
    log(f"{W.name}: {steps} steps: {ms_per_step:.3f} ms/step ({query_gbs:.0f} GB/s algorithmic over the "
        f"query), {kernel} {avg_launch_ms:.3f} ms per {blk}-row launch over {launches} launches, "
        f"{achieved:.0f} GB/s algorithmic")
    log(f"{W.name}: table state: {stats}")

    # the last step's groups, on the host
    host = gb_result_host(qk, gb, r)
    del r
    passed = w.passing()
    cnt_sum = int(np.asarray(host["values"][w.aggs.index(qk.AGG_COUNT)]).sum())
    traffic_file = os.path.join("profiles", W.counters)
    out = {
        "metric": METRIC,
        "value": rows_per_s,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": W.dtype,
        "data": "synthetic",
        "config": {
            "workload": W.description,
            "rows": n, "block_rows": blk, "blocks": nblocks, "groups": host["count"],
            "accumulator": "f64" if W.dtype == "f32" else "int64", "passed_rows": passed,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": read_traffic(os.path.join(ROOT, traffic_file)),
            "traffic_source": f"{traffic_file} (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this "
                              "workload, scaled by a 1 GiB device copy)",
            "kernel": kernel,
            "avg_launch_ms": avg_launch_ms,
            "algorithmic_bytes_per_launch": blk * W.bytes_per_row,
            "query_achieved": query_gbs,
            "query_frac": query_gbs / HBM_PEAK_GBS,
        },
        "cpu_baseline": None,
        "parity": "not checked",
    }
    if cnt_sum != passed:
        out["parity"] = f"MISMATCH: counts sum to {cnt_sum}, {passed} rows pass the filter"
    elif not args.no_cpu_baseline:
        log(f"{W.name}: running the CPU oracle on all {n} rows ({host_cpu()['threads']} threads) and "
            f"comparing every group")
        out["cpu_baseline"], out["parity"] = cpu_baseline_and_parity(w, host, args.cpu_single_rows)
        log(f"{W.name}: CPU {out['cpu_baseline']['value'] / 1e6:.0f} M rows/s on "
            f"{out['cpu_baseline']['cores']} threads; parity: {out['parity']}")
    gb.close()
    del host
    if keep:
        out["_w"] = w
    else:
        w.free()
    return out


def gb_result_host(qk, gb, r) -> dict:
    """A GroupByResult copied to the host (numpy keys / values, bool NULL flags)."""
    from qikkdb_amd.memory import unpack_nulls
    g = r.count
    out = {"count": g, "keys": [], "key_nulls": [], "values": [], "value_nulls": []}
    for i, t in enumerate(gb.key_types):
        out["keys"].append(r.keys[i].to_numpy(qk._abi.NP_OF[t], g) if g else np.empty(0, qk._abi.NP_OF[t]))
        kn = r.key_nulls[i]
        out["key_nulls"].append(unpack_nulls(kn.to_numpy(np.uint64), g) if (kn is not None and g) else np.zeros(g, bool))
    for a, t in enumerate(gb.out_types):
        if r.values[a] is None:
            out["values"].append(None)
            out["value_nulls"].append(None)
            continue
        out["values"].append(r.values[a].to_numpy(qk._abi.NP_OF[t], g))
        vn = r.value_nulls[a]
        out["value_nulls"].append(unpack_nulls(vn.to_numpy(np.uint64), g) if vn is not None else np.zeros(g, bool))
    return out


def batched_queries(nq: int, nblocks: int, seed: int):
    """The batched query mix: per query a contiguous block range of log-uniform length, a ts
    window of selectivity U(0.1%, 50%) placed uniformly in the year, and a key modulus
    log-uniform in [1, 1e6]."""
    from qikkdb_amd import batch as B, datagen
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(nq):
        length = int(min(nblocks, max(1, np.floor(np.exp(rng.uniform(0, np.log(nblocks + 1)))))))
        first = int(rng.integers(0, nblocks - length + 1))
        sel = float(rng.uniform(0.001, 0.5))
        lo = datagen.YEAR_START_MS + int(rng.uniform(0, 1 - sel) * datagen.YEAR_MS)
        hi = lo + int(sel * datagen.YEAR_MS)
        mod = int(max(1, np.floor(10 ** rng.uniform(0, 6))))
        qs.append(B.Query(first, length, lo, hi, mod))
    return qs


def plain_query(qk, B, cols, q, blk, keybuf):
    """One query of the batched mix alone through the drop-in GPUGroupBy (per block: key
    column, then fused WHERE + GROUP BY); returns its GroupByResult and table."""
    gb = qk.GPUGroupBy([qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT], qk.LONG, qk.LONG,
                       q.key_mod, 1, flags=qk.GB_KEYS_NOT_NULL)
    for bi in range(q.first_block, q.first_block + q.nblocks):
        off = bi * blk
        B.key_column(cols["a"].ptr + off * 4, cols["b"].ptr + off * 4, 4096, q.key_mod, blk, keybuf.ptr)
        ts = qk.DeviceColumn(qk.DeviceBuffer(blk * 8, ptr=cols["ts"].ptr + off * 8, owned=False), qk.LONG, blk)
        preds = [qk.where(qk.GREATER_EQUAL, ts, qk.Const(q.ts_lo, qk.LONG)),
                 qk.where(qk.LESS, ts, qk.Const(q.ts_hi, qk.LONG))]
        gb.ProcessBlockWhere([keybuf.ptr], cols["value"].ptr + off * 8, blk, preds)
    return gb.GetResults(), gb


def batch_matches_plain(qk, eng, i: int, r) -> bool:
    """Every group of query i of the engine's last run equals the plain path's group."""
    dense = eng.result_arrays(i)
    live = np.flatnonzero(dense[0])
    if r.count != live.size:
        return False
    if r.count == 0:
        return True
    keys = r.keys[0].to_numpy(np.int64, r.count)
    order = np.argsort(keys)
    keys = keys[order]
    if not np.array_equal(keys, live):
        return False
    # plain aggregates in order SUM, MIN, MAX, COUNT; dense arrays in order COUNT, SUM, MIN, MAX
    for a, d in ((0, 1), (1, 2), (2, 3), (3, 0)):
        v = r.values[a].to_numpy(np.int64, r.count)[order]
        if not np.array_equal(v, dense[d][live]):
            return False
    return True


def plain_path(cols, qs, blk, eng=None):
    """Runs each query alone through the drop-in GPUGroupBy and returns the per-query wall
    times (s) and, with `eng`, how many differ in any group from query i of the engine's
    last run."""
    import qikkdb_amd as qk
    from qikkdb_amd import batch as B

    keybuf = qk.DeviceBuffer(blk * 8)
    lat = []
    mism = 0
    for i, q in enumerate(qs):
        t0 = time.perf_counter()
        r, gb = plain_query(qk, B, cols, q, blk, keybuf)
        lat.append(time.perf_counter() - t0)
        if eng is not None and not batch_matches_plain(qk, eng, i, r):
            mism += 1
        del r
        gb.close()
    return lat, mism


def run_batched_on(qk, cols, n: int, args) -> dict:
    """4096 queries over the device-resident 512M-row table in 2^24-row blocks, all in one
    persistent kernel (autotuned), against the plain one-query-at-a-time path; the plain path's
    queries are also compared group by group with the batch's."""
    from qikkdb_amd import batch as B

    blk = 1 << 24
    nblocks = n // blk
    qs = batched_queries(args.queries, nblocks, seed=7)
    eng = B.BatchEngine(cols["a"].ptr, cols["b"].ptr, cols["ts"].ptr, cols["value"].ptr, n, blk, 4096)
    t0 = time.perf_counter()
    if args.batch_tuning:
        tile, groups, wpc = (int(x) for x in args.batch_tuning.split(","))
        eng.set_tuning(B.Tuning(tile, groups, wpc, 1))
        tuned = eng.tuning()
        log(f"batched: tuning set: {tuned}")
    else:
        tuned = eng.autotune(qs[:256])
        log(f"batched: autotuned on 256 queries in {time.perf_counter() - t0:.1f} s: {tuned}")
    for _ in range(args.batched_warmup):
        eng.run(qs)
    qk._abi.call("qk_device_synchronize")
    times = []
    for _ in range(args.batched_steps):
        t0 = time.perf_counter()
        eng.run(qs)
        qk._abi.call("qk_device_synchronize")
        times.append(time.perf_counter() - t0)
    comp = eng.completion_ms()
    elapsed = float(np.median(times))
    qps = len(qs) / elapsed
    rows_scanned = sum(min((q.first_block + q.nblocks) * blk, n) - q.first_block * blk for q in qs)
    log(f"batched: {len(qs)} queries in {elapsed * 1e3:.1f} ms (median of {len(times)}): {qps:.0f} q/s, "
        f"{rows_scanned / elapsed / 1e9:.1f} G rows/s scanned; completion p50 {np.percentile(comp, 50):.1f} ms, "
        f"p99 {np.percentile(comp, 99):.1f} ms")
    nplain = min(args.plain_queries, len(qs))
    t_all = time.perf_counter()
    lat, mism = plain_path(cols, qs[:nplain], blk, eng)
    plain_elapsed = time.perf_counter() - t_all
    lat_ms = np.array(lat) * 1e3
    plain_qps = nplain / plain_elapsed
    log(f"batched: plain path: {nplain} queries in {plain_elapsed:.2f} s: {plain_qps:.0f} q/s, per-query p50 "
        f"{np.percentile(lat_ms, 50):.1f} ms, p99 {np.percentile(lat_ms, 99):.1f} ms; "
        f"{mism} of {nplain} differ from the batch in some group")
    eng.close()
    return {
        "workload": "batched-4096: 4096 queries over a device-resident 512M-row table (int32 a, int32 b, "
                    "int64 ts, int64 value) in 2^24-row blocks; per query a log-uniform block range, ts "
                    "selectivity U(0.1%,50%), key (a*4096+b) mod C with C log-uniform in [1,1e6]; "
                    "SUM/MIN/MAX/COUNT",
        "queries": len(qs), "rows": n, "block_rows": blk, "rows_scanned": rows_scanned,
        "q_per_s": qps, "ms_per_batch": elapsed * 1e3, "steps": len(times),
        "p50_ms": float(np.percentile(comp, 50)), "p99_ms": float(np.percentile(comp, 99)),
        "latency_basis": "completion time of each query from the start of its batch",
        "plain_q_per_s": plain_qps, "plain_queries": nplain,
        "plain_p50_ms": float(np.percentile(lat_ms, 50)), "plain_p99_ms": float(np.percentile(lat_ms, 99)),
        "ratio": qps / plain_qps,
        "parity": (f"all groups of {nplain} queries equal the plain path's" if mism == 0
                   else f"MISMATCH: {mism} of {nplain} queries differ from the plain path"),
        "tuning": tuned.__dict__,
    }


def plain_ladder(S, ptrs, n, blk, nblocks, args, rng) -> list:
    """The plain one-request-at-a-time path measured under open-loop Poisson arrivals
    (qk_serve_replay_plain) at rates up to its capacity: [(rate, p50, p99, served q/s)]."""
    # capacity: requests back to back
    qs = batched_queries(args.plain_queries, nblocks, seed=99)
    lat = S.replay_plain(*ptrs, n, blk, 4096, qs, np.zeros(len(qs)))
    cap = len(qs) / (lat.max() / 1e3)
    log(f"plain path: {len(qs)} requests back to back in {lat.max():.0f} ms: capacity {cap:.0f} q/s")
    out = []
    for frac in (0.25, 0.5, 0.7, 0.85, 0.95):
        rate = frac * cap
        m = max(64, int(rate * args.serve_seconds))
        qs = batched_queries(m, nblocks, seed=2000 + int(rate))
        arrival = np.cumsum(rng.exponential(1000.0 / rate, m))
        lat = S.replay_plain(*ptrs, n, blk, 4096, qs, arrival)
        served = m / ((arrival + lat).max() / 1e3)
        p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
        out.append({"offered_qps": rate, "requests": m, "served_qps": served, "p50_ms": p50, "p99_ms": p99})
        log(f"plain path at {rate:.0f} q/s offered: served {served:.0f} q/s, p50 {p50:.1f} ms, p99 {p99:.1f} ms")
    return out


def run_serving(args) -> int:
    """Open-loop Poisson arrivals from the batched query mix through the serving loop at several
    rates, against the plain one-request-at-a-time path measured under Poisson arrivals of its
    own up to its capacity: the comparison is the plain path's highest measured rate whose p99
    is within the serving loop's p99."""
    import qikkdb_amd as qk
    from qikkdb_amd import batch as B, datagen, serve as S

    qk.ensure_device(0)
    n = args.rows or (1 << 29)
    blk = 1 << 24
    nblocks = n // blk
    t0 = time.perf_counter()
    d = datagen.composite(n, seed=3)
    cols = {k: qk.DeviceBuffer.from_numpy(v) for k, v in d.items()}
    del d
    log(f"generated and copied {n} rows ({n * 24 / 1e9:.1f} GB) in {time.perf_counter() - t0:.1f} s")
    ptrs = (cols["a"].ptr, cols["b"].ptr, cols["ts"].ptr, cols["value"].ptr)
    tune_qs = batched_queries(256, nblocks, seed=7)
    eng = B.BatchEngine(*ptrs, n, blk, 4096)
    if args.batch_tuning:
        tile, groups, wpc = (int(x) for x in args.batch_tuning.split(","))
        tuned = B.Tuning(tile, groups, wpc, 0)
    else:
        tuned = eng.autotune(tune_qs)
    eng.close()
    tuned = B.Tuning(tuned.tile_rows, tuned.lds_groups, tuned.wgs_per_cu, 0, 0)
    log(f"engine tuning: {tuned}")
    rng = np.random.default_rng(11)
    ladder = plain_ladder(S, ptrs, n, blk, nblocks, args, rng)
    rates = [float(x) for x in args.rates.split(",")]
    rows = []
    with S.Server(*ptrs, n, blk, 4096, S.Config(engines=args.engines, tuning=tuned, wait_ms=args.wait_ms)) as srv:
        srv.replay(tune_qs, np.zeros(len(tune_qs)))  # warm-up: buffers and pools
        for lam in rates:
            m = int(lam * args.serve_seconds)
            qs = batched_queries(m, nblocks, seed=1000 + int(lam))
            arrival = np.cumsum(rng.exponential(1000.0 / lam, m))
            before = srv.stats()
            t0 = time.perf_counter()
            lat_ms = srv.replay(qs, arrival)
            wall = time.perf_counter() - t0
            after = srv.stats()
            qps = m / wall
            p50, p99 = float(np.percentile(lat_ms, 50)), float(np.percentile(lat_ms, 99))
            ok = [r for r in ladder if r["p99_ms"] <= p99]
            best = max(ok, key=lambda r: r["served_qps"]) if ok else None
            plain_qps = best["served_qps"] if best else 0.0
            rows.append({"offered_qps": lam, "requests": m, "served_qps": qps, "p50_ms": p50, "p99_ms": p99,
                         "batches": after["batches"] - before["batches"],
                         "peak_admitted_gb": after["peak_admitted_bytes"] / 1e9,
                         "plain_qps_at_equal_p99": plain_qps,
                         "ratio_at_equal_p99": qps / plain_qps if plain_qps > 0 else None})
            cmp = (f"plain path sustains {plain_qps:.0f} q/s within p99 {p99:.0f} ms (measured at "
                   f"{best['offered_qps']:.0f} q/s offered, p99 {best['p99_ms']:.0f} ms) -> {qps / plain_qps:.1f}x"
                   if best else f"no measured plain rate reaches a p99 of {p99:.0f} ms")
            log(f"rate {lam:.0f} q/s: {m} requests served at {qps:.0f} q/s in {after['batches'] - before['batches']} "
                f"batches, latency p50 {p50:.0f} ms p99 {p99:.0f} ms; {cmp}")
    best = max(rows, key=lambda r: r["served_qps"])
    out = {
        "metric": METRIC,
        "value": best["served_qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 1,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "serving: open-loop Poisson arrivals drawn from the batched query mix over "
                               "the device-resident 512M-row table; continuous batching, admission by "
                               "projected table bytes under a 200 GB cap",
                   "rows": n, "block_rows": blk, "engines": args.engines, "wait_ms": args.wait_ms,
                   "seconds_per_rate": args.serve_seconds,
                   "tuning": tuned.__dict__},
        "serving": rows,
        "plain_path": {"discipline": "one request at a time in arrival order (qk_serve_replay_plain), "
                                     "Poisson arrivals at fractions of its measured capacity",
                       "rates": ladder},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def summary(sub: dict) -> dict:
    """The fields of a query workload's result kept as a sub-result of the main line."""
    keep = ("value", "unit", "ms_per_step", "steps", "warmup", "dtype", "config", "roofline",
            "cpu_baseline", "parity")
    return {k: sub[k] for k in keep}


def failed(res: dict) -> bool:
    return str(res.get("parity", "")).startswith("MISMATCH")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["all"] + sorted(WORKLOADS) + ["batched", "serving"], default="all")
    ap.add_argument("--queries", type=int, default=4096, help="batched workload: queries per batch")
    ap.add_argument("--batched-steps", type=int, default=3, help="batched workload: timed batches")
    ap.add_argument("--batched-warmup", type=int, default=1, help="batched workload: untimed batches")
    ap.add_argument("--plain-queries", type=int, default=128,
                    help="batched workload: queries timed on the one-at-a-time path (and compared)")
    ap.add_argument("--rates", default="500,1000,1500,2000,3000,4000", help="serving workload: arrival rates (q/s)")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="serving workload: seconds of arrivals per rate")
    ap.add_argument("--engines", type=int, default=2, help="serving workload: batch engines")
    ap.add_argument("--wait-ms", type=float, default=0.0,
                    help="serving workload: wait of a second batch (0 = half the last batch's time)")
    ap.add_argument("--batch-tuning", default="",
                    help="batched workload: 'tile_rows,lds_groups,wgs_per_cu' instead of autotuning")
    ap.add_argument("--rows", type=int, default=0, help="rows (default: the workload's size)")
    ap.add_argument("--block-rows", type=int, default=1 << 26)
    ap.add_argument("--cpu-single-rows", type=int, default=1 << 24,
                    help="rows of the single-threaded CPU oracle timing")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU oracle (and with it the check of every group)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run one plain device-to-device copy of this size before the steps "
                         "(a known-size reference for HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by row blocks across GPUs by design, but only the single-GPU "
              "kernel is built and measured)", file=sys.stderr)
        return 2

    if args.workload == "serving":
        return run_serving(args)

    import qikkdb_amd as qk

    qk.ensure_device(0)
    if args.workload == "batched":
        from qikkdb_amd import datagen
        n = args.rows or Composite.rows
        d = datagen.composite(n, seed=3)
        cols = {k: qk.DeviceBuffer.from_numpy(v) for k, v in d.items()}
        del d
        b = run_batched_on(qk, cols, n, args)
        out = {"metric": METRIC, "value": b["q_per_s"], "unit": "queries/s", "n_gpus": 1,
               "steps": b["steps"], "warmup": args.batched_warmup, "ms_per_step": b["ms_per_batch"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
               "data": "synthetic", "config": {"workload": b["workload"], "rows": n},
               "roofline": None, "cpu_baseline": None, "batched": b, "parity": b["parity"]}
        print(json.dumps(out), flush=True)
        return 1 if failed(b) else 0

    if args.workload in WORKLOADS:
        out = run_query(WORKLOADS[args.workload], qk, args, args.block_rows, args.steps, args.warmup)
        print(json.dumps(out), flush=True)
        return 1 if failed(out) else 0

    # all: zipf-256M is the line; composite-512M and the batch over the same 512M-row table
    # are sub-results of it
    out = run_query(Zipf, qk, args, args.block_rows, args.steps, args.warmup)
    comp = run_query(Composite, qk, args, args.block_rows, max(3, args.steps // 2), 1, keep=True)
    w = comp.pop("_w")
    out["composite"] = summary(comp)
    out["batched"] = run_batched_on(qk, w.cols, w.n, args)
    w.free()
    print(json.dumps(out), flush=True)
    return 1 if (failed(out) or failed(comp) or failed(out["batched"])) else 0


if __name__ == "__main__":
    sys.exit(main())
