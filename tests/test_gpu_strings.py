"""GPU GROUP BY over string keys (qk_gbs_*) against the reference's known answers and the
oracle: multiple blocks, NULL keys and values, WHERE, the empty string, byte strings that differ
only at the end, many distinct keys, and the full-table error."""
import numpy as np
import pytest

import oracle as O
import qikkdb_amd as qk
from qikkdb_amd.strings import DeviceStrings, GPUGroupByString
from helpers import AGG_NAMES, golden, pack_bits

pytestmark = pytest.mark.gpu
G = golden()


def gpu_string_group_by(aggs, blocks, values, key_nulls=None, value_nulls=None, vtype=qk.INT, where=None,
                        capacity=1 << 16):
    gb = GPUGroupByString(aggs, vtype, capacity)
    try:
        for b, (keys, vals) in enumerate(zip(blocks, values)):
            n = len(keys)
            ks = DeviceStrings.from_list(keys)
            vb = qk.DeviceBuffer.from_numpy(np.asarray(vals))
            kn = qk.DeviceBuffer.from_numpy(pack_bits(key_nulls[b])) if key_nulls is not None else None
            vn = qk.DeviceBuffer.from_numpy(pack_bits(value_nulls[b])) if value_nulls is not None else None
            if where is None:
                gb.ProcessBlock(ks, vb, n, kn, vn)
            else:
                col = qk.DeviceColumn(vb, qk._abi.TYPE_OF[np.asarray(vals).dtype], n, vn)
                gb.ProcessBlockWhere(ks, vb, n, [qk.where(where[0], col, qk.Const(where[1], qk.INT))], kn, vn)
        return gb.results_host()
    finally:
        gb.close()


@pytest.mark.parametrize("case", G["group_by_string"], ids=lambda c: c["source"].split()[1])
def test_string_known_answers(case):
    r = gpu_string_group_by(AGG_NAMES[case["agg"]], case["blocks"],
                            [np.array(v, np.int32) for v in case["values"]], capacity=8)
    got = {k.decode(): int(v) for k, v in zip(r["keys"], r["values"][0])}
    assert r["count"] == len(case["expect"]) and got == case["expect"]


def as_map(r):
    out = {}
    for i, k in enumerate(r["keys"]):
        assert k not in out, f"duplicate group {k!r}"
        out[k] = tuple(None if (r["value_nulls"][a] is not None and r["value_nulls"][a][i]) else
                       (float(r["values"][a][i]) if r["values"][a].dtype.kind == "f" else int(r["values"][a][i]))
                       for a in range(len(r["values"])))
    return out


def random_strings(rng, n, distinct, maxlen=24):
    alphabet = np.frombuffer(b"abcdefghijklmnopqrstuvwxyzABCDEFGHIJ0123456789 _-\x00\xc3\xa9", np.uint8)
    pool = []
    for d in range(distinct):
        ln = int(rng.integers(0, maxlen + 1)) if d else 0  # the empty string is one of the keys
        pool.append(bytes(alphabet[rng.integers(0, alphabet.size, ln)]))
    # suffix variants of a few keys: equal up to the last byte
    pool += [pool[1] + b"x", pool[1] + b"\x00", pool[2][:-1] if pool[2] else b"z"]
    pool = list(dict.fromkeys(pool))
    return [pool[i] for i in rng.integers(0, len(pool), n)]


@pytest.mark.parametrize("blocks", [1, 3])
@pytest.mark.parametrize("distinct", [5, 2000])
def test_random_strings_match_oracle(blocks, distinct):
    rng = np.random.default_rng(distinct + blocks)
    n = 30_000
    keys = random_strings(rng, n, distinct)
    vals = rng.integers(-1000, 1000, n).astype(np.int32)
    knull = rng.random(n) < 0.05
    vnull = rng.random(n) < 0.1
    aggs = [qk.AGG_SUM, qk.AGG_MIN, qk.AGG_MAX, qk.AGG_COUNT, qk.AGG_AVG]
    cuts = np.linspace(0, n, blocks + 1).astype(int)
    kb = [keys[a:b] for a, b in zip(cuts[:-1], cuts[1:])]
    vb = [vals[a:b] for a, b in zip(cuts[:-1], cuts[1:])]
    knb = [knull[a:b] for a, b in zip(cuts[:-1], cuts[1:])]
    vnb = [vnull[a:b] for a, b in zip(cuts[:-1], cuts[1:])]
    print(f"case: strings n={n} distinct={distinct} blocks={blocks}", flush=True)
    got = gpu_string_group_by(aggs, kb, vb, knb, vnb)
    o = O.GroupByStrings(aggs, O.INT)
    for k, v, kn_, vn_ in zip(kb, vb, knb, vnb):
        o.process(k, v, pack_bits(vn_), pack_bits(kn_))
    want = o.results()
    assert got["count"] == want["count"]
    assert as_map(got) == as_map(want)


def test_string_where_and_table_full():
    rng = np.random.default_rng(3)
    n = 5000
    keys = random_strings(rng, n, 300)
    vals = rng.integers(-50, 50, n).astype(np.int32)
    got = gpu_string_group_by([qk.AGG_SUM, qk.AGG_COUNT], [keys], [vals], where=(qk.GREATER, 0))
    o = O.GroupByStrings([O.AGG_SUM, O.AGG_COUNT], O.INT)
    o.process(keys, vals, preds=[O.pred(0, vals, 0, O.INT)])
    assert as_map(got) == as_map(o.results())
    # more distinct strings than the table's capacity
    with pytest.raises(qk.QueryEngineError) as e:
        gpu_string_group_by(qk.AGG_COUNT, [[b"k%d" % i for i in range(100)]], [np.ones(100, np.int32)],
                            capacity=10)
    assert e.value.code == qk.QK_HASH_TABLE_FULL
