"""The reader of qikkDB's on-disk format (qikkdb_amd.storage): a block file written here in
the documented layout reads back exactly (NULL bits, padding, several blocks, every numeric
type), and, where the reference tree is present, the sample database's column files agree with
the same rows' text in its CSV export and with the committed fixture."""
import csv
import datetime as dt
import os
import struct

import numpy as np
import pytest

from qikkdb_amd import storage as S

REF = "/root/reference/sampleData"
HERE = os.path.dirname(os.path.abspath(__file__))


def write_blocks(path, ctype, block_size, blocks):
    """blocks: list of (values, null flags) in the column file layout of Database.cpp."""
    fmt = {S.INT: "<iifi", S.LONG: "<qqfq", S.FLOAT: "<ffff", S.DOUBLE: "<ddfd", S.INT8: "<bbfb"}[ctype]
    dtp = np.dtype(S._NP[ctype])
    with open(path, "wb") as f:
        for i, (vals, nulls) in enumerate(blocks):
            mask = np.packbits(np.asarray(nulls, bool), bitorder="little").tobytes()
            f.write(struct.pack("<IiI", i, 0, len(mask)))
            f.write(mask)
            f.write(struct.pack("<Q?", len(vals), False))
            v = np.asarray(vals, dtp)
            lo = v.min() if v.size else 0
            hi = v.max() if v.size else 0
            f.write(struct.pack(fmt, lo, hi, float(v.mean()) if v.size else 0.0, v.sum(dtype=dtp).item() if v.size else 0))
            pad = np.zeros(block_size, dtp)
            pad[:v.size] = v
            f.write(pad.tobytes())


@pytest.mark.parametrize("ctype", [S.INT, S.LONG, S.FLOAT, S.DOUBLE, S.INT8])
def test_block_file_round_trip(tmp_path, ctype):
    rng = np.random.default_rng(ctype)
    bs = 64
    dtp = S._NP[ctype]
    blocks = [(rng.integers(-100, 100, n).astype(dtp), rng.random(n) < 0.2) for n in (64, 13, 0, 40)]
    p = tmp_path / "c.data"
    write_blocks(p, ctype, bs, blocks)
    got = S.read_column(str(p), ctype, bs)
    assert [b.index for b in got] == [0, 1, 2, 3]
    for (vals, nulls), b in zip(blocks, got):
        np.testing.assert_array_equal(b.data, vals)
        np.testing.assert_array_equal(b.nulls, nulls)
    with pytest.raises(ValueError):
        S.read_column(str(p), ctype, 32)  # a block larger than the block size
    with pytest.raises(NotImplementedError):
        S.read_column(str(p), S.STRING, bs)


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's sample database is not here")
def test_sample_database_matches_its_csv_and_the_fixture():
    t = S.load_database(os.path.join(REF, "db", "trips.db"))["trips_sample"]
    with open(os.path.join(REF, "csv", "trips_sample.csv")) as f:
        rows = list(csv.DictReader(f))
    n = len(rows)
    for name in ("total_amount", "trip_distance", "pickup_longitude", "pickup_latitude"):
        np.testing.assert_array_equal(t.columns[name].values(), np.array([r[name] for r in rows], np.float32))
    for name in ("cab_type", "passenger_count"):
        np.testing.assert_array_equal(t.columns[name].values(), np.array([r[name] for r in rows], np.int32))
    epoch = [int(dt.datetime.strptime(r["pickup_datetime"][:19], "%Y-%m-%d %H:%M:%S")
                 .replace(tzinfo=dt.timezone.utc).timestamp()) for r in rows]
    np.testing.assert_array_equal(t.columns["pickup_datetime"].values(), np.array(epoch, np.int64))
    for c in t.columns.values():
        if c.blocks:
            assert sum(b.data.shape[0] for b in c.blocks) == n
    fx = np.load(os.path.join(HERE, "golden", "trips_sample.npz"))
    for name in ("total_amount", "trip_distance", "cab_type", "passenger_count", "pickup_datetime"):
        np.testing.assert_array_equal(fx[name], t.columns[name].values())
    assert [d.decode() for d in fx["pickup_date"][:3]] == [r["pickup_datetime"][:10] for r in rows[:3]]
