"""Writes tests/golden/trips_sample.npz from the reference's sample database.

The columns are read from the binary column files of qikkDB's sample database
(sampleData/db/trips.db and its .data files) with qikkdb_amd.storage; the date part of each
pickup time is taken from the same rows' text in sampleData/csv/trips_sample.csv, as a string
column for string-key GROUP BY.  Run from the repository root where the reference tree exists:

    python tests/golden/make_trips_fixture.py /root/reference
"""
import csv
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from qikkdb_amd import storage  # noqa: E402


def main(ref: str) -> None:
    t = storage.load_database(os.path.join(ref, "sampleData", "db", "trips.db"))["trips_sample"]
    out = {}
    for name, col in t.columns.items():
        if col.type == storage.POINT:
            continue
        out[name] = col.values()
        out[name + "__nulls"] = col.null_flags()
    with open(os.path.join(ref, "sampleData", "csv", "trips_sample.csv")) as f:
        dates = [row["pickup_datetime"][:10] for row in csv.DictReader(f)]
    out["pickup_date"] = np.array(dates, dtype="S10")
    np.savez_compressed(os.path.join(HERE, "trips_sample.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
