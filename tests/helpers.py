"""Shared test helpers: golden fixtures, result comparison as maps (output order is
unspecified in the reference, S10)."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

AGG_NAMES = {"MIN": 0, "MAX": 1, "SUM": 2, "AVG": 3, "COUNT": 4, "NONE": 5, "COUNT_ALL": 6}
OP_NAMES = {"GREATER": 0, "GREATER_EQUAL": 1, "LESS": 2, "LESS_EQUAL": 3, "EQUAL": 4,
            "NOT_EQUAL": 5, "LOGICAL_AND": 6, "LOGICAL_OR": 7}


def golden() -> dict:
    with open(os.path.join(GOLDEN, "reference_known_answers.json")) as f:
        return json.load(f)


def pack_bits(isnull) -> np.ndarray:
    isnull = np.asarray(isnull, dtype=bool)
    words = (isnull.size + 63) // 64
    padded = np.zeros(max(1, words) * 64, dtype=bool)
    padded[: isnull.size] = isnull
    return np.packbits(padded, bitorder="little").view(np.uint64)


def unpack_bits(words, n) -> np.ndarray:
    return np.unpackbits(np.ascontiguousarray(words, np.uint64).view(np.uint8),
                         bitorder="little")[:n].astype(bool)


def as_map(res: dict, agg_index: int = 0) -> dict:
    """{key tuple (None for NULL columns): value or None} from a results dict with keys /
    key_nulls / values / value_nulls lists (oracle or GPU)."""
    g = res["count"]
    out = {}
    nk = len(res["keys"])
    for i in range(g):
        key = tuple(None if res["key_nulls"][c][i] else _py(res["keys"][c][i]) for c in range(nk))
        v = res["values"][agg_index]
        vn = res["value_nulls"][agg_index]
        val = None if (vn is not None and vn[i]) else (_py(v[i]) if v is not None else None)
        assert key not in out, f"duplicate group {key}"
        out[key] = val
    return out


def _py(x):
    return x.item() if hasattr(x, "item") else x


def expect_map(expect, composite: bool) -> dict:
    out = {}
    for k, v in expect:
        key = tuple(k) if composite else (k,)
        out[key] = v
    return out


def maps_close(got: dict, want: dict, rel: float = 0.0) -> None:
    assert set(got) == set(want), f"group keys differ: missing {set(want) - set(got)}, extra {set(got) - set(want)}"
    for k in want:
        a, b = got[k], want[k]
        if b is None or a is None:
            assert a is None and b is None, f"group {k}: {a} vs {b}"
        elif rel == 0.0:
            assert a == b, f"group {k}: {a} != {b}"
        else:
            assert abs(a - b) <= rel * max(abs(b), 1e-30), f"group {k}: {a} vs {b} (rel {rel})"


# the vectorised all-groups comparison lives with the oracle (the checker)
from oracle import compare_groups as compare_all_groups, sort_groups  # noqa: E402,F401
