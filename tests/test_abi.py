"""CPU checks of the drop-in boundary: the HIP library loads (no GPU call) and exports every
entry point that include/qikkdb_amd.h declares; the host bindings cover all of them; the
data generator is deterministic and has the distributions the workloads name."""
import ctypes
import os
import re

import numpy as np
import pytest

from qikkdb_amd import _abi, datagen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "qikkdb_amd.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(qk_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "qk_gb_process_block" in names and "qk_reconstruct_col_keep" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)  # loading does not touch the GPU
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_bindings_cover_header():
    names = set(declared_functions()) - {"qk_last_error_message"}
    assert names == set(_abi.PROTOTYPES), names ^ set(_abi.PROTOTYPES)


def test_predicate_struct_layout_matches_c():
    # qk_predicate: 2 ints, 2 pointers, 2 ints, 2 pointers, int64, double
    assert ctypes.sizeof(_abi.Predicate) == 4 + 4 + 8 + 8 + 4 + 4 + 8 + 8 + 8 + 8


def test_datagen_deterministic_and_sliceable():
    a = datagen.uniform_i32(5, 0, 1000, 100_000)
    b = datagen.uniform_i32(5, 0, 1000, 100_000)
    np.testing.assert_array_equal(a, b)
    assert a.min() >= 0 and a.max() < 1000
    z = datagen.zipf_keys(1, 1.1, 1000, 1000)
    z2 = datagen.zipf_keys(1, 1.1, 1000, 500)
    np.testing.assert_array_equal(z[:500], z2)


def test_datagen_distributions():
    n = 2_000_000
    keys, ranks = datagen.zipf_keys(7, 1.1, 1_000_000, n, with_ranks=True)
    h = np.sum(np.arange(1, 1_000_001, dtype=np.float64) ** -1.1)
    top = np.mean(ranks == 1)
    assert abs(top - 1.0 / h) < 0.005, top           # top rank ~ 12.4% of rows
    assert (keys != np.iinfo(np.int64).min).all()
    v = datagen.lognormal_f32(8, 0.0, 1.0, n)
    sel = np.mean((v >= np.float32(0.5)) & (v <= np.float32(4.0)))
    assert abs(sel - 0.6726) < 0.005, sel            # Phi(ln 4) - Phi(ln 0.5)
    nm = datagen.null_mask(9, 0.05, n)
    frac = np.unpackbits(nm.view(np.uint8), bitorder="little")[:n].mean()
    assert abs(frac - 0.05) < 0.002


def test_cpp_dropin_headers_compile(tmp_path):
    """The C++ drop-in layer (include/qikkdb/) compiles on its own and with the driver."""
    import shutil
    import subprocess
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    inc = os.path.join(ROOT, "include")
    tu = tmp_path / "all.cpp"
    headers = sorted(f for f in os.listdir(os.path.join(inc, "qikkdb")) if f.endswith(".h"))
    tu.write_text("".join(f'#include "qikkdb/{h}"\n' for h in headers) +
                  "int main() { GPUGroupBy<AggregationFunctions::sum, int64_t, int64_t, float> *p = nullptr;"
                  " (void)p; return 0; }\n")
    for src in (str(tu), os.path.join(ROOT, "tests", "cpp", "groupby_driver.cpp")):
        r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", inc, src],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_struct_layouts_match_c(tmp_path):
    """The ctypes mirrors of the header's structs have the C sizes and field offsets."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {
        "qk_batch_columns": _abi.BatchColumns, "qk_batch_query": _abi.BatchQuery,
        "qk_batch_tuning": _abi.BatchTuning, "qk_serve_config": _abi.ServeConfig,
        "qk_serve_stats": _abi.ServeStats, "qk_predicate": _abi.Predicate,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "qikkdb_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"
