"""Build the native libraries of qikkdb_amd in-tree.

* ``libqikkdb_amd.so`` - the HIP kernels and the C ABI declared in include/qikkdb_amd.h,
  compiled for gfx950 only.
* ``libqk_datagen.so`` - host-side seeded column generators used by the benchmark and tests.

Run ``python -m qikkdb_amd.build`` (or call ``build()``).  Objects go to qikkdb_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
HIP_SOURCES = ["runtime.hip", "filter.hip", "reconstruct.hip", "groupby.hip", "twophase.hip", "batch.hip", "serve.hip",
               "cache.hip", "strings.hip"]
LIB = os.path.join(PKG, "libqikkdb_amd.so")
DATAGEN_LIB = os.path.join(PKG, "libqk_datagen.so")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{msg}")


def _newer(target: str, deps: list[str]) -> bool:
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(d) <= t for d in deps)


def build_hip(force: bool = True, jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    headers.append(os.path.join(ROOT, "include", "qikkdb_amd.h"))
    objs, todo = [], []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(o)
        if force or not _newer(o, [s] + headers):
            todo.append([hipcc, *HIPCC_FLAGS, "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for fut in [ex.submit(_run, c) for c in todo]:
            fut.result()
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_datagen(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "datagen.cpp")
    tmp = DATAGEN_LIB + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", tmp, src])
    os.replace(tmp, DATAGEN_LIB)
    if verbose:
        print(f"built {DATAGEN_LIB}")
    return DATAGEN_LIB


DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "groupby_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "cpp", "groupby_driver")
DISPATCHER_SRC = os.path.join(ROOT, "tests", "cpp", "dispatcher_driver.cpp")
DISPATCHER = os.path.join(ROOT, "tests", "cpp", "dispatcher_driver")
STRING_DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "string_driver.cpp")
STRING_DRIVER = os.path.join(ROOT, "tests", "cpp", "string_driver")


def build_driver(verbose: bool = False) -> str | None:
    """The C++ callers of the drop-in headers (include/qikkdb/) used by the parity tests."""
    out = None
    for src, exe in ((DRIVER_SRC, DRIVER), (DISPATCHER_SRC, DISPATCHER), (STRING_DRIVER_SRC, STRING_DRIVER)):
        if not os.path.exists(src):
            continue
        tmp = exe + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src,
              "-L", PKG, "-lqikkdb_amd", "-Wl,-rpath,$ORIGIN/../../qikkdb_amd", "-o", tmp])
        os.replace(tmp, exe)
        if verbose:
            print(f"built {exe}")
        out = exe
    return out


def build(force: bool = True, jobs: int = 4, verbose: bool = False) -> None:
    build_datagen(verbose)
    build_hip(force=force, jobs=jobs, verbose=verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build(force="--incremental" not in sys.argv, verbose=True)
