"""Host checks of the block cache (qikkdb_amd/csrc/cache.hpp): its bookkeeping runs over a host
allocator in tests/cpp/cache_test.cpp (hits, misses, LRU eviction, the lock list, one-block
removal, the allocator cap), and the C++ drop-in header GPUMemoryCache.h compiles."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_cache_bookkeeping_on_the_host(tmp_path):
    exe = tmp_path / "cache_test"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpp", "cache_test.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and "cache ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_drop_in_cache_header_compiles(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "qikkdb/GPUMemoryCache.h"\n'
                   "int f(GPUMemoryCache& c) {\n"
                   '    auto r = c.GetColumn<int64_t>("db", "t.c", 0, 10, -1, 0);\n'
                   "    GPUMemoryCache::SetLockList({\"db.t.c\"});\n"
                   "    c.SetCurrentBlockIndex(0);\n"
                   '    return std::get<2>(r) && c.ContainsColumn("db", "t.c", 0, -1, 0) && c.Evict();\n'
                   "}\n")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)
