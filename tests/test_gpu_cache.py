"""GPU checks of the block cache (qk_cache_*) and the pool cap (qk_mem_set_limit): blocks keep
their contents while cached, least recently used blocks go first, the lock list holds, the
pool's cap drives eviction, and the reference's errors come back for a block larger than the
cache or a cache with nothing left to evict."""
import numpy as np
import pytest

import qikkdb_amd as qk
from qikkdb_amd import BlockCache, DeviceBuffer

pytestmark = pytest.mark.gpu
MB = 1 << 20


def fill(buf, value, nbytes):
    arr = np.full(nbytes // 8, value, np.int64)
    qk._abi.call("qk_memcpy_htod", buf.ptr, arr.ctypes.data, arr.nbytes, None)


def read(buf, nbytes):
    return buf.to_numpy(np.int64, nbytes // 8)


def test_hits_keep_contents_and_lru_eviction():
    qk.ensure_device(0)
    c = BlockCache(3 * MB)
    try:
        b0, n0, hit = c.get_column("db", "t.a", 0, MB)
        assert not hit and n0 == MB
        fill(b0, 7, MB)
        b1, _, hit = c.get_column("db", "t.a", 0, MB)
        assert hit and b1.ptr == b0.ptr and (read(b1, MB) == 7).all()
        c.get_column("db", "t.b", 0, MB)
        c.get_column("db", "t.c", 0, MB)
        c.get_column("db", "t.a", 0, MB)      # a most recent, b least
        c.get_column("db", "t.d", 0, MB)      # evicts b
        assert not c.contains("db", "t.b", 0) and c.contains("db", "t.a", 0)
        assert (read(c.get_column("db", "t.a", 0, MB)[0], MB) == 7).all()
        st = c.stats()
        assert st["entries"] == 3 and st["used"] == 3 * MB and st["evictions"] == 1
        c.clear_block("db", "t.a", 0)
        assert not c.contains("db", "t.a", 0) and c.stats()["entries"] == 2
    finally:
        c.close()


def test_lock_list_and_errors():
    qk.ensure_device(0)
    c = BlockCache(2 * MB)
    try:
        BlockCache.set_lock_list(["db.t.a", "db.t.b"])
        c.set_current_block(4)
        c.get_column("db", "t.a", 4, MB)
        c.get_column("db", "t.b", 4, MB)
        with pytest.raises(qk.QueryEngineError) as e:
            c.get_column("db", "t.c", 4, MB)
        assert e.value.code == qk._abi.QK_OUT_OF_MEMORY
        assert not c.evict()
        c.set_current_block(5)
        assert c.evict() and c.stats()["entries"] == 1
        with pytest.raises(qk.QueryEngineError) as e:
            c.get_column("db", "t.big", 5, 2 * MB + 1)
        assert e.value.code == qk._abi.QK_ERROR
    finally:
        BlockCache.set_lock_list([])
        c.close()


def test_pool_cap_drives_eviction_and_refuses_allocations():
    qk.ensure_device(0)
    qk._abi.call("qk_device_synchronize")
    qk._abi.call("qk_mem_release_cached")
    in_use, _ = qk.memory_stats()
    c = BlockCache(64 * MB)
    try:
        qk.set_memory_limit(in_use + 4 * MB)
        assert qk.memory_limit() == in_use + 4 * MB
        for i in range(6):  # each needs room under the cap: older blocks are evicted
            b, _, hit = c.get_column("db", "t.v", i, MB)
            fill(b, i, MB)
        st = c.stats()
        assert st["entries"] <= 4 and st["evictions"] >= 2
        assert c.contains("db", "t.v", 5) and not c.contains("db", "t.v", 0)
        assert (read(c.get_column("db", "t.v", 5, MB)[0], MB) == 5).all()
        assert not qk.can_allocate(8 * MB)
        with pytest.raises(qk.QueryEngineError) as e:
            DeviceBuffer(8 * MB)
        assert e.value.code == qk._abi.QK_OUT_OF_MEMORY
    finally:
        qk.set_memory_limit(0)
        c.close()
    assert qk.memory_limit() == 0 and qk.can_allocate(8 * MB)
    with pytest.raises(qk.QueryEngineError):
        qk.set_memory_limit_percent(0)
