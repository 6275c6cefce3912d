// groupby_driver.cpp - a C++ caller of the drop-in headers that runs the reference's
// per-block sequence for `SELECT k, AGG(v) FROM t WHERE v > c GROUP BY k`
// (GpuSqlDispatcherBinaryFunctions.h:80 filter, GpuSqlDispatcherAggregationFunctions.h:
// 422-455 compaction of key and value by the mask, GPUGroupBy::ProcessBlock per block,
// GetResults after the last block), with one table per aggregate as the reference keeps.
//
//   groupby_driver <rows> <block_rows> <keys.i32> <values.i32> <threshold>
//
// Reads int32 key and value columns from raw files, prints one line per group:
//   key sum min max count
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "qikkdb/GPUBinary.h"
#include "qikkdb/GPUGroupBy.h"
#include "qikkdb/GPUReconstruct.h"

namespace {

std::vector<int32_t> read_i32(const char* path, size_t n) {
    std::vector<int32_t> v(n);
    FILE* f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(int32_t), n, f) != n) {
        std::fprintf(stderr, "cannot read %zu int32 from %s\n", n, path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <typename K, typename O>
std::map<K, O> collect(K* dkeys, O* dvals, int32_t count) {
    std::vector<K> k(count);
    std::vector<O> v(count);
    if (count > 0) {
        GPUMemory::copyDeviceToHost(k.data(), dkeys, count);
        GPUMemory::copyDeviceToHost(v.data(), dvals, count);
    }
    GPUMemory::free(dkeys);
    GPUMemory::free(dvals);
    std::map<K, O> m;
    for (int32_t i = 0; i < count; i++) m[k[i]] = v[i];
    return m;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s rows block_rows keys.i32 values.i32 threshold\n", argv[0]);
        return 2;
    }
    const int32_t n = std::atoi(argv[1]), block = std::atoi(argv[2]);
    const int32_t threshold = std::atoi(argv[5]);
    const std::vector<int32_t> hk = read_i32(argv[3], n), hv = read_i32(argv[4], n);
    try {
        qikkdb::check(qk_init(0));
        GPUGroupBy<AggregationFunctions::sum, int64_t, int32_t, int32_t> gsum(1000, 1);
        GPUGroupBy<AggregationFunctions::min, int32_t, int32_t, int32_t> gmin(1000, 1);
        GPUGroupBy<AggregationFunctions::max, int32_t, int32_t, int32_t> gmax(1000, 1);
        GPUGroupBy<AggregationFunctions::count, int64_t, int32_t, int32_t> gcnt(1000, 1);
        for (int32_t s = 0; s < n; s += block) {
            const int32_t m = std::min(block, n - s);
            int32_t *dk = nullptr, *dv = nullptr;
            int8_t* mask = nullptr;
            GPUMemory::alloc(&dk, m);
            GPUMemory::alloc(&dv, m);
            GPUMemory::alloc(&mask, m);
            GPUMemory::copyHostToDevice(dk, hk.data() + s, m);
            GPUMemory::copyHostToDevice(dv, hv.data() + s, m);
            GPUBinary<FilterConditions::greater, int8_t, int32_t*, int32_t>::Binary(mask, dv, threshold, m);
            int32_t *ck = nullptr, *cv = nullptr;
            int32_t kept = 0, kept2 = 0;
            GPUReconstruct::reconstructColKeep(&ck, &kept, dk, mask, m);
            GPUReconstruct::reconstructColKeep(&cv, &kept2, dv, mask, m);
            if (kept != kept2) {
                std::fprintf(stderr, "compaction counts differ: %d vs %d\n", kept, kept2);
                return 1;
            }
            gsum.ProcessBlock(ck, cv, kept);
            gmin.ProcessBlock(ck, cv, kept);
            gmax.ProcessBlock(ck, cv, kept);
            gcnt.ProcessBlock(ck, cv, kept);
            GPUMemory::free(ck);
            GPUMemory::free(cv);
            GPUMemory::free(mask);
            GPUMemory::free(dk);
            GPUMemory::free(dv);
        }
        int32_t c1 = 0, c2 = 0, c3 = 0, c4 = 0;
        int32_t *k1, *k2, *k3, *k4, *mn, *mx;
        int64_t *sm, *ct;
        gsum.GetResults(&k1, &sm, &c1);
        gmin.GetResults(&k2, &mn, &c2);
        gmax.GetResults(&k3, &mx, &c3);
        gcnt.GetResults(&k4, &ct, &c4);
        const auto msum = collect(k1, sm, c1);
        const auto mmin = collect(k2, mn, c2);
        const auto mmax = collect(k3, mx, c3);
        const auto mcnt = collect(k4, ct, c4);
        for (const auto& kv : msum)
            std::printf("%d %lld %d %d %lld\n", kv.first, (long long)kv.second, mmin.at(kv.first), mmax.at(kv.first),
                        (long long)mcnt.at(kv.first));
    } catch (const query_engine_error& e) {
        std::fprintf(stderr, "query_engine_error %d: %s\n", (int)e.GetQueryEngineError(), e.what());
        return 3;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 4;
    }
    return 0;
}
