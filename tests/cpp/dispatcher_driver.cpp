// dispatcher_driver.cpp - a C++ caller of the drop-in headers shaped like the reference's
// dispatcher: GroupByHelper<OP,O,K,V> with CreateInstance / ProcessBlock / GetResults over
// `std::vector<std::unique_ptr<IGroupBy>> groupByTables_` (GpuSqlDispatcherAggregationFunctions.h:
// 99-159 single key, :161-365 composite key), the per-block WHERE -> reconstructColKeep ->
// ProcessBlock sequence (:377-507), and GetResults(..., groupByTables_, ...) after the last block.
// Blocks are dealt round-robin to two dispatcher "threads", each with its own table, so the
// tables-vector GetResults really merges.  It also drives the GPUReconstruct statics
// (reconstructCol, GenerateIndexes, PrefixSum, Sum/SumKeep, CompressNullMask), the pre-seeded
// GPUGroupBy constructor and ReconstructRawNumbers, each checked against a host computation.
//
//   dispatcher_driver <dir> <rows> <block_rows> <threshold>
//
// Reads <dir>/a.i32, b.i32, v.i32, f.i32 (int32 columns) and an.u8, bn.u8, vn.u8 (one NULL byte
// per row for a, b, v).  WHERE f > threshold.  Prints one line per group:
//   S <agg> <key null> <key> <value null> <value>                      (GROUP BY a)
//   M <agg> <a null> <a> <b null> <b> <value null> <value>             (GROUP BY a, b)
// with agg in {min, max, sum, avg, count}, then "CHECK <name> ok" lines; any failed check exits 1.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "qikkdb/GPUBinary.h"
#include "qikkdb/GPUGroupBy.h"
#include "qikkdb/GPUMemoryCache.h"
#include "qikkdb/GPUReconstruct.h"

namespace {

template <typename T>
std::vector<T> read_col(const std::string& path, size_t n) {
    std::vector<T> v(n);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(T), n, f) != n) {
        std::fprintf(stderr, "cannot read %zu values from %s\n", n, path.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

void fail_check(const char* what) {
    std::fprintf(stderr, "CHECK FAILED: %s\n", what);
    std::exit(1);
}

std::vector<nullmask_t> pack(const uint8_t* bytes, int32_t n) {
    std::vector<nullmask_t> w(NullValues::GetNullBitMaskSize(n), 0);
    for (int32_t i = 0; i < n; i++)
        if (bytes[i]) w[i >> 6] |= 1ull << (i & 63);
    return w;
}

bool bit(const std::vector<nullmask_t>& w, int32_t i) { return (w[i >> 6] >> (i & 63)) & 1ull; }

// A device column of one block and its null bit mask.
struct Block {
    int32_t* data = nullptr;
    nullmask_t* nulls = nullptr;
    Block(const int32_t* host, const uint8_t* nbytes, int32_t m) {
        GPUMemory::alloc(&data, m);
        GPUMemory::copyHostToDevice(data, host, m);
        if (nbytes) {
            std::vector<nullmask_t> w = pack(nbytes, m);
            GPUMemory::alloc(&nulls, w.size());
            GPUMemory::copyHostToDevice(nulls, w.data(), w.size());
        }
    }
    ~Block() {
        GPUMemory::free(data);
        GPUMemory::free(nulls);
    }
};

// A block of a column loaded through the device's block cache, as GpuSqlDispatcher::LoadCol
// does: GetColumn<T>, and on a miss the block (and its null mask, cached as
// "<column>_nullMask") is copied from the host.  The cache owns the memory.
struct CachedBlock {
    int32_t* data = nullptr;
    nullmask_t* nulls = nullptr;
    CachedBlock(GPUMemoryCache& cache, const std::string& column, const int32_t* host, const uint8_t* nbytes,
                int32_t blockIndex, int32_t m, int64_t& hits, int64_t& misses) {
        auto d = cache.GetColumn<int32_t>("db", column, blockIndex, m, -1, 0);
        data = std::get<0>(d);
        if (std::get<2>(d)) hits++; else misses++;
        if (!std::get<2>(d)) GPUMemory::copyHostToDevice(data, host, m);
        if (nbytes) {
            std::vector<nullmask_t> w = pack(nbytes, m);
            auto nm = cache.GetColumn<nullmask_t>("db", column + "_nullMask", blockIndex, w.size(), -1, 0);
            nulls = std::get<0>(nm);
            if (std::get<2>(nm)) hits++; else misses++;
            if (!std::get<2>(nm)) GPUMemory::copyHostToDevice(nulls, w.data(), w.size());
        }
    }
};

struct CacheCounts {
    int64_t hits = 0, misses = 0;
};

// The part of the reference dispatcher that the GROUP BY helpers touch.
struct Dispatcher {
    std::vector<std::unique_ptr<IGroupBy>> groupByTables_;
    int32_t dispatcherThreadId_ = 0;
};

// GROUP BY one key column (the shape of GpuSqlDispatcher::GroupByHelper<OP,O,K,V>).
template <typename OP, typename O, typename K, typename V>
struct GroupByHelper {
    static std::unique_ptr<IGroupBy> CreateInstance(int32_t groupByBuckets, int32_t hashTableMultiplier) {
        return std::make_unique<GPUGroupBy<OP, O, K, V>>(groupByBuckets, hashTableMultiplier);
    }
    static void ProcessBlock(K* keys, nullmask_t* keyNulls, V* values, nullmask_t* valueNulls, int32_t n,
                             Dispatcher& d) {
        reinterpret_cast<GPUGroupBy<OP, O, K, V>*>(d.groupByTables_[d.dispatcherThreadId_].get())
            ->ProcessBlock(keys, values, n, keyNulls, valueNulls);
    }
    static void GetResults(const char* agg, Dispatcher& d) {
        int32_t outSize = 0;
        K* outKeys = nullptr;
        nullmask_t* outKeyNullMask = nullptr;
        O* outValues = nullptr;
        nullmask_t* outValueNullMask = nullptr;
        reinterpret_cast<GPUGroupBy<OP, O, K, V>*>(d.groupByTables_[d.dispatcherThreadId_].get())
            ->GetResults(&outKeys, &outValues, &outSize, d.groupByTables_, &outKeyNullMask, &outValueNullMask);
        std::vector<K> k(outSize);
        std::vector<O> v(outSize);
        std::vector<nullmask_t> kn(NullValues::GetNullBitMaskSize(outSize)), vn(kn.size());
        if (outSize > 0) {
            GPUMemory::copyDeviceToHost(k.data(), outKeys, outSize);
            GPUMemory::copyDeviceToHost(v.data(), outValues, outSize);
            GPUMemory::copyDeviceToHost(kn.data(), outKeyNullMask, kn.size());
            GPUMemory::copyDeviceToHost(vn.data(), outValueNullMask, vn.size());
        }
        for (int32_t i = 0; i < outSize; i++)
            std::printf("S %s %d %lld %d %lld\n", agg, (int)bit(kn, i), (long long)k[i], (int)bit(vn, i),
                        (long long)v[i]);
        GPUMemory::free(outKeys);
        GPUMemory::free(outValues);
        GPUMemory::free(outKeyNullMask);
        GPUMemory::free(outValueNullMask);
    }
};

// GROUP BY several key columns (the shape of GroupByHelper<OP,O,std::vector<void*>,V>).
template <typename OP, typename O, typename V>
struct MultiGroupByHelper {
    static std::unique_ptr<IGroupBy> CreateInstance(int32_t groupByBuckets, int32_t hashTableMultiplier,
                                                    std::vector<DataType>& keyTypes) {
        return std::make_unique<GPUGroupBy<OP, O, std::vector<void*>, V>>(groupByBuckets, hashTableMultiplier,
                                                                          keyTypes);
    }
    static void ProcessBlock(std::vector<void*> keys, std::vector<nullmask_t*> keyNulls, V* values,
                             nullmask_t* valueNulls, int32_t n, Dispatcher& d) {
        reinterpret_cast<GPUGroupBy<OP, O, std::vector<void*>, V>*>(d.groupByTables_[d.dispatcherThreadId_].get())
            ->ProcessBlock(keys, keyNulls, values, n, valueNulls);
    }
    static void GetResults(const char* agg, Dispatcher& d) {
        int32_t outSize = 0;
        std::vector<void*> outKeys;
        std::vector<nullmask_t*> outKeyNullMasks;
        O* outValues = nullptr;
        nullmask_t* outValueNullMask = nullptr;
        reinterpret_cast<GPUGroupBy<OP, O, std::vector<void*>, V>*>(d.groupByTables_[d.dispatcherThreadId_].get())
            ->GetResults(&outKeys, &outValues, &outSize, d.groupByTables_, &outKeyNullMasks, &outValueNullMask);
        const size_t words = NullValues::GetNullBitMaskSize(outSize);
        std::vector<int32_t> a(outSize), b(outSize);
        std::vector<O> v(outSize);
        std::vector<nullmask_t> an(words), bn(words), vn(words);
        if (outSize > 0) {
            GPUMemory::copyDeviceToHost(a.data(), static_cast<int32_t*>(outKeys[0]), outSize);
            GPUMemory::copyDeviceToHost(b.data(), static_cast<int32_t*>(outKeys[1]), outSize);
            GPUMemory::copyDeviceToHost(v.data(), outValues, outSize);
            GPUMemory::copyDeviceToHost(an.data(), outKeyNullMasks[0], words);
            GPUMemory::copyDeviceToHost(bn.data(), outKeyNullMasks[1], words);
            GPUMemory::copyDeviceToHost(vn.data(), outValueNullMask, words);
        }
        for (int32_t i = 0; i < outSize; i++)
            std::printf("M %s %d %d %d %d %d %lld\n", agg, (int)bit(an, i), a[i], (int)bit(bn, i), b[i],
                        (int)bit(vn, i), (long long)v[i]);
        for (void* p : outKeys) GPUMemory::free(p);
        for (nullmask_t* p : outKeyNullMasks) GPUMemory::free(p);
        GPUMemory::free(outValues);
        GPUMemory::free(outValueNullMask);
    }
};

struct Columns {
    int32_t n;
    std::vector<int32_t> a, b, v, f;
    std::vector<uint8_t> an, bn, vn;
};

// The dispatcher's per-block sequence for one aggregate: WHERE mask, compaction of every
// column the GROUP BY reads (with null masks), ProcessBlock on the block's dispatcher thread.
template <typename Helper, bool MULTI>
void run_query(const char* agg, const Columns& c, int32_t block, int32_t threshold, GPUMemoryCache& cache,
               CacheCounts& cc) {
    Dispatcher d;
    std::vector<DataType> kt = {COLUMN_INT, COLUMN_INT};
    for (int t = 0; t < 2; t++) {
        if constexpr (MULTI)
            d.groupByTables_.push_back(Helper::CreateInstance(262144, 1, kt));
        else
            d.groupByTables_.push_back(Helper::CreateInstance(262144, 1));
    }
    int32_t blockIndex = 0;
    for (int32_t s = 0; s < c.n; s += block, blockIndex++) {
        const int32_t m = std::min(block, c.n - s);
        d.dispatcherThreadId_ = blockIndex % 2;
        cache.SetCurrentBlockIndex(blockIndex);
        CachedBlock a(cache, "t.a", c.a.data() + s, c.an.data() + s, blockIndex, m, cc.hits, cc.misses);
        CachedBlock b(cache, "t.b", c.b.data() + s, c.bn.data() + s, blockIndex, m, cc.hits, cc.misses);
        CachedBlock v(cache, "t.v", c.v.data() + s, c.vn.data() + s, blockIndex, m, cc.hits, cc.misses);
        CachedBlock f(cache, "t.f", c.f.data() + s, nullptr, blockIndex, m, cc.hits, cc.misses);
        int8_t* mask = nullptr;
        GPUMemory::alloc(&mask, m);
        GPUBinary<FilterConditions::greater, int8_t, int32_t*, int32_t>::Binary(mask, f.data, threshold, m);
        int32_t *ca = nullptr, *cb = nullptr, *cv = nullptr;
        nullmask_t *can = nullptr, *cbn = nullptr, *cvn = nullptr;
        int32_t ka = 0, kb = 0, kv = 0;
        GPUReconstruct::reconstructColKeep(&ca, &ka, a.data, mask, m, &can, a.nulls);
        GPUReconstruct::reconstructColKeep(&cb, &kb, b.data, mask, m, &cbn, b.nulls);
        GPUReconstruct::reconstructColKeep(&cv, &kv, v.data, mask, m, &cvn, v.nulls);
        if (ka != kv || kb != kv) fail_check("compaction counts differ");
        if constexpr (MULTI)
            Helper::ProcessBlock({ca, cb}, {can, cbn}, cv, cvn, kv, d);
        else
            Helper::ProcessBlock(ca, can, cv, cvn, kv, d);
        for (void* p : {(void*)ca, (void*)cb, (void*)cv, (void*)can, (void*)cbn, (void*)cvn, (void*)mask})
            GPUMemory::free(p);
    }
    d.dispatcherThreadId_ = 0;
    Helper::GetResults(agg, d);
}

// reconstructCol, GenerateIndexes, PrefixSum, PrefixSumExclusive, Sum, SumKeep and
// CompressNullMask on one block, each against the host computation.
void check_reconstruct(const Columns& c, int32_t threshold) {
    const int32_t m = c.n;
    Block v(c.v.data(), c.vn.data(), m), f(c.f.data(), nullptr, m);
    int8_t* mask = nullptr;
    GPUMemory::alloc(&mask, m);
    GPUBinary<FilterConditions::greater, int8_t, int32_t*, int32_t>::Binary(mask, f.data, threshold, m);
    std::vector<int32_t> want;
    std::vector<uint8_t> wantNull;
    std::vector<int32_t> wantIdx;
    for (int32_t i = 0; i < m; i++)
        if (c.f[i] > threshold) {
            want.push_back(c.v[i]);
            wantNull.push_back(c.vn[i]);
            wantIdx.push_back(i);
        }
    const int32_t k = (int32_t)want.size();
    // reconstructCol into host buffers, with the null mask
    std::vector<int32_t> got(m);
    std::vector<nullmask_t> gotNull(NullValues::GetNullBitMaskSize(m));
    int32_t cnt = -1;
    GPUReconstruct::reconstructCol(got.data(), &cnt, v.data, mask, m, gotNull.data(), v.nulls);
    if (cnt != k) fail_check("reconstructCol count");
    for (int32_t i = 0; i < k; i++)
        if (got[i] != want[i] || bit(gotNull, i) != (wantNull[i] != 0)) fail_check("reconstructCol data / nulls");
    std::printf("CHECK reconstructCol ok\n");
    // GenerateIndexes into a host buffer
    std::vector<int32_t> idx(m);
    cnt = -1;
    GPUReconstruct::GenerateIndexes(idx.data(), &cnt, mask, m);
    if (cnt != k || !std::equal(wantIdx.begin(), wantIdx.end(), idx.begin())) fail_check("GenerateIndexes");
    std::printf("CHECK GenerateIndexes ok\n");
    // inclusive and exclusive prefix sums of the mask
    int32_t* ps = nullptr;
    GPUMemory::alloc(&ps, m);
    std::vector<int32_t> hps(m);
    GPUReconstruct::PrefixSum(ps, mask, m);
    GPUMemory::copyDeviceToHost(hps.data(), ps, m);
    int32_t run = 0;
    for (int32_t i = 0; i < m; i++) {
        run += c.f[i] > threshold ? 1 : 0;
        if (hps[i] != run) fail_check("PrefixSum");
    }
    GPUReconstruct::PrefixSumExclusive(ps, mask, m);
    GPUMemory::copyDeviceToHost(hps.data(), ps, m);
    run = 0;
    for (int32_t i = 0; i < m; i++) {
        if (hps[i] != run) fail_check("PrefixSumExclusive");
        run += c.f[i] > threshold ? 1 : 0;
    }
    GPUMemory::free(ps);
    std::printf("CHECK PrefixSum ok\n");
    // Sum into a host value and SumKeep into a device scalar
    int64_t hsum = -1;
    GPUReconstruct::Sum(hsum, mask, m);
    int64_t* dsum = nullptr;
    GPUMemory::alloc(&dsum, 1);
    GPUReconstruct::SumKeep(dsum, mask, m);
    int64_t ksum = -1;
    GPUMemory::copyDeviceToHost(&ksum, dsum, 1);
    GPUMemory::free(dsum);
    if (hsum != k || ksum != k) fail_check("Sum / SumKeep");
    std::printf("CHECK Sum ok\n");
    // CompressNullMask of one byte per row
    uint8_t* bytes = nullptr;
    GPUMemory::alloc(&bytes, m);
    GPUMemory::copyHostToDevice(bytes, c.vn.data(), m);
    cuda_ptr<nullmask_t> packed = GPUReconstruct::CompressNullMask(bytes, m);
    std::vector<nullmask_t> hp(NullValues::GetNullBitMaskSize(m));
    GPUMemory::copyDeviceToHost(hp.data(), packed.get(), hp.size());
    if (hp != pack(c.vn.data(), m)) fail_check("CompressNullMask");
    GPUMemory::free(bytes);
    GPUMemory::free(mask);
    std::printf("CHECK CompressNullMask ok\n");
}

// The pre-seeded constructor the way the reference's own caller uses it (GPUGroupBy.cuh:
// 718-733): the keys of one table's results seed a second table, whose results then follow
// their order.  And ReconstructRawNumbers of an AVG table against its own GetResults.
void check_seeded_and_raw(const Columns& c) {
    const int32_t m = c.n;
    Block a(c.a.data(), c.an.data(), m), v(c.v.data(), c.vn.data(), m);
    GPUGroupBy<AggregationFunctions::sum, int32_t, int32_t, int32_t> sumTable(262144, 1);
    sumTable.ProcessBlock(a.data, v.data, m, a.nulls, v.nulls);
    int32_t* keys = nullptr;
    int32_t* sums = nullptr;
    nullmask_t *keyNulls = nullptr, *sumNulls = nullptr;
    int32_t g = 0;
    sumTable.GetResults(&keys, &sums, &g, &keyNulls, &sumNulls);
    nullmask_t firstChar = 0;
    GPUMemory::copyDeviceToHost(&firstChar, keyNulls, 1);
    GPUGroupBy<AggregationFunctions::count, int64_t, int32_t, int32_t> countTable(g, keys, (firstChar & 1) == 1);
    countTable.ProcessBlock(a.data, v.data, m, a.nulls, v.nulls);
    int32_t* keys2 = nullptr;
    int64_t* counts = nullptr;
    nullmask_t *keyNulls2 = nullptr, *countNulls = nullptr;
    int32_t g2 = 0;
    countTable.GetResults(&keys2, &counts, &g2, &keyNulls2, &countNulls);
    if (g2 != g) fail_check("seeded table group count");
    std::vector<int32_t> hk(g), hk2(g);
    std::vector<nullmask_t> hn(NullValues::GetNullBitMaskSize(g)), hn2(hn.size());
    GPUMemory::copyDeviceToHost(hk.data(), keys, g);
    GPUMemory::copyDeviceToHost(hk2.data(), keys2, g);
    GPUMemory::copyDeviceToHost(hn.data(), keyNulls, hn.size());
    GPUMemory::copyDeviceToHost(hn2.data(), keyNulls2, hn.size());
    for (int32_t i = 0; i < g; i++)
        if (bit(hn, i) != bit(hn2, i) || (!bit(hn, i) && hk[i] != hk2[i])) fail_check("seeded key order");
    // the counts at each position are those of the key at that position
    std::vector<int64_t> hc(g);
    GPUMemory::copyDeviceToHost(hc.data(), counts, g);
    for (int32_t i = 0; i < g; i++) {
        int64_t want = 0;
        for (int32_t r = 0; r < m; r++) {
            const bool rn = c.an[r] != 0;
            if (bit(hn, i) ? rn : (!rn && c.a[r] == hk[i])) want += c.vn[r] ? 0 : 1;
        }
        if (hc[i] != want) fail_check("seeded table counts");
    }
    for (void* p : {(void*)keys, (void*)sums, (void*)keyNulls, (void*)sumNulls, (void*)keys2, (void*)counts,
                    (void*)keyNulls2, (void*)countNulls})
        GPUMemory::free(p);
    std::printf("CHECK seeded ok\n");

    // raw numbers of an AVG table: sums and occurrences whose quotient is the table's AVG
    GPUGroupBy<AggregationFunctions::avg, int32_t, int32_t, int32_t> avgTable(262144, 1);
    avgTable.ProcessBlock(a.data, v.data, m, a.nulls, v.nulls);
    const int32_t cap = avgTable.GetMaxHashCount();
    std::vector<int32_t> rk(cap), rv(cap);
    std::vector<nullarray_t> rkn(cap), rvn(cap);
    std::vector<int64_t> rocc(cap);
    int32_t rg = 0;
    avgTable.ReconstructRawNumbers(rk.data(), rkn.data(), rv.data(), rvn.data(), rocc.data(), &rg);
    int32_t* ak = nullptr;
    int32_t* av = nullptr;
    nullmask_t *akn = nullptr, *avn = nullptr;
    int32_t ag = 0;
    avgTable.GetResults(&ak, &av, &ag, &akn, &avn);
    if (ag != rg) fail_check("ReconstructRawNumbers count");
    std::vector<int32_t> hak(ag), hav(ag);
    std::vector<nullmask_t> hakn(NullValues::GetNullBitMaskSize(ag)), havn(hakn.size());
    GPUMemory::copyDeviceToHost(hak.data(), ak, ag);
    GPUMemory::copyDeviceToHost(hav.data(), av, ag);
    GPUMemory::copyDeviceToHost(hakn.data(), akn, hakn.size());
    GPUMemory::copyDeviceToHost(havn.data(), avn, havn.size());
    for (int32_t i = 0; i < ag; i++) {
        if ((rkn[i] != 0) != bit(hakn, i) || (!rkn[i] && rk[i] != hak[i])) fail_check("raw keys");
        if ((rvn[i] != 0) != bit(havn, i)) fail_check("raw value nulls");
        if (!rvn[i] && (rocc[i] <= 0 || (int32_t)(rv[i] / rocc[i]) != hav[i])) fail_check("raw sum / occurrences");
    }
    for (void* p : {(void*)ak, (void*)av, (void*)akn, (void*)avn}) GPUMemory::free(p);
    std::printf("CHECK ReconstructRawNumbers ok\n");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s dir rows block_rows threshold\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    Columns c;
    c.n = std::atoi(argv[2]);
    const int32_t block = std::atoi(argv[3]), threshold = std::atoi(argv[4]);
    c.a = read_col<int32_t>(dir + "/a.i32", c.n);
    c.b = read_col<int32_t>(dir + "/b.i32", c.n);
    c.v = read_col<int32_t>(dir + "/v.i32", c.n);
    c.f = read_col<int32_t>(dir + "/f.i32", c.n);
    c.an = read_col<uint8_t>(dir + "/an.u8", c.n);
    c.bn = read_col<uint8_t>(dir + "/bn.u8", c.n);
    c.vn = read_col<uint8_t>(dir + "/vn.u8", c.n);
    try {
        qikkdb::check(qk_init(0));
        using namespace AggregationFunctions;
        {
            // every block of the four columns and three null masks fits: the first query loads
            // them, the other nine find them cached
            const int32_t nblocks = (c.n + block - 1) / block;
            GPUMemoryCache cache(0, (size_t)c.n * 4 * 4 + (size_t)nblocks * 3 * 8 * (block / 64 + 1));
            CacheCounts cc;
            run_query<GroupByHelper<min, int32_t, int32_t, int32_t>, false>("min", c, block, threshold, cache, cc);
            run_query<GroupByHelper<max, int32_t, int32_t, int32_t>, false>("max", c, block, threshold, cache, cc);
            run_query<GroupByHelper<sum, int32_t, int32_t, int32_t>, false>("sum", c, block, threshold, cache, cc);
            run_query<GroupByHelper<avg, int32_t, int32_t, int32_t>, false>("avg", c, block, threshold, cache, cc);
            run_query<GroupByHelper<count, int64_t, int32_t, int32_t>, false>("count", c, block, threshold, cache, cc);
            run_query<MultiGroupByHelper<min, int32_t, int32_t>, true>("min", c, block, threshold, cache, cc);
            run_query<MultiGroupByHelper<max, int32_t, int32_t>, true>("max", c, block, threshold, cache, cc);
            run_query<MultiGroupByHelper<sum, int32_t, int32_t>, true>("sum", c, block, threshold, cache, cc);
            run_query<MultiGroupByHelper<avg, int32_t, int32_t>, true>("avg", c, block, threshold, cache, cc);
            run_query<MultiGroupByHelper<count, int64_t, int32_t>, true>("count", c, block, threshold, cache, cc);
            if (cc.misses != (int64_t)nblocks * 7 || cc.hits != 9 * cc.misses) fail_check("block cache hits");
            std::printf("CHECK block cache ok\n");
        }
        check_reconstruct(c, threshold);
        check_seeded_and_raw(c);
        // an empty tables vector is refused as in the reference
        std::vector<std::unique_ptr<IGroupBy>> none;
        GPUGroupBy<sum, int32_t, int32_t, int32_t> t(16, 1);
        int32_t *ok = nullptr, *ov = nullptr, og = 0;
        bool threw = false;
        try {
            t.GetResults(&ok, &ov, &og, none);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        if (!threw) fail_check("empty tables vector");
        std::printf("CHECK empty tables ok\n");
    } catch (const query_engine_error& e) {
        std::fprintf(stderr, "query_engine_error %d: %s\n", (int)e.GetQueryEngineError(), e.what());
        return 3;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 4;
    }
    return 0;
}
