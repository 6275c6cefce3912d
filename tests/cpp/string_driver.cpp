// string_driver.cpp - a C++ caller of the drop-in string GROUP BY (GPUGroupBy<AGG,O,std::string,V>)
// shaped like the reference's TestGroupByString (qikkDB_test/GPUGroupByTests.cu:15-66): keys of
// each block become a GPUString, ProcessBlock per block, GetResults, then the string column and
// the values are read back.  Runs the six known answers of GPUGroupByTests.cu:474-516 and prints
// "CHECK <name> ok" per case; a wrong group exits 1.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "qikkdb/GPUGroupBy.h"

namespace {

GPUMemory::GPUString prepare(const std::vector<std::string>& keys) {
    std::string all;
    std::vector<int64_t> idx;
    for (const std::string& k : keys) {
        all += k;
        idx.push_back((int64_t)all.size());
    }
    GPUMemory::GPUString s;
    GPUMemory::alloc(&s.allChars, all.size() ? all.size() : 1);
    if (!all.empty()) GPUMemory::copyHostToDevice(s.allChars, all.data(), all.size());
    GPUMemory::alloc(&s.stringIndices, idx.size());
    GPUMemory::copyHostToDevice(s.stringIndices, idx.data(), idx.size());
    return s;
}

template <typename AGG>
void run(const char* name, const std::vector<std::vector<std::string>>& keys,
         const std::vector<std::vector<int32_t>>& values, const std::map<std::string, int32_t>& want) {
    GPUGroupBy<AGG, int32_t, std::string, int32_t> groupBy(8, 1);
    for (size_t b = 0; b < keys.size(); b++) {
        GPUMemory::GPUString k = prepare(keys[b]);
        int32_t* v = nullptr;
        GPUMemory::alloc(&v, values[b].size());
        GPUMemory::copyHostToDevice(v, values[b].data(), values[b].size());
        groupBy.ProcessBlock(k, v, (int32_t)keys[b].size());
        GPUMemory::free(k);
        GPUMemory::free(v);
    }
    GPUMemory::GPUString outKeys;
    int32_t* outValues = nullptr;
    int32_t n = 0;
    groupBy.GetResults(&outKeys, &outValues, &n);
    std::vector<int64_t> idx(n);
    std::vector<int32_t> vals(n);
    GPUMemory::copyDeviceToHost(idx.data(), outKeys.stringIndices, n);
    GPUMemory::copyDeviceToHost(vals.data(), outValues, n);
    std::string all(n ? idx[n - 1] : 0, '\0');
    if (!all.empty()) GPUMemory::copyDeviceToHost(&all[0], outKeys.allChars, all.size());
    std::map<std::string, int32_t> got;
    for (int32_t i = 0; i < n; i++) {
        const int64_t s = i ? idx[i - 1] : 0;
        got[all.substr(s, idx[i] - s)] = vals[i];
    }
    GPUMemory::free(outKeys);
    GPUMemory::free(outValues);
    if (got != want) {
        std::fprintf(stderr, "CHECK FAILED: %s\n", name);
        std::exit(1);
    }
    std::printf("CHECK %s ok\n", name);
}

}  // namespace

int main() {
    try {
        qikkdb::check(qk_init(0));
        using namespace AggregationFunctions;
        run<sum>("StringUnique", {{"Apple", "Abcd", "XYZ"}}, {{1, 2, -1}}, {{"Apple", 1}, {"Abcd", 2}, {"XYZ", -1}});
        run<sum>("StringSimple", {{"Apple", "Abcd", "XYZ", "Abcd", "ZYX", "XYZ"}}, {{1, 2, -1, 3, 7, 5}},
                 {{"Apple", 1}, {"Abcd", 5}, {"XYZ", 4}, {"ZYX", 7}});
        run<sum>("StringMultiBlockSimple", {{"Apple", "Abcd"}, {"XYZ", "Abcd"}, {"ZYX", "XYZ"}, {"Apple", "Apple"}},
                 {{1, 1}, {1, 2}, {1, 2}, {2, 4}}, {{"Apple", 7}, {"Abcd", 3}, {"XYZ", 3}, {"ZYX", 1}});
        const std::vector<std::vector<std::string>> mk = {
            {"Apple", "Abcd", "Apple", "XYZ"}, {"Banana", "XYZ", "Abcd", "0"}, {"XYZ", "XYZ"}};
        const std::vector<std::vector<int32_t>> mv = {{1, 2, 3, 4}, {5, 6, 7, 10}, {13, 15}};
        run<sum>("StringMultiBlockMediumSum", mk, mv, {{"Apple", 4}, {"Abcd", 9}, {"Banana", 5}, {"XYZ", 38}, {"0", 10}});
        run<min>("StringMultiBlockMediumMin", mk, mv, {{"Apple", 1}, {"Abcd", 2}, {"Banana", 5}, {"XYZ", 4}, {"0", 10}});
        run<max>("StringMultiBlockMediumMax", mk, mv, {{"Apple", 3}, {"Abcd", 7}, {"Banana", 5}, {"XYZ", 15}, {"0", 10}});
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}
