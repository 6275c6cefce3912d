// datagen.cpp - seeded synthetic column generators for the benchmark workloads.
//
// Every value is a pure function of (seed, row index), so any slice of a column can be
// regenerated without the rest and the result does not depend on the thread count.
// Methods: SplitMix64 finaliser as a counter-based generator (Steele, Lea, Flood 2014),
// Lemire's multiply-shift for bounded integers, Box-Muller (1958) for normal variates,
// inverse-CDF sampling on an FP64 table for Zipf ranks.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rnd(uint64_t seed, uint64_t i) {
    return mix64(seed * 0xD1B54A32D192ED03ull + (i + 1) * 0x9E3779B97F4A7C15ull);
}

// Uniform double in the open interval (0, 1).
inline double unit(uint64_t r) { return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

template <class F>
void parallel_rows(int64_t n, int nthreads, F&& f) {
    if (nthreads < 1) nthreads = 1;
    if (n < (int64_t)1 << 16) nthreads = 1;
    int64_t per = ((n + nthreads - 1) / nthreads + 63) / 64 * 64;
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; t++) {
        int64_t b = std::min<int64_t>((int64_t)t * per, n), e = std::min<int64_t>(b + per, n);
        if (b >= e) break;
        ts.emplace_back([=, &f] { f(b, e); });
    }
    for (auto& t : ts) t.join();
}

}  // namespace

extern "C" {

// out[i] uniform in [lo, hi)
int qkgen_uniform_i32(uint64_t seed, int32_t lo, int32_t hi, int64_t n, int32_t* out,
                      int nthreads) {
    if (hi <= lo) return 1;
    uint64_t range = (uint64_t)((int64_t)hi - lo);
    parallel_rows(n, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++)
            out[i] = (int32_t)(lo + (int64_t)(((unsigned __int128)rnd(seed, i) * range) >> 64));
    });
    return 0;
}

int qkgen_uniform_i64(uint64_t seed, int64_t lo, int64_t hi, int64_t n, int64_t* out,
                      int nthreads) {
    if (hi <= lo) return 1;
    uint64_t range = (uint64_t)hi - (uint64_t)lo;
    parallel_rows(n, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++)
            out[i] = (int64_t)((uint64_t)lo + (uint64_t)(((unsigned __int128)rnd(seed, i) * range) >> 64));
    });
    return 0;
}

// Zipf(s) ranks over 1..ndistinct by inverse CDF (FP64 table, binary search), scattered
// over the int64 domain by the bijective mix64 so that keys are not small dense integers.
// A key that would equal INT64_MIN is remapped (INT64_MIN is never emitted).
int qkgen_zipf_keys(uint64_t seed, double s, int64_t ndistinct, int64_t n, int64_t* out,
                    int64_t* ranks_out, int nthreads) {
    if (ndistinct < 1) return 1;
    std::vector<double> cdf((size_t)ndistinct);
    double acc = 0.0;
    for (int64_t k = 1; k <= ndistinct; k++) {
        acc += std::pow((double)k, -s);
        cdf[(size_t)(k - 1)] = acc;
    }
    for (auto& c : cdf) c /= acc;
    cdf.back() = 1.0;
    parallel_rows(n, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++) {
            double u = unit(rnd(seed, i));
            int64_t r = (int64_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
            uint64_t key = mix64((uint64_t)r ^ 0x5DEECE66Dull);
            if (key == 0x8000000000000000ull) key = mix64((uint64_t)(r + ndistinct) ^ 0x5DEECE66Dull);
            out[i] = (int64_t)key;
            if (ranks_out) ranks_out[i] = r;
        }
    });
    return 0;
}

// lognormal(mu, sigma) as float via Box-Muller on two independent uniforms per row.
int qkgen_lognormal_f32(uint64_t seed, double mu, double sigma, int64_t n, float* out,
                        int nthreads) {
    const double two_pi = 6.283185307179586476925286766559;
    parallel_rows(n, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++) {
            double u1 = unit(rnd(seed, 2 * (uint64_t)i));
            double u2 = unit(rnd(seed, 2 * (uint64_t)i + 1));
            double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
            out[i] = (float)std::exp(mu + sigma * z);
        }
    });
    return 0;
}

// Bernoulli(p) null bits, LSB-first in uint64 words (ceil(n/64) words written).
int qkgen_null_mask(uint64_t seed, double p, int64_t n, uint64_t* out, int nthreads) {
    int64_t words = (n + 63) / 64;
    parallel_rows(words, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t w = b; w < e; w++) {
            uint64_t bits = 0;
            for (int j = 0; j < 64; j++) {
                int64_t i = w * 64 + j;
                if (i < n && unit(rnd(seed, (uint64_t)i)) < p) bits |= 1ull << j;
            }
            out[w] = bits;
        }
    });
    return 0;
}

// out[i] = (a[i] << shift) | b[i] as int64 (composite key packing helper for tests).
int qkgen_pack_i32_pair(const int32_t* a, const int32_t* b, int64_t n, int64_t* out,
                        int nthreads) {
    parallel_rows(n, nthreads, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++)
            out[i] = (int64_t)(((uint64_t)(uint32_t)a[i] << 32) | (uint32_t)b[i]);
    });
    return 0;
}

}  // extern "C"
