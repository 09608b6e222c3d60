// Drives the reference-compatible C++ API (include/gpumultisplit.hpp) the way
// the reference's drivers do -- main_multisplit.cu (multisplit_context,
// allocate, key-value call with delta_func and h_offsets), main_sort.cu /
// main_sort_pairs.cu (ms_sort_context, 7/8-bit sorts), main_compaction.cu
// (compaction_context, test_bucket<1>) -- and compares every result with the
// C oracle.  Built twice: against libgpumultisplit.so with device memory from
// the HIP runtime (GPU test), and against the CPU emulation of the same
// kernels with host memory (GMS_DRIVER_HOST, CPU test).
//
// usage: api_driver <n>     exit status 0 when every case matches
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "gpumultisplit.hpp"

#ifndef GMS_DRIVER_HOST
#include <hip/hip_runtime_api.h>
#endif

extern "C" {
int ms_oracle_multisplit(uint32_t kind, uint32_t B, uint32_t param, const uint32_t* kin,
                         const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint64_t n,
                         uint32_t* offsets);
int ms_oracle_multisplit_wide(uint32_t kind, uint32_t B, uint64_t param, uint32_t width,
                              uint32_t key_words, uint32_t value_words, const uint32_t* kin,
                              const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint64_t n,
                              uint32_t* offsets);
uint64_t ms_oracle_compaction(uint32_t kind, uint32_t param, const uint32_t* kin,
                              const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint64_t n);
}

static int failures = 0;
#define EXPECT(cond, ...)                 \
    do {                                  \
        if (!(cond)) {                    \
            printf("FAIL: " __VA_ARGS__); \
            printf("\n");                 \
            ++failures;                   \
        }                                 \
    } while (0)

// Device buffer of u32 (host memory in the emulation build).
struct DevBuf {
    uint32_t* p = nullptr;
    size_t n;
    explicit DevBuf(size_t n_) : n(n_ ? n_ : 1) {
#ifdef GMS_DRIVER_HOST
        p = (uint32_t*)malloc(n * 4);
#else
        if (hipMalloc((void**)&p, n * 4) != hipSuccess) p = nullptr;
#endif
        if (!p) {
            printf("allocation of %zu words failed\n", n);
            exit(2);
        }
    }
    ~DevBuf() {
#ifdef GMS_DRIVER_HOST
        free(p);
#else
        (void)hipFree(p);
#endif
    }
    void put(const std::vector<uint32_t>& h) {
#ifdef GMS_DRIVER_HOST
        memcpy(p, h.data(), h.size() * 4);
#else
        (void)hipMemcpy(p, h.data(), h.size() * 4, hipMemcpyHostToDevice);
#endif
    }
    std::vector<uint32_t> get(size_t m) const {
        std::vector<uint32_t> h(m);
#ifdef GMS_DRIVER_HOST
        memcpy(h.data(), p, m * 4);
#else
        (void)hipDeviceSynchronize();
        (void)hipMemcpy(h.data(), p, m * 4, hipMemcpyDeviceToHost);
#endif
        return h;
    }
};

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd() {
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
}

// main_multisplit.cu: keys in [0, B*delta), delta_func, key-value with
// values = original index, h_offsets requested.
static void multisplit_delta(uint32_t n, uint32_t B) {
    const uint32_t delta = 1000;
    std::vector<uint32_t> k(n), v(n);
    for (uint32_t i = 0; i < n; ++i) {
        k[i] = rnd() % (B * delta);
        v[i] = i;
    }
    DevBuf dk(n), dv(n), dko(n), dvo(n);
    dk.put(k);
    dv.put(v);
    multisplit_context ctx(B);
    multisplit_allocate_key_value(n, ctx);
    std::vector<uint32_t> offs(B, 0xFFFFFFFFu);
    multisplit_key_value(dk.p, dv.p, dko.p, dvo.p, n, ctx, gms::delta_func(delta), offs.data());
    std::vector<uint32_t> wk(n), wv(n), wo(B);
    ms_oracle_multisplit(1 /*delta*/, B, delta, k.data(), v.data(), wk.data(), wv.data(), n,
                         wo.data());
    EXPECT(dko.get(n) == wk, "multisplit_key_value delta keys B=%u n=%u", B, n);
    EXPECT(dvo.get(n) == wv, "multisplit_key_value delta values B=%u n=%u", B, n);
    EXPECT(offs == wo, "multisplit_key_value h_offsets B=%u n=%u", B, n);

    // Key-only on the same context at a smaller n (the reference would need a
    // new allocation here; this API derives the grid from the call's n).
    const uint32_t m = n / 3 + 1;
    multisplit_key_only(dk.p, dko.p, m, ctx, gms::delta_func(delta));
    std::vector<uint32_t> wk2(m);
    ms_oracle_multisplit(1, B, delta, k.data(), nullptr, wk2.data(), nullptr, m, wo.data());
    EXPECT(dko.get(m) == wk2, "multisplit_key_only delta B=%u m=%u", B, m);
    multisplit_release_memory(ctx);
}

static void sort_case(uint32_t n, int bits, bool pairs) {
    std::vector<uint32_t> k(n), v(n);
    for (uint32_t i = 0; i < n; ++i) {
        k[i] = rnd() >> 1;  // rand()-like 31-bit keys, as main_sort.cu
        v[i] = i;
    }
    for (uint32_t i = 0; i + 7 < n; i += 7) k[i] = k[0];  // duplicates expose instability
    DevBuf dk(n), dv(n), dko(n), dvo(n);
    dk.put(k);
    dv.put(v);
    ms_sort_context ctx;
    if (bits == 7) {
        if (pairs) {
            multisplit_sort_7bit_allocate_pairs(n, ctx);
            multisplit_sort_7bit_pairs(dk.p, dko.p, dv.p, dvo.p, n, ctx);
        } else {
            multisplit_sort_7bit_allocate(n, ctx);
            multisplit_sort_7bit(dk.p, dko.p, n, ctx);
        }
    } else {
        if (pairs) {
            multisplit_sort_8bit_allocate_pairs(n, ctx);
            multisplit_sort_8bit_pairs(dk.p, dko.p, dv.p, dvo.p, n, ctx);
        } else {
            multisplit_sort_8bit_allocate(n, ctx);
            multisplit_sort_8bit(dk.p, dko.p, n, ctx);
        }
    }
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
    std::vector<uint32_t> wk(n);
    for (uint32_t i = 0; i < n; ++i) wk[i] = k[order[i]];
    EXPECT(dko.get(n) == wk, "multisplit_sort_%dbit%s keys n=%u", bits, pairs ? "_pairs" : "", n);
    if (pairs) EXPECT(dvo.get(n) == order, "multisplit_sort_%dbit_pairs values n=%u", bits, n);
    EXPECT(dk.get(n) == k, "sort wrote its input n=%u", n);
    multisplit_sort_release_memory(ctx);
}

// main_compaction.cu: keys = a permutation, test_bucket<1> (odd keys last).
static void compaction_case(uint32_t n) {
    std::vector<uint32_t> k(n), v(n);
    std::iota(k.begin(), k.end(), 0u);
    for (uint32_t i = n; i > 1; --i) std::swap(k[i - 1], k[rnd() % i]);
    for (uint32_t i = 0; i < n; ++i) v[i] = i;
    DevBuf dk(n), dv(n), dko(n), dvo(n);
    dk.put(k);
    dv.put(v);
    compaction_context ctx;
    compaction_allocate_key_value(n, ctx);
    const unsigned c = compaction_key_value(dk.p, dv.p, dko.p, dvo.p, n, ctx, gms::test_bucket<1, uint32_t>());
    std::vector<uint32_t> wk(n), wv(n);
    const uint64_t wc = ms_oracle_compaction(4 /*lowbits*/, 1, k.data(), v.data(), wk.data(), wv.data(), n);
    EXPECT(c == wc, "compaction_key_value count %u vs %llu", c, (unsigned long long)wc);
    EXPECT(dko.get(n) == wk, "compaction_key_value keys n=%u", n);
    EXPECT(dvo.get(n) == wv, "compaction_key_value values n=%u", n);
    const unsigned c2 = compaction_key_only(dk.p, dko.p, n, ctx, gms::test_bucket<1, uint32_t>());
    EXPECT(c2 == wc, "compaction_key_only count");
    EXPECT(dko.get(n) == wk, "compaction_key_only keys n=%u", n);
    compaction_release_memory(ctx);
}

// multisplit_key_value<uint64_t, uint64_t> with a 64-bit delta, and the
// typed sort of 64-bit keys carrying 32-bit values.
static void wide_cases(uint32_t n) {
    const uint32_t B = 200;
    const uint64_t delta = 0xFFFFFFFFFFFFFFFFull / 150;
    std::vector<uint32_t> k(2 * n), v(2 * n), v32(n);
    for (uint32_t i = 0; i < 2 * n; ++i) k[i] = rnd();
    for (uint32_t i = 0; i < n; ++i) {
        v[2 * i] = i;
        v[2 * i + 1] = ~i;
        v32[i] = i;
    }
    DevBuf dk(2 * n), dv(2 * n), dko(2 * n), dvo(2 * n);
    dk.put(k);
    dv.put(v);
    multisplit_context ctx(B);
    multisplit_allocate_key_value(n, ctx);
    std::vector<uint32_t> offs(B, 0xFFFFFFFFu);
    multisplit_key_value(reinterpret_cast<uint64_t*>(dk.p), reinterpret_cast<uint64_t*>(dv.p),
                         reinterpret_cast<uint64_t*>(dko.p), reinterpret_cast<uint64_t*>(dvo.p), n,
                         ctx, gms::delta_func64(delta), offs.data());
    std::vector<uint32_t> wk(2 * n), wv(2 * n), wo(B);
    ms_oracle_multisplit_wide(1, B, delta, 0, 2, 2, k.data(), v.data(), wk.data(), wv.data(), n,
                              wo.data());
    EXPECT(dko.get(2 * n) == wk, "multisplit_key_value<u64,u64> keys n=%u", n);
    EXPECT(dvo.get(2 * n) == wv, "multisplit_key_value<u64,u64> values n=%u", n);
    EXPECT(offs == wo, "multisplit_key_value<u64,u64> h_offsets n=%u", n);
    multisplit_release_memory(ctx);

    // The same 64-bit identifier through the functor histogram: the counts are
    // the differences of the oracle's bucket starts.
    DevBuf dh(B);
    dh.put(std::vector<uint32_t>(B, 0u));
    multisplit_histogram(reinterpret_cast<const uint64_t*>(dk.p), n, dh.p, B,
                         gms::delta_func64(delta));
    std::vector<uint32_t> wh(B);
    for (uint32_t b = 0; b < B; ++b) wh[b] = (b + 1 < B ? wo[b + 1] : n) - wo[b];
    EXPECT(dh.get(B) == wh, "multisplit_histogram<u64> n=%u", n);

    DevBuf dv32(n), dvo32(n);
    dv32.put(v32);
    ms_sort_context sc;
    multisplit_sort_8bit_allocate_pairs<uint64_t, uint32_t>(n, sc);
    multisplit_sort_8bit_pairs(reinterpret_cast<uint64_t*>(dk.p), reinterpret_cast<uint64_t*>(dko.p),
                               dv32.p, dvo32.p, n, sc);
    std::vector<uint64_t> k64(n);
    for (uint32_t i = 0; i < n; ++i) k64[i] = (uint64_t)k[2 * i] | ((uint64_t)k[2 * i + 1] << 32);
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return k64[a] < k64[b]; });
    std::vector<uint32_t> sk(2 * n);
    for (uint32_t i = 0; i < n; ++i) {
        sk[2 * i] = (uint32_t)k64[order[i]];
        sk[2 * i + 1] = (uint32_t)(k64[order[i]] >> 32);
    }
    EXPECT(dko.get(2 * n) == sk, "multisplit_sort_8bit_pairs<u64,u32> keys n=%u", n);
    EXPECT(dvo32.get(n) == order, "multisplit_sort_8bit_pairs<u64,u32> values n=%u", n);
    multisplit_sort_release_memory(sc);
}

static void error_cases() {
    multisplit_context ctx(256);
    bool threw = false;
    try {
        DevBuf a(16), b(16);
        multisplit_key_only(a.p, b.p, 16, ctx, gms::range_bucket(256));  // not allocated
    } catch (const gms::error& e) {
        threw = e.status == GMS_ERR_NOT_ALLOCATED;
    }
    EXPECT(threw, "call before allocate must throw GMS_ERR_NOT_ALLOCATED");
    threw = false;
    try {
        multisplit_context bad((1u << 20) + 1u);
    } catch (const gms::error& e) {
        threw = e.status == GMS_ERR_UNSUPPORTED_BUCKETS;
    }
    EXPECT(threw, "2^20+1 buckets must throw GMS_ERR_UNSUPPORTED_BUCKETS");
}

int main(int argc, char** argv) {
    const uint32_t n = argc > 1 ? (uint32_t)strtoul(argv[1], nullptr, 0) : 100000u;
    for (uint32_t B : {2u, 32u, 256u, 1000u}) multisplit_delta(n, B);
    for (int bits : {7, 8})
        for (bool pairs : {false, true}) sort_case(n, bits, pairs);
    compaction_case(n);
    wide_cases(n);
    error_cases();
    if (failures) {
        printf("API DRIVER: %d failure(s)\n", failures);
        return 1;
    }
    printf("API DRIVER OK n=%u\n", n);
    return 0;
}
