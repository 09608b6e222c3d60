// Shared by the functor drivers (functor_*.hip): the caller-compiled
// multisplit path of gpumultisplit_device.hpp driven with functors that none
// of the built-in identifiers expresses, each result checked against a
// stable sort of the input by host-computed bucket ids (std::stable_sort of
// the positions, the textbook definition of a stable multisplit), with the
// same clamp of ids >= B into bucket B-1 that the library applies.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "gpumultisplit_device.hpp"

namespace fdrv {

inline uint64_t& rng_state() {
    static uint64_t s = 0x2545F4914F6CDD1Dull;
    return s;
}
inline uint64_t rnd64() {  // splitmix64
    uint64_t z = (rng_state() += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <class T>
struct Dev {
    T* p = nullptr;
    size_t n;
    explicit Dev(size_t n_) : n(n_ ? n_ : 1) {
        if (hipMalloc((void**)&p, n * sizeof(T)) != hipSuccess) {
            printf("device allocation of %zu bytes failed\n", n * sizeof(T));
            exit(2);
        }
    }
    ~Dev() { (void)hipFree(p); }
    void put(const std::vector<T>& h) {
        if (!h.empty() && hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            exit(3);
    }
    std::vector<T> get(size_t m) const {
        std::vector<T> h(m);
        if (m && hipMemcpy(h.data(), p, m * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) exit(3);
        return h;
    }
    Dev(const Dev&) = delete;
};

// Stable multisplit by host ids: the output order and the bucket starts.
template <class KeyT, class F>
inline void expected(const std::vector<KeyT>& keys, uint32_t B, F f, std::vector<uint32_t>& order,
                     std::vector<uint32_t>& starts) {
    const size_t n = keys.size();
    std::vector<uint32_t> id(n);
    for (size_t i = 0; i < n; ++i) {
        const uint32_t b = (uint32_t)f(keys[i]);
        id[i] = b < B ? b : B - 1;
    }
    order.resize(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return id[a] < id[b]; });
    starts.assign(B, 0);
    std::vector<uint64_t> cnt(B, 0);
    for (uint32_t b : id) ++cnt[b];
    uint64_t run = 0;
    for (uint32_t b = 0; b < B; ++b) {
        starts[b] = (uint32_t)run;
        run += cnt[b];
    }
}

// One multisplit through the C++ API (multisplit_key_only / _key_value with
// the functor as bucket_t); values are the original positions.  Returns 1 on
// a mismatch.
template <class KeyT, class ValueT, class F>
inline int multisplit_case(const char* name, uint32_t B, uint32_t n, bool kv, F f) {
    std::vector<KeyT> keys(n);
    for (auto& k : keys) k = (KeyT)rnd64();
    std::vector<ValueT> vals(n);
    for (uint32_t i = 0; i < n; ++i) vals[i] = (ValueT)i;
    Dev<KeyT> kin(n), kout(n);
    Dev<ValueT> vin(n), vout(n);
    kin.put(keys);
    vin.put(vals);
    std::vector<uint32_t> offs(B, 0xDEADBEEFu);
    multisplit_context ctx(B);
    try {
        if (kv) {
            multisplit_allocate_key_value(n, ctx);
            multisplit_key_value(kin.p, vin.p, kout.p, vout.p, n, ctx, f, offs.data());
        } else {
            multisplit_allocate_key_only(n, ctx);
            multisplit_key_only(kin.p, kout.p, n, ctx, f, offs.data());
        }
    } catch (const gms::error& e) {
        printf("FAIL %s B=%u n=%u kv=%d: %s\n", name, B, n, (int)kv, e.what());
        return 1;
    }
    std::vector<uint32_t> order, starts;
    expected(keys, B, f, order, starts);
    const std::vector<KeyT> ko = kout.get(n);
    const std::vector<ValueT> vo = kv ? vout.get(n) : std::vector<ValueT>();
    for (uint32_t j = 0; j < n; ++j) {
        if (ko[j] != keys[order[j]] || (kv && vo[j] != (ValueT)order[j])) {
            printf("FAIL %s B=%u n=%u kv=%d: first difference at output %u\n", name, B, n, (int)kv, j);
            return 1;
        }
    }
    if (offs != starts) {
        printf("FAIL %s B=%u n=%u kv=%d: bucket starts\n", name, B, n, (int)kv);
        return 1;
    }
    return 0;
}

// One compaction (compaction_key_only / _key_value with the functor as the
// 2-bucket identifier): stable, bucket 0 first, count = bucket-0 keys.
template <class KeyT, class F>
inline int compaction_case(const char* name, uint32_t n, bool kv, F f) {
    std::vector<KeyT> keys(n);
    for (auto& k : keys) k = (KeyT)rnd64();
    std::vector<uint32_t> vals(n);
    std::iota(vals.begin(), vals.end(), 0u);
    Dev<KeyT> kin(n), kout(n);
    Dev<uint32_t> vin(n), vout(n);
    kin.put(keys);
    vin.put(vals);
    compaction_context ctx;
    unsigned cnt = 0;
    try {
        if (kv) {
            compaction_allocate_key_value(n, ctx);
            cnt = compaction_key_value(kin.p, vin.p, kout.p, vout.p, n, ctx, f);
        } else {
            compaction_allocate_key_only(n, ctx);
            cnt = compaction_key_only(kin.p, kout.p, n, ctx, f);
        }
    } catch (const gms::error& e) {
        printf("FAIL %s compaction n=%u kv=%d: %s\n", name, n, (int)kv, e.what());
        return 1;
    }
    std::vector<uint32_t> order, starts;
    expected(keys, 2, f, order, starts);
    const uint32_t want = n ? (uint32_t)(std::count_if(keys.begin(), keys.end(),
                                                      [&](KeyT k) { return f(k) == 0; }))
                            : 0u;
    const std::vector<KeyT> ko = kout.get(n);
    const std::vector<uint32_t> vo = kv ? vout.get(n) : std::vector<uint32_t>();
    bool ok = cnt == want;
    for (uint32_t j = 0; ok && j < n; ++j)
        ok = ko[j] == keys[order[j]] && (!kv || vo[j] == order[j]);
    if (!ok) {
        printf("FAIL %s compaction n=%u kv=%d (count %u, want %u)\n", name, n, (int)kv, cnt, want);
        return 1;
    }
    return 0;
}

}  // namespace fdrv

// The drivers' case lists (one translation unit each, compiled in parallel):
// every case for each n in ns; returns the number of failures.
int functor_hash_cases(const std::vector<uint32_t>& ns);
int functor_popc_cases(const std::vector<uint32_t>& ns);
int functor_wide_cases(const std::vector<uint32_t>& ns);
