// 64-bit keys (and values) with a 64-bit multiplicative hash,
// (k * 0x9E3779B97F4A7C15) >> (64 - w), and 32-bit keys with 64-bit values.
#include "functor_common.hpp"

struct hash64_bucket {
    uint32_t shift;  // 64 - log2(B)
    __host__ __device__ uint32_t operator()(uint64_t k) const {
        return (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> shift);
    }
};

struct hash32_bucket {
    uint32_t shift;
    __host__ __device__ uint32_t operator()(uint32_t k) const { return (k * 0x85EBCA6Bu) >> shift; }
};

int functor_wide_cases(const std::vector<uint32_t>& ns) {
    int fails = 0;
    for (uint32_t n : ns) {
        for (uint32_t lb : {1u, 5u, 8u}) {
            const hash64_bucket f{64u - lb};
            fails += fdrv::multisplit_case<uint64_t, uint64_t>("hash64 KO", 1u << lb, n, false, f);
            fails += fdrv::multisplit_case<uint64_t, uint64_t>("hash64/64", 1u << lb, n, true, f);
            fails += fdrv::multisplit_case<uint64_t, uint32_t>("hash64/32", 1u << lb, n, true, f);
            fails += fdrv::multisplit_case<uint32_t, uint64_t>("hash32/64", 1u << lb, n, true,
                                                               hash32_bucket{32u - lb});
        }
        fails += fdrv::compaction_case<uint64_t>("hash64-compaction", n, true, hash64_bucket{63u});
    }
    return fails;
}
