// Multiplicative hash into 2^w buckets, (k * 0x9E3779B1) >> (32 - w), and a
// functor whose ids run past the bucket count (clamped into the last bucket).
#include "functor_common.hpp"

struct hash_bucket {
    uint32_t shift;  // 32 - log2(B)
    __host__ __device__ uint32_t operator()(uint32_t k) const { return (k * 0x9E3779B1u) >> shift; }
};

struct wide_id_bucket {  // ids up to 4095 for any bucket count
    __host__ __device__ uint32_t operator()(uint32_t k) const { return k >> 20; }
};

int functor_hash_cases(const std::vector<uint32_t>& ns) {
    int fails = 0;
    for (uint32_t n : ns)
        for (uint32_t lb : {1u, 3u, 6u, 8u, 10u})
            for (bool kv : {false, true})
                fails += fdrv::multisplit_case<uint32_t, uint32_t>("hash", 1u << lb, n, kv,
                                                                   hash_bucket{32u - lb});
    for (uint32_t n : ns)
        for (uint32_t B : {1u, 5u, 256u})
            fails += fdrv::multisplit_case<uint32_t, uint32_t>("clamped", B, n, true, wide_id_bucket{});
    return fails;
}
