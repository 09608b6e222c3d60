// Population count mod B (a non-monotone identifier), any bucket count, and
// its parity as a compaction predicate.
#include "functor_common.hpp"

struct popc_bucket {
    uint32_t B;
    __host__ __device__ uint32_t operator()(uint32_t k) const {
        return (uint32_t)__builtin_popcount(k) % B;
    }
};

int functor_popc_cases(const std::vector<uint32_t>& ns) {
    int fails = 0;
    for (uint32_t n : ns) {
        for (uint32_t B : {2u, 3u, 17u, 33u, 100u})
            for (bool kv : {false, true})
                fails += fdrv::multisplit_case<uint32_t, uint32_t>("popc", B, n, kv, popc_bucket{B});
        for (bool kv : {false, true})
            fails += fdrv::compaction_case<uint32_t>("popc-parity", n, kv, popc_bucket{2});
    }
    return fails;
}
