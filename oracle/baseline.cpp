// CPU baselines of the radix sort, histogram and compaction rows, timed by
// bench.py's cpu_baseline leg on the host cores of the GPU machine.
//
// TEST INFRASTRUCTURE ONLY, like oracle.c: only tests/ and bench.py's
// cpu_baseline leg call these, never the product path.  They are the
// reference drivers' own CPU comparisons, restated with the standard library:
//   sort        std::stable_sort of (key, value) pairs by key (the reference
//               checks its multisplit sort against a sort, src/main/main_sort.cu:
//               162-229, main_sort_pairs.cu);
//   histogram   a counting loop over (key >> shift) & (bins - 1)
//               (src/main/main_histogram.cu:140-220);
//   compaction  std::stable_partition of (key, value) pairs by the 2-bucket
//               predicate, count = partition point
//               (src/main/main_compaction.cu:77-181).
// Single-threaded.  Pinned against numpy in tests/test_oracle.py.
#include <stdint.h>

#include <algorithm>
#include <utility>
#include <vector>

extern "C" {

// Key-only (vin == NULL) or key-value stable sort.
void ms_baseline_stable_sort(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                             uint32_t* vout, uint64_t n) {
    if (!vin) {
        std::copy(kin, kin + n, kout);
        std::stable_sort(kout, kout + n);
        return;
    }
    std::vector<std::pair<uint32_t, uint32_t>> kv(n);
    for (uint64_t i = 0; i < n; ++i) kv[i] = {kin[i], vin[i]};
    std::stable_sort(kv.begin(), kv.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    for (uint64_t i = 0; i < n; ++i) {
        kout[i] = kv[i].first;
        vout[i] = kv[i].second;
    }
}

// hist[b] += #{i : (key_i >> shift) & (bins - 1) == b}, bins a power of two.
void ms_baseline_histogram(const uint32_t* keys, uint64_t n, uint32_t shift, uint32_t bins,
                           uint32_t* hist) {
    const uint32_t mask = bins - 1u;
    for (uint64_t i = 0; i < n; ++i) ++hist[(keys[i] >> shift) & mask];
}

// Stable partition: keys below `threshold` first (bucket 0 of the predicate
// key / threshold), each side in input order.  Returns the bucket-0 count.
uint64_t ms_baseline_stable_partition(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                      uint32_t* vout, uint64_t n, uint64_t threshold) {
    if (!vin) {
        std::copy(kin, kin + n, kout);
        return (uint64_t)(std::stable_partition(kout, kout + n,
                                                [&](uint32_t k) { return k < threshold; }) -
                          kout);
    }
    std::vector<std::pair<uint32_t, uint32_t>> kv(n);
    for (uint64_t i = 0; i < n; ++i) kv[i] = {kin[i], vin[i]};
    const auto mid = std::stable_partition(kv.begin(), kv.end(),
                                           [&](const auto& p) { return p.first < threshold; });
    for (uint64_t i = 0; i < n; ++i) {
        kout[i] = kv[i].first;
        vout[i] = kv[i].second;
    }
    return (uint64_t)(mid - kv.begin());
}

}  // extern "C"
