// Tile shapes of the plain multisplit, shared by host code and kernels.
//
// One workgroup of 256 threads (4 waves of 64) owns one tile of 256*K keys.
// K keys per thread are held in registers in blocked order; the local stable
// reorder ranks digits of at most LEVEL_BITS bits per LDS pass.  The prescan
// uses the same tile so that both kernels agree on the histogram matrix.
#pragma once

#include <stdint.h>

namespace gms {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// Keys per thread, by log2(bucket count) and key-value-ness.
constexpr int tune_items(int log_b, bool kv) {
    (void)log_b;
    (void)kv;
    return 24;
}

// Widest digit ranked in one LDS pass (2^w x 256 u32 counters).  Up to 5 bits
// in one pass; 6 bits as 3+3, 7 as 4+3, 8 as 4+4.
constexpr int tune_level_bits(int log_b) {
    return log_b <= 5 ? (log_b < 1 ? 1 : log_b) : (log_b == 6 ? 3 : 4);
}

constexpr uint32_t tile_keys(int log_b, bool kv) { return (uint32_t)(kThreads * tune_items(log_b, kv)); }

inline int ceil_log2_u32(uint32_t x) {
    int l = 0;
    while ((1u << l) < x && l < 31) ++l;
    return l;
}

}  // namespace gms
