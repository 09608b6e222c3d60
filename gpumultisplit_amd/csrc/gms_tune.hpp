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

// Keys per thread, by log2(bucket count), key-value-ness and input size
// (measured on MI355X).  24 everywhere, except for inputs of at least
// kLongTileMin keys where longer tiles -- longer per-bucket output runs --
// measured faster: 64-bucket key-value (32) and key-only from 128 buckets
// (48).  Below that size the longer tiles leave too few tiles per CU.
constexpr uint32_t kLongTileMin = 1u << 26;

constexpr int tune_items(int log_b, bool kv, bool large) {
    return !large ? 24 : (kv ? (log_b == 6 ? 32 : 24) : (log_b >= 7 ? 48 : 24));
}

// Widest digit ranked in one LDS pass (2^w x 256 u32 counters): 1..4 bits in
// one pass; 5 bits as 3+2, 6 as 3+3, 7 as 4+3, 8 as 4+4.
constexpr int tune_level_bits(int log_b) {
    return log_b <= 4 ? (log_b < 1 ? 1 : log_b) : (log_b <= 6 ? 3 : 4);
}

// Whether a key-value postscan keeps its digit counters inside the value
// staging area (fewer LDS bytes, more resident workgroups).  Not for 128/256
// buckets with short tiles: three resident workgroups per CU then keep more
// partially written output lines open than L2 merges well, which measured
// 35 % slower than two.
constexpr bool tune_alias(int log_b, int items, bool kv, int level_bits) {
    return kv && (1 << level_bits) <= items && !(log_b >= 7 && items < 28);
}

// Keys per tile for an input of n keys.  The shortest tile, used to size the
// histogram matrix for every n up to a capacity, is tile_keys(log_b, kv, 0).
constexpr uint32_t tile_keys(int log_b, bool kv, uint32_t n) {
    return (uint32_t)(kThreads * tune_items(log_b, kv, n >= kLongTileMin));
}

inline int ceil_log2_u32(uint32_t x) {
    int l = 0;
    while ((1u << l) < x && l < 31) ++l;
    return l;
}

}  // namespace gms
