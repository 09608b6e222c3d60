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
// measured faster: 64- and 256-bucket key-value (28: at 2^27 pairs the
// 256-bucket call takes 0.564 ms against 0.571 at 24, 0.576 at 20, 0.586 at
// 32; the 64-bucket one 0.515 against 0.523 at 24 and 0.532 at 32), 512 and
// 1024 buckets key-value (36, the longest tile that keeps two workgroups per
// CU: 1024 buckets 0.820 ms against 0.939 at 24, 0.882 at 28, 0.901 at 32)
// and key-only: 128 and 256 buckets 56 (256 buckets 0.379 ms against 0.388
// at 48 and 0.394 at 40), 512 and 1024 buckets 48 (at 56 a 1024-bucket
// workgroup no longer fits twice in a CU: 0.705 against 0.521 ms).  Below that
// size the longer tiles leave too few tiles per CU, except for 1024-bucket
// key-value, whose long tiles pay from 2^24 keys (0.116 ms against 0.132 at
// 2^24, 0.229 against 0.266 at 2^25; 512 buckets there are within 2 % either
// way, 256 buckets at 2^25 0.161 against 0.156).
constexpr uint32_t kLongTileMin = 1u << 26;
constexpr uint32_t long_tile_min(int log_b, bool kv) {
    return kv && log_b >= 10 ? kLongTileMin / 4 : kLongTileMin;
}
// The prescan streams its key loads past the caches from this many keys;
// smaller inputs are re-read by the postscan while still in the last-level
// cache.
constexpr uint32_t kStreamLoadMin = 1u << 26;

constexpr int tune_items(int log_b, bool kv, bool large) {
    if (!large) return 24;
    if (!kv) return log_b >= 9 ? 48 : (log_b >= 7 ? 56 : 24);
    return log_b >= 9 ? 36 : (log_b == 6 || log_b == 8 ? 28 : 24);
}

// Widest digit ranked in one LDS pass (2^w x 256 u32 counters): 1..4 bits in
// one pass; 5 bits as 3+2, 6 as 3+3, 7 as 4+3, 8 as 4+4.
constexpr int tune_level_bits(int log_b) {
    return log_b <= 4 ? (log_b < 1 ? 1 : log_b) : (log_b <= 6 ? 3 : 4);
}

// Digit counters of a ranking pass with D digits: logically D x 256 u32,
// digit-major (word d*256 + t), and each thread scans the D consecutive
// logical words [t*D, t*D+D) -- its "row".  Rows are padded in LDS to
// D + cnt_pad(D) words so that the rows of 8 consecutive lanes start on
// different banks (row stride / 4 odd for 16-byte accesses; stride 3 for
// D = 2): unpadded, the 16-byte row reads of D = 16 are 16-way bank conflicts.
constexpr int cnt_pad(int D) { return D % 4 == 0 ? (((D / 4) & 1) ? 0 : 4) : (D == 2 ? 1 : 0); }
constexpr int cnt_row(int D) { return D + cnt_pad(D); }
// LDS words of the counters of passes of up to level_bits bits.
constexpr uint32_t cnt_words(int level_bits) { return (uint32_t)cnt_row(1 << level_bits) * kThreads; }

// Whether a key-value postscan keeps its digit counters inside the value
// staging area (fewer LDS bytes, more resident workgroups).  Not for 128/256
// buckets with short tiles: three resident workgroups per CU then keep more
// partially written output lines open than L2 merges well, which measured
// 35 % slower than two.
constexpr bool tune_alias(int log_b, int items, bool kv, int level_bits) {
    return kv && cnt_row(1 << level_bits) <= items && !(log_b >= 7 && items < 28);
}

// Bucket counts up to this run in one prescan/scan/postscan pass (three LDS
// ranking passes of at most 4 bits above 256); more take LSD passes over
// 8-bit digits of the bucket id.
constexpr uint32_t kSinglePassBuckets = 1024;

// Keys per tile for an input of n keys.  The shortest tile, used to size the
// histogram matrix for every n up to a capacity, is tile_keys(log_b, kv, 0).
constexpr uint32_t tile_keys(int log_b, bool kv, uint32_t n) {
    return (uint32_t)(kThreads * tune_items(log_b, kv, n >= long_tile_min(log_b, kv)));
}

inline int ceil_log2_u32(uint32_t x) {
    int l = 0;
    while ((1u << l) < x && l < 31) ++l;
    return l;
}

}  // namespace gms
