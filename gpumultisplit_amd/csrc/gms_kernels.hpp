// The three multisplit stages for gfx950, plus the helpers they share.
//
//   prescan   H[b * T + t] = #{keys of tile t in bucket b}       (bucket-major)
//   scan      G = exclusive prefix sum of H in (bucket, tile) order
//   postscan  stable reorder of every tile in LDS, then write-out:
//             key i of tile t in bucket b lands at G[b*T+t] + its rank among
//             the tile's bucket-b keys.
//
// Same decomposition as the reference (BMS prescan, CUB scan, BMS postscan;
// src/api/multisplit.cuh:199-313).  The in-tile ranking differs: instead of
// warp ballots per bucket bit, each thread owns a column of digit counters in
// LDS and ranks its K consecutive keys with ds_add_rtn; a block-wide scan of
// the counters in memory (digit-major) order turns them into positions.
// Bucket ids wider than LEVEL_BITS are ranked least-significant part first in
// up to three stable LDS passes.
//
// One workgroup per tile, tiles ordered so that every XCD works on one
// contiguous range (xcd_tile).  Full tiles take a branch-free path whose loads
// are all unconditional, so the key and value loads of a thread are in flight
// together; only the ragged last tile takes guarded loads.
#pragma once

#include "gms_device.hpp"
#include "gms_tune.hpp"

namespace gms {

// Keys of tile t.  32-bit arithmetic on purpose (t * tile < n < 2^32 for
// every launched tile): a uniform 64-bit unsigned compare has no scalar
// instruction on gfx950, and in that form the compiler has been seen to pick
// the result with a condition flag clobbered by an address add, which made
// the ragged last tile count as full.
__device__ __forceinline__ uint32_t tile_valid(uint32_t n, uint32_t t, uint32_t tile) {
    const uint32_t rem = n - t * tile;
    return rem < tile ? rem : tile;
}

// Tile loads.  Full tiles use unconditional loads (16-byte vectors when VEC);
// only the ragged last tile of an input takes the guarded path.
//
// Blocked: thread t owns [t*K, t*K+K).
template <int K, bool VEC>
__device__ __forceinline__ void load_blocked_full(const uint32_t* __restrict__ p, size_t base,
                                                  uint32_t (&x)[K]) {
    const uint32_t off = threadIdx.x * (uint32_t)K;
    if constexpr (VEC) {
        const uint4* q = reinterpret_cast<const uint4*>(p + base + off);
#pragma unroll
        for (int i = 0; i < K / 4; ++i) {
            const uint4 v = q[i];
            x[4 * i + 0] = v.x;
            x[4 * i + 1] = v.y;
            x[4 * i + 2] = v.z;
            x[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) x[k] = p[base + off + k];
    }
}

template <int K>
__device__ __forceinline__ void load_blocked_guarded(const uint32_t* __restrict__ p, size_t base,
                                                     uint32_t nvalid, uint32_t (&x)[K]) {
    const uint32_t off = threadIdx.x * (uint32_t)K;
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = (off + (uint32_t)k < nvalid) ? p[base + off + k] : 0u;
}

// Striped by 16 bytes: thread t owns [4(j*256+t), +4) for j < K/4.  Every
// line is read by one instruction, so the vector loads may stream
// (STREAM: ld_stream, the line not kept for reuse).  That pays for inputs
// larger than the caches: 2^27 keys in 256 buckets, prescan 136 -> 99 us on
// MI355X.  Smaller inputs are better left cached for the postscan's re-read.
// Blocked loads touch each line from K/4 instructions and need it cached
// between them.
template <int K, bool VEC, bool STREAM = false>
__device__ __forceinline__ void load_striped4_full(const uint32_t* __restrict__ p, size_t base,
                                                   uint32_t (&x)[K]) {
#pragma unroll
    for (int j = 0; j < K / 4; ++j) {
        const uint32_t w4 = 4u * ((uint32_t)j * kThreads + threadIdx.x);
        if constexpr (VEC) {
            const uint4* q = reinterpret_cast<const uint4*>(p + base + w4);
            const uint4 v = STREAM ? ld_stream(q) : *q;
            x[4 * j + 0] = v.x;
            x[4 * j + 1] = v.y;
            x[4 * j + 2] = v.z;
            x[4 * j + 3] = v.w;
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) x[4 * j + i] = p[base + w4 + i];
        }
    }
}

template <int K>
__device__ __forceinline__ void load_striped4_guarded(const uint32_t* __restrict__ p, size_t base,
                                                      uint32_t nvalid, uint32_t (&x)[K]) {
#pragma unroll
    for (int j = 0; j < K / 4; ++j) {
        const uint32_t w4 = 4u * ((uint32_t)j * kThreads + threadIdx.x);
#pragma unroll
        for (int i = 0; i < 4; ++i) x[4 * j + i] = (w4 + i < nvalid) ? p[base + w4 + i] : 0u;
    }
}

// ---------------------------------------------------------------------------
// Prescan: per-tile bucket counts.
// ---------------------------------------------------------------------------
// Counts of one tile into s_h, kWaves rows of B words: row q receives wave
// q's counts.  Up to 8 buckets the rows are overwritten; above that the
// waves add into them with LDS atomics, so the rows must be zero on entry.
template <int LOG_B, int K, bool FULL, class F>
__device__ __forceinline__ void prescan_count(const uint32_t (&x)[K], uint32_t nvalid, F f,
                                              uint32_t* s_h) {
    constexpr int B = 1 << LOG_B;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    if constexpr (LOG_B <= 3) {
        // Byte counters for up to 8 buckets in one 64-bit register (K <= 255).
        uint64_t pc = 0;
#pragma unroll
        for (int j = 0; j < K / 4; ++j) {
            const uint32_t w4 = 4u * ((uint32_t)j * kThreads + tid);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (FULL || w4 + i < nvalid) pc += (uint64_t)1 << (8u * f(x[4 * j + i]));
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const uint32_t c = wave_incl_scan((uint32_t)(pc >> (8 * b)) & 0xFFu);
            if (lane == 63u) s_h[wid * B + b] = c;
        }
    } else {
        uint32_t* my = s_h + wid * B;
        // Probed on the wave's first key slot (key 4*tid of the tile).
        const bool coincide = wave_ids_coincide(f(x[0]), FULL || 4u * tid < nvalid);
#pragma unroll
        for (int j = 0; j < K / 4; ++j) {
            const uint32_t w4 = 4u * ((uint32_t)j * kThreads + tid);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t d = f(x[4 * j + i]);
                const bool v = FULL || w4 + i < nvalid;
                if (coincide) wave_count_add(my, d, v);
                else if (v) atomicAdd(&my[d], 1u);
            }
        }
    }
}

// Counts of one tile into H[:, t] (one column of the bucket-major per-tile
// count matrix; the batched kernels' layout).
template <int LOG_B, int K, bool FULL, class F>
__device__ __forceinline__ void prescan_tile(const uint32_t (&x)[K], uint32_t nvalid, uint32_t t,
                                             uint32_t num_tiles, uint32_t* __restrict__ H, F f,
                                             uint32_t* s_h) {
    constexpr int B = 1 << LOG_B;
    const uint32_t tid = threadIdx.x;
    if constexpr (LOG_B > 3) {
        // Per-wave sub-histograms; a wave zeroes and fills only its own.
        uint32_t* my = s_h + (tid >> 6) * B;
        for (uint32_t i = tid & 63u; i < (uint32_t)B; i += 64) my[i] = 0;
        GMS_WAVE_SYNC();
    }
    prescan_count<LOG_B, K, FULL>(x, nvalid, f, s_h);
    __syncthreads();
    for (uint32_t b = tid; b < (uint32_t)B; b += kThreads) {
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q) s += s_h[q * B + b];
        H[(size_t)b * num_tiles + t] = s;
    }
    __syncthreads();
}

// Prescan, one workgroup per tile: H[b * T + t] = count of bucket b in tile
// t.  Tiles are handed to workgroups in the XCD order (xcd_tile), so that
// neighbouring tiles' columns of H share L2 lines.  STREAM: streamed key
// loads (long-tile inputs, larger than the caches).
// `clear` (nclear 64-bit words, or NULL): the status words of the
// single-pass scan that follows, zeroed by workgroup 0.
template <int LOG_B, int K, bool VEC, bool STREAM, class F>
__global__ __launch_bounds__(kThreads) void ms_prescan_kernel(const uint32_t* __restrict__ kin,
                                                              uint32_t n, uint32_t num_tiles,
                                                              uint32_t* __restrict__ H, F f,
                                                              uint64_t* __restrict__ clear,
                                                              uint32_t nclear) {
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < nclear; i += kThreads) store_device_u64(&clear[i], 0ull);
    constexpr int B = 1 << LOG_B;
    constexpr uint32_t TILE = kThreads * K;
    __shared__ __attribute__((aligned(16))) uint32_t s_h[kWaves * B];
    const uint32_t t = xcd_tile(blockIdx.x, num_tiles);
    const uint32_t nvalid = tile_valid(n, t, TILE);
    uint32_t x[K];
    if (nvalid == TILE) {
        load_striped4_full<K, VEC, STREAM>(kin, (size_t)t * TILE, x);
        prescan_tile<LOG_B, K, true>(x, TILE, t, num_tiles, H, f, s_h);
    } else {
        load_striped4_guarded<K>(kin, (size_t)t * TILE, nvalid, x);
        prescan_tile<LOG_B, K, false>(x, nvalid, t, num_tiles, H, f, s_h);
    }
}

// Counts of one tile into s_sum[b * TPW + j], through the per-wave rows of s_h.
template <int LOG_B, int K, int TPW, bool FULL, class F>
__device__ __forceinline__ void group_count(const uint32_t (&x)[K], uint32_t nvalid, F f,
                                            uint32_t* s_h, uint32_t* s_sum, uint32_t j) {
    constexpr int B = 1 << LOG_B;
    const uint32_t tid = threadIdx.x;
    if constexpr (LOG_B > 3) {
        uint32_t* my = s_h + (tid >> 6) * B;
        for (uint32_t i = tid & 63u; i < (uint32_t)B; i += 64) my[i] = 0;
        GMS_WAVE_SYNC();
    }
    prescan_count<LOG_B, K, FULL>(x, nvalid, f, s_h);
    __syncthreads();
    for (uint32_t b = tid; b < (uint32_t)B; b += kThreads) {
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q) s += s_h[q * B + b];
        s_sum[b * TPW + j] = s;
    }
    __syncthreads();
}

// Prescan of TPW adjacent tiles per workgroup (groups in the XCD order),
// counted one after the other; the counts of one bucket in the group's tiles
// are TPW adjacent words of H, written by TPW adjacent lanes of one store.
// Pays where the matrix is large against the keys (512 and 1024 buckets:
// 9216-key tiles write 512 or 1024 words each).  At 2^27 key-value pairs on
// MI355X: 1024 buckets 116 us in groups of 8 tiles, 123 in groups of 4, 137
// one tile per workgroup; 512 buckets 96 in groups of 8, 101 in groups of 4,
// 112 one tile per workgroup; key-only 1024 buckets 106 against 120, 512
// buckets 90 against 102.  At 64 buckets it loses (86 against 79 us).
template <int LOG_B, int K, bool VEC, bool STREAM, int TPW, class F>
__global__ __launch_bounds__(kThreads) void ms_prescan_group_kernel(
    const uint32_t* __restrict__ kin, uint32_t n, uint32_t num_tiles, uint32_t* __restrict__ H,
    F f, uint64_t* __restrict__ clear, uint32_t nclear) {
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < nclear; i += kThreads) store_device_u64(&clear[i], 0ull);
    constexpr int B = 1 << LOG_B;
    constexpr uint32_t TILE = kThreads * K;
    __shared__ __attribute__((aligned(16))) uint32_t s_h[kWaves * B];
    __shared__ uint32_t s_sum[B * TPW];
    const uint32_t tid = threadIdx.x;
    const uint32_t ngroups = (num_tiles + TPW - 1) / TPW;
    const uint32_t t0 = xcd_tile(blockIdx.x, ngroups) * (uint32_t)TPW;
    const uint32_t nt = num_tiles - t0 < (uint32_t)TPW ? num_tiles - t0 : (uint32_t)TPW;
    for (uint32_t j = 0; j < nt; ++j) {
        uint32_t x[K];
        const uint32_t t = t0 + j;
        const uint32_t nv = tile_valid(n, t, TILE);
        if (nv == TILE) {
            load_striped4_full<K, VEC, STREAM>(kin, (size_t)t * TILE, x);
            group_count<LOG_B, K, TPW, true>(x, TILE, f, s_h, s_sum, j);
        } else {
            load_striped4_guarded<K>(kin, (size_t)t * TILE, nv, x);
            group_count<LOG_B, K, TPW, false>(x, nv, f, s_h, s_sum, j);
        }
    }
    for (uint32_t i = tid; i < (uint32_t)(B * TPW); i += kThreads) {
        const uint32_t b = i / TPW, j = i % TPW;
        if (j < nt) H[(size_t)b * num_tiles + t0 + j] = s_sum[i];
    }
}

// ---------------------------------------------------------------------------
// One stable LDS ranking pass over the digit (bucket >> SHIFT) & (2^W - 1).
// On entry key/val hold the thread's K consecutive keys of the current order
// (thread t owns positions [t*K, t*K+K)); on exit s_keys/s_vals hold the tile
// stably reordered by that digit.  Ends with a barrier.
//
// s_cnt is D x kThreads u32, digit-major: word d*kThreads + t counts digit d among the
// keys of thread t.  Counting hits 64 distinct banks per wave; the scan reads
// D consecutive words per thread, so position = exclusive prefix in memory
// order = (keys of smaller digits) + (keys of this digit in earlier threads).
//
// ALIAS: s_cnt lies inside s_vals (key-value tiles), which keeps the LDS of a
// 256-bucket key-value workgroup under a third of the CU's 160 KB.  Every
// thread then reads all its positions into registers, and a barrier separates
// those reads from the scatter that overwrites the counters.
// ---------------------------------------------------------------------------
template <int SHIFT, int W, int K, bool KV, bool FULL, bool ALIAS, class F>
__device__ __forceinline__ void rank_pass(const uint32_t (&key)[K], const uint32_t (&val)[K],
                                          uint32_t nvalid, F f, uint32_t* s_keys,
                                          uint32_t* s_vals, uint32_t* s_cnt, uint32_t* s_ws) {
    constexpr int D = 1 << W;
    constexpr uint32_t DMASK = (uint32_t)D - 1u;
    constexpr uint32_t PAD = (uint32_t)cnt_pad(D), ROW = (uint32_t)cnt_row(D);
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t first = tid * (uint32_t)K;
    // LDS word of logical counter w = d*kThreads + t (rows of D padded to ROW).
    auto cidx = [](uint32_t w) { return w + PAD * (w / (uint32_t)D); };

#pragma unroll
    for (int d = 0; d < D; ++d) s_cnt[cidx((uint32_t)d * kThreads + tid)] = 0;

    // slot[k] = (LDS word of key k's counter) << 8 | (its rank among this
    // thread's keys of the same digit, < K <= 255): one register per key.
    static_assert(K <= 255, "ranks among a thread's keys fit in 8 bits");
    uint32_t slot[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        slot[k] = cidx(tid) << 8;
        if (FULL || first + (uint32_t)k < nvalid) {
            const uint32_t d = (f(key[k]) >> SHIFT) & DMASK;
            const uint32_t w = cidx(d * (uint32_t)kThreads + tid);
            slot[k] = (w << 8) | atomicAdd(&s_cnt[w], 1u);
        }
    }
    __syncthreads();

    uint32_t c[D];
    uint32_t* mine = s_cnt + tid * ROW;
    if constexpr (D % 4 == 0) {
#pragma unroll
        for (int i = 0; i < D / 4; ++i) {
            const uint4 v = reinterpret_cast<const uint4*>(mine)[i];
            c[4 * i + 0] = v.x; c[4 * i + 1] = v.y; c[4 * i + 2] = v.z; c[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < D; ++i) c[i] = mine[i];
    }
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) s += c[i];
    const uint32_t inc = wave_incl_scan(s);
    if (lane == 63u) s_ws[wid] = inc;
    __syncthreads();
    uint32_t run = inc - s;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) run += ((uint32_t)i < wid) ? s_ws[i] : 0u;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const uint32_t v = c[i];
        c[i] = run;
        run += v;
    }
    if constexpr (D % 4 == 0) {
#pragma unroll
        for (int i = 0; i < D / 4; ++i)
            reinterpret_cast<uint4*>(mine)[i] =
                make_uint4(c[4 * i + 0], c[4 * i + 1], c[4 * i + 2], c[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < D; ++i) mine[i] = c[i];
    }
    __syncthreads();

    if constexpr (ALIAS) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t pos = s_cnt[slot[k] >> 8] + (slot[k] & 0xFFu);
            slot[k] = pos < nvalid ? pos : nvalid - 1u;  // in bounds even if counts were inconsistent
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (FULL || first + (uint32_t)k < nvalid) {
                s_keys[slot[k]] = key[k];
                if constexpr (KV) s_vals[slot[k]] = val[k];
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (FULL || first + (uint32_t)k < nvalid) {
                uint32_t pos = s_cnt[slot[k] >> 8] + (slot[k] & 0xFFu);
                pos = pos < nvalid ? pos : nvalid - 1u;  // in bounds even if counts were inconsistent
                s_keys[pos] = key[k];
                if constexpr (KV) s_vals[pos] = val[k];
            }
        }
    }
    __syncthreads();
}

template <int K, bool KV>
__device__ __forceinline__ void reload_blocked(uint32_t (&key)[K], uint32_t (&val)[K],
                                               const uint32_t* s_keys, const uint32_t* s_vals) {
    static_assert(K % 4 == 0, "16-byte LDS reads of whole rows");
    const uint32_t off = threadIdx.x * (uint32_t)K;
#pragma unroll
    for (int i = 0; i < K / 4; ++i) {
        uint4 a = *reinterpret_cast<const uint4*>(s_keys + off + 4 * i);
        key[4 * i + 0] = a.x; key[4 * i + 1] = a.y; key[4 * i + 2] = a.z; key[4 * i + 3] = a.w;
        if constexpr (KV) {
            uint4 b = *reinterpret_cast<const uint4*>(s_vals + off + 4 * i);
            val[4 * i + 0] = b.x; val[4 * i + 1] = b.y; val[4 * i + 2] = b.z; val[4 * i + 3] = b.w;
        }
    }
}

// ---------------------------------------------------------------------------
// Postscan: stable in-tile reorder + write-out, one workgroup per tile.
// ---------------------------------------------------------------------------
// In-tile ranking of one tile: delta[b] = global start - local start of every
// bucket run, then the stable LDS passes; leaves the tile in bucket order in
// s_keys/s_vals.
// Thread tid holds buckets tid*R .. tid*R+R-1 (R = 1 up to 256 buckets): their
// global starts g0[] and the tile's counts cnt[] (0 past the bucket count).
template <int LOG_B, int K, int LEVEL_BITS, bool KV, bool FULL, bool ALIAS, int R, class F>
__device__ __forceinline__ void postscan_rank(uint32_t (&key)[K], uint32_t (&val)[K],
                                              const uint32_t (&g0)[R], const uint32_t (&cnt)[R],
                                              uint32_t nvalid, F f,
                                              uint32_t* s_keys, uint32_t* s_vals,
                                              uint32_t* s_cnt, uint32_t* s_delta, uint32_t* s_bs,
                                              uint32_t* s_ws) {
    constexpr int B = 1 << LOG_B;
    constexpr int NLEV = (LOG_B + LEVEL_BITS - 1) / LEVEL_BITS;
    static_assert(R * kThreads >= B, "R buckets per thread cover every bucket");
    const uint32_t tid = threadIdx.x;
    {
        uint32_t sum = 0, tot;
#pragma unroll
        for (int r = 0; r < R; ++r) sum += cnt[r];
        uint32_t ls = block_excl_scan<kWaves>(sum, s_bs, tot);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t b = tid * (uint32_t)R + (uint32_t)r;
            if (b < (uint32_t)B) s_delta[b] = g0[r] - ls;
            ls += cnt[r];
        }
    }
    constexpr int W0 = LOG_B < LEVEL_BITS ? LOG_B : LEVEL_BITS;
    rank_pass<0, W0, K, KV, FULL, ALIAS>(key, val, nvalid, f, s_keys, s_vals, s_cnt, s_ws);
    if constexpr (NLEV >= 2) {
        constexpr int S1 = LEVEL_BITS;
        constexpr int W1 = (LOG_B - S1) < LEVEL_BITS ? (LOG_B - S1) : LEVEL_BITS;
        reload_blocked<K, KV>(key, val, s_keys, s_vals);
        if constexpr (ALIAS) __syncthreads();  // values read before the counters are cleared
        rank_pass<S1, W1, K, KV, FULL, ALIAS>(key, val, nvalid, f, s_keys, s_vals, s_cnt, s_ws);
    }
    if constexpr (NLEV >= 3) {
        constexpr int S2 = 2 * LEVEL_BITS;
        constexpr int W2 = LOG_B - S2;
        reload_blocked<K, KV>(key, val, s_keys, s_vals);
        if constexpr (ALIAS) __syncthreads();
        rank_pass<S2, W2, K, KV, FULL, ALIAS>(key, val, nvalid, f, s_keys, s_vals, s_cnt, s_ws);
    }
}

// Write-out in slot order: consecutive threads write consecutive slots, which
// are consecutive addresses inside each bucket run.  The index is clamped
// into [0, n): a no-op when the scanned matrix is consistent, and every store
// stays in bounds even if it were not.
template <int K, bool KV, bool FULL, class F>
__device__ __forceinline__ void postscan_write(uint32_t nvalid, uint32_t n,
                                               uint32_t* __restrict__ kout,
                                               uint32_t* __restrict__ vout, F f,
                                               const uint32_t* s_keys, const uint32_t* s_vals,
                                               const uint32_t* s_delta) {
    const uint32_t tid = threadIdx.x;
    // Unrolled by 4, not fully: the compiler would otherwise compute every
    // store's 64-bit address up front, which costs the registers a caller
    // keeps its next tile in.
#pragma unroll 4
    for (int j = 0; j < K; ++j) {
        const uint32_t p = (uint32_t)j * kThreads + tid;
        if (FULL || p < nvalid) {
            const uint32_t k = s_keys[p];
            uint32_t gp = s_delta[f(k)] + p;
            gp = gp < n ? gp : n - 1u;
            kout[gp] = k;
            if constexpr (KV) vout[gp] = s_vals[p];
        }
    }
}

// Word `off` of the count matrix.  The matrix of one call is below 4 GB (at
// most 1024 rows of 2^32 / 6144 tiles), so the byte offset is 32-bit: one
// register per address instead of two.
__device__ __forceinline__ uint32_t matrix_word(const uint32_t* __restrict__ m, uint32_t off) {
    return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(m) + (off << 2));
}

// Postscan, one workgroup per tile, tiles in the XCD order (xcd_tile).  The
// same launch shape as the prescan: on MI355X a workgroup per tile, handed
// out by the dispatcher in tile order, measured faster than resident
// workgroups that walk several tiles each with the next one prefetched
// (426 against 800 us at 2^27 key-value pairs, 256 buckets).
template <int LOG_B, int K, int LEVEL_BITS, bool KV, bool VEC, class F>
__global__ __launch_bounds__(kThreads) void ms_postscan_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, uint32_t num_tiles,
    const uint32_t* __restrict__ G, F f) {
    constexpr int B = 1 << LOG_B;
    constexpr uint32_t TILE = kThreads * K;
    constexpr int NLEV = (LOG_B + LEVEL_BITS - 1) / LEVEL_BITS;
    static_assert(NLEV >= 1 && NLEV <= 3, "at most three ranking passes");
    static_assert(K % 4 == 0, "vector loads");

    // Key-value tiles keep the digit counters inside the value staging area
    // (ALIAS in rank_pass); key-only tiles have no values and a separate area.
    constexpr uint32_t CNT = cnt_words(LEVEL_BITS);
    constexpr bool ALIAS = tune_alias(LOG_B, K, KV, LEVEL_BITS);
    __shared__ __attribute__((aligned(16))) uint32_t s_keys[TILE];
    __shared__ __attribute__((aligned(16))) uint32_t s_vc[KV ? TILE : CNT];
    __shared__ __attribute__((aligned(16))) uint32_t s_cnt_sep[(KV && !ALIAS) ? CNT : 4];
    __shared__ uint32_t s_delta[B];
    __shared__ uint32_t s_bs[kWaves];
    __shared__ uint32_t s_ws[kWaves];
    uint32_t* s_vals = s_vc;
    uint32_t* s_cnt = KV ? (ALIAS ? s_vc : s_cnt_sep) : s_vc;

    const uint32_t tid = threadIdx.x;
    const uint32_t t = xcd_tile(blockIdx.x, num_tiles);
    const uint32_t nvalid = tile_valid(n, t, TILE);
    const size_t base = (size_t)t * TILE;

    // Scanned entry G[b][t] and the next one in scan order (n past the end):
    // their difference is the tile's count in bucket b.  Thread tid loads
    // buckets tid*R .. tid*R+R-1 (threads past B reload bucket B-1 and ignore
    // it).
    constexpr int R = (B + kThreads - 1) / kThreads;
    const uint32_t mtot = (uint32_t)B * num_tiles;
    uint32_t g0[R], g1raw[R];
    bool last[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t b = tid * (uint32_t)R + (uint32_t)r;
        const uint32_t idx = (b < (uint32_t)B ? b : (uint32_t)B - 1u) * num_tiles + t;
        last[r] = idx + 1 >= mtot;
        g0[r] = matrix_word(G, idx);
        g1raw[r] = matrix_word(G, last[r] ? idx : idx + 1);
    }
    auto tile_counts = [&](uint32_t (&cnt)[R]) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            cnt[r] = tid * (uint32_t)R + (uint32_t)r < (uint32_t)B
                         ? (last[r] ? n : g1raw[r]) - g0[r]
                         : 0u;
    };

    uint32_t key[K], val[K];
    if (nvalid == TILE) {
        load_blocked_full<K, VEC>(kin, base, key);
        if constexpr (KV) load_blocked_full<K, VEC>(vin, base, val);
        else {
#pragma unroll
            for (int k = 0; k < K; ++k) val[k] = 0;
        }
        uint32_t cnt[R];
        tile_counts(cnt);
        postscan_rank<LOG_B, K, LEVEL_BITS, KV, true, ALIAS>(key, val, g0, cnt, TILE, f, s_keys,
                                                             s_vals, s_cnt, s_delta, s_bs, s_ws);
        postscan_write<K, KV, true>(TILE, n, kout, vout, f, s_keys, s_vals, s_delta);
    } else {
        load_blocked_guarded<K>(kin, base, nvalid, key);
        if constexpr (KV) load_blocked_guarded<K>(vin, base, nvalid, val);
        else {
#pragma unroll
            for (int k = 0; k < K; ++k) val[k] = 0;
        }
        uint32_t cnt[R];
        tile_counts(cnt);
        postscan_rank<LOG_B, K, LEVEL_BITS, KV, false, ALIAS>(key, val, g0, cnt, nvalid, f, s_keys,
                                                              s_vals, s_cnt, s_delta, s_bs, s_ws);
        postscan_write<K, KV, false>(nvalid, n, kout, vout, f, s_keys, s_vals, s_delta);
    }
}

// ---------------------------------------------------------------------------
// Bucket starts of a finished multisplit output of n elements of KW words:
// start[b] = the first i with f(element i) >= b, which is n when no element
// is.  The ids along the output are non-decreasing, so one thread per bucket
// binary-searches for it: at most ceil(log2(n + 1)) steps whatever the
// distribution (no thread walks a run of empty buckets).  Neighbouring
// buckets follow the same first steps, which share cache lines.
// ---------------------------------------------------------------------------
template <int KW, class F>
__global__ __launch_bounds__(kThreads) void bucket_starts_kernel(const uint32_t* __restrict__ out,
                                                                 uint32_t n, uint32_t num_buckets,
                                                                 F f, uint32_t* __restrict__ start) {
    const uint32_t b = blockIdx.x * kThreads + threadIdx.x;
    if (b >= num_buckets) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t id;
        if constexpr (KW == 2)
            id = f((uint64_t)out[2 * (size_t)mid] | ((uint64_t)out[2 * (size_t)mid + 1] << 32));
        else
            id = f(out[mid]);
        if (id < b) lo = mid + 1u;
        else hi = mid;
    }
    start[b] = lo;
}

template <int KW, class F>
hipError_t launch_bucket_starts(const uint32_t* out, uint32_t n, uint32_t num_buckets, const F& f,
                                uint32_t* start, hipStream_t stream) {
    hipLaunchKernelGGL((bucket_starts_kernel<KW, F>), dim3((num_buckets + kThreads - 1) / kThreads),
                       dim3(kThreads), 0, stream, out, n, num_buckets, f, start);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of u32 (reduce, scan of partials, downsweep).
// ---------------------------------------------------------------------------
constexpr int kScanItems = 16;
constexpr uint32_t kScanChunk = kThreads * kScanItems;  // 4096 elements per workgroup
constexpr uint32_t kScanFusedParts = 2048;  // chunks up to which the partials pass is folded

__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                               uint32_t m,
                                                               uint32_t* __restrict__ part);
__global__ __launch_bounds__(1024) void scan_partials_kernel(uint32_t* __restrict__ part,
                                                             uint32_t np);
template <bool SUM_PARTS>
__global__ __launch_bounds__(kThreads) void scan_down_kernel(const uint32_t* in, uint32_t* out,
                                                             uint32_t m,
                                                             const uint32_t* __restrict__ part);
__global__ void gather_offsets_kernel(const uint32_t* __restrict__ G, uint32_t num_tiles,
                                      uint32_t num_buckets, uint32_t* __restrict__ out);
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, uint32_t n);
__global__ __launch_bounds__(kThreads) void copy_u32_kernel(const uint32_t* __restrict__ src,
                                                            uint32_t* __restrict__ dst, size_t n);

}  // namespace gms
