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
// warp ballots per bucket bit, each thread owns a column of packed 16-bit
// digit counters in LDS and ranks its K consecutive keys with ds_add_rtn, a
// block-wide column scan turns the counters into positions, and digits wider
// than LEVEL_BITS are ranked least-significant part first in up to three
// stable LDS passes.
#pragma once

#include "gms_device.hpp"
#include "gms_tune.hpp"

namespace gms {

// K consecutive u32 of a tile into registers (thread t owns [t*K, t*K+K)).
template <int K>
__device__ __forceinline__ void load_blocked(const uint32_t* __restrict__ p, size_t base,
                                             uint32_t nvalid, uint32_t (&x)[K], bool vec) {
    const uint32_t off = threadIdx.x * (uint32_t)K;
    if (vec) {
        const uint4* q = reinterpret_cast<const uint4*>(p + base + off);
#pragma unroll
        for (int i = 0; i < K / 4; ++i) {
            uint4 v = q[i];
            x[4 * i + 0] = v.x;
            x[4 * i + 1] = v.y;
            x[4 * i + 2] = v.z;
            x[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) x[k] = (off + (uint32_t)k < nvalid) ? p[base + off + k] : 0u;
    }
}

// ---------------------------------------------------------------------------
// Prescan: per-tile bucket counts.
// ---------------------------------------------------------------------------
template <int LOG_B, int K, class F>
__global__ __launch_bounds__(kThreads) void ms_prescan_kernel(const uint32_t* __restrict__ kin,
                                                              uint32_t n, uint32_t num_tiles,
                                                              uint32_t* __restrict__ H, F f,
                                                              int vec_ok) {
    constexpr int B = 1 << LOG_B;
    constexpr uint32_t TILE = kThreads * K;
    __shared__ uint32_t s_h[kWaves * B];

    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t tile = xcd_tile(blockIdx.x, num_tiles);
    const size_t base = (size_t)tile * TILE;
    const uint32_t nvalid = (uint32_t)(((size_t)n - base) < TILE ? ((size_t)n - base) : TILE);
    const bool vec = vec_ok && nvalid == TILE;

    if constexpr (LOG_B <= 3) {
        // Byte counters for up to 8 buckets in one 64-bit register (K <= 255).
        uint64_t pc = 0;
#pragma unroll
        for (int j = 0; j < K / 4; ++j) {
            const uint32_t w4 = 4u * ((uint32_t)j * kThreads + tid);
            uint32_t x[4];
            if (vec) {
                uint4 v = *reinterpret_cast<const uint4*>(kin + base + w4);
                x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) x[i] = (w4 + i < nvalid) ? kin[base + w4 + i] : 0u;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (vec || w4 + i < nvalid) pc += (uint64_t)1 << (8u * f(x[i]));
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            uint32_t c = wave_incl_scan((uint32_t)(pc >> (8 * b)) & 0xFFu);
            if (lane == 63u) s_h[wid * B + b] = c;
        }
    } else {
        for (uint32_t i = tid; i < (uint32_t)(kWaves * B); i += kThreads) s_h[i] = 0;
        __syncthreads();
        uint32_t* my = s_h + wid * B;
#pragma unroll
        for (int j = 0; j < K / 4; ++j) {
            const uint32_t w4 = 4u * ((uint32_t)j * kThreads + tid);
            uint32_t x[4];
            if (vec) {
                uint4 v = *reinterpret_cast<const uint4*>(kin + base + w4);
                x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) x[i] = (w4 + i < nvalid) ? kin[base + w4 + i] : 0u;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (vec || w4 + i < nvalid) atomicAdd(&my[f(x[i])], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = tid; b < (uint32_t)B; b += kThreads) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += s_h[w * B + b];
        H[(size_t)b * num_tiles + tile] = s;
    }
}

// ---------------------------------------------------------------------------
// One stable LDS ranking pass over the digit (bucket >> SHIFT) & (2^W - 1).
// On entry key/val hold the thread's K consecutive keys of the current order
// (thread t owns positions [t*K, t*K+K)); on exit s_keys/s_vals hold the tile
// stably reordered by that digit.  Ends with a barrier.
// ---------------------------------------------------------------------------
template <int SHIFT, int W, int K, bool KV, class F>
__device__ __forceinline__ void rank_pass(const uint32_t (&key)[K], const uint32_t (&val)[K],
                                          uint32_t nvalid, F f, uint32_t* s_keys,
                                          uint32_t* s_vals, uint32_t* s_cnt, uint32_t* s_wsum) {
    constexpr int RW = (W >= 1) ? (1 << (W - 1)) : 1;  // packed counter words per thread
    constexpr uint32_t DMASK = (1u << W) - 1u;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t first = tid * (uint32_t)K;

    // Thread-private column of counters: word r holds digits 2r (low half)
    // and 2r+1 (high half) of this thread.  Column layout r*256+tid keeps
    // every ds access of a wave on 64 distinct banks.
#pragma unroll
    for (int r = 0; r < RW; ++r) s_cnt[r * kThreads + tid] = 0;

    uint32_t lr[K], dg[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        dg[k] = 0;
        lr[k] = 0;
        if (first + (uint32_t)k < nvalid) {
            const uint32_t d = (f(key[k]) >> SHIFT) & DMASK;
            const uint32_t sh = (d & 1u) << 4;
            const uint32_t old = atomicAdd(&s_cnt[(d >> 1) * kThreads + tid], 1u << sh);
            lr[k] = (old >> sh) & 0xFFFFu;
            dg[k] = d;
        }
    }

    // Column scan: for every digit, exclusive prefix over threads, plus the
    // digit's start in the tile.  Both halves of a packed word scan at once
    // (every partial sum is <= tile size < 2^16).
    uint32_t c[RW], inc[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        c[r] = s_cnt[r * kThreads + tid];
        inc[r] = wave_incl_scan(c[r]);
        if (lane == 63u) s_wsum[wid * RW + r] = inc[r];
    }
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < kWaves; ++i) {
            const uint32_t v = s_wsum[i * RW + r];
            pre += ((uint32_t)i < wid) ? v : 0u;
            tot += v;
        }
        const uint32_t ex = pre + inc[r] - c[r];
        const uint32_t lo_start = run;
        run += tot & 0xFFFFu;
        const uint32_t hi_start = run;
        run += tot >> 16;
        s_cnt[r * kThreads + tid] = ((lo_start + (ex & 0xFFFFu)) & 0xFFFFu) |
                                    ((hi_start + (ex >> 16)) << 16);
    }

#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (first + (uint32_t)k < nvalid) {
            const uint32_t d = dg[k];
            const uint32_t v = s_cnt[(d >> 1) * kThreads + tid];
            const uint32_t pos = ((v >> ((d & 1u) << 4)) & 0xFFFFu) + lr[k];
            if (pos < nvalid) {  // always true when counts are consistent
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
// Postscan: stable in-tile reorder + write-out.
// ---------------------------------------------------------------------------
template <int LOG_B, int K, int LEVEL_BITS, bool KV, class F>
__global__ __launch_bounds__(kThreads) void ms_postscan_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, uint32_t num_tiles,
    const uint32_t* __restrict__ G, F f, int vec_ok) {
    constexpr int B = 1 << LOG_B;
    constexpr uint32_t TILE = kThreads * K;
    constexpr int NLEV = (LOG_B + LEVEL_BITS - 1) / LEVEL_BITS;
    constexpr int RW = 1 << (LEVEL_BITS - 1);
    static_assert(NLEV >= 1 && NLEV <= 3, "at most three ranking passes");
    static_assert(TILE <= 32768, "packed 16-bit positions");
    static_assert(K % 4 == 0, "vector loads");

    __shared__ __attribute__((aligned(16))) uint32_t s_keys[TILE];
    __shared__ __attribute__((aligned(16))) uint32_t s_vals[KV ? TILE : 4];
    __shared__ uint32_t s_cnt[RW * kThreads];
    __shared__ uint32_t s_wsum[kWaves * RW];
    __shared__ uint32_t s_delta[B];
    __shared__ uint32_t s_bs[kWaves];

    const uint32_t tid = threadIdx.x;
    const uint32_t tile = xcd_tile(blockIdx.x, num_tiles);
    const size_t base = (size_t)tile * TILE;
    const uint32_t nvalid = (uint32_t)(((size_t)n - base) < TILE ? ((size_t)n - base) : TILE);
    const bool vec = vec_ok && nvalid == TILE;

    // Global start of each of this tile's bucket runs, and the tile's count
    // per bucket as the difference of consecutive scanned entries.
    uint32_t g0 = 0, cntb = 0;
    if (tid < (uint32_t)B) {
        const size_t idx = (size_t)tid * num_tiles + tile;
        g0 = G[idx];
        const uint32_t g1 = (idx + 1 < (size_t)B * num_tiles) ? G[idx + 1] : n;
        cntb = g1 - g0;
    }

    uint32_t key[K], val[K];
    load_blocked<K>(kin, base, nvalid, key, vec);
    if constexpr (KV) load_blocked<K>(vin, base, nvalid, val, vec);
    else {
#pragma unroll
        for (int k = 0; k < K; ++k) val[k] = 0;
    }

    {
        uint32_t tot;
        const uint32_t ls = block_excl_scan<kWaves>(cntb, s_bs, tot);
        if (tid < (uint32_t)B) s_delta[tid] = g0 - ls;
    }

    constexpr int W0 = LOG_B < LEVEL_BITS ? LOG_B : LEVEL_BITS;
    rank_pass<0, W0, K, KV>(key, val, nvalid, f, s_keys, s_vals, s_cnt, s_wsum);
    if constexpr (NLEV >= 2) {
        constexpr int S1 = LEVEL_BITS;
        constexpr int W1 = (LOG_B - S1) < LEVEL_BITS ? (LOG_B - S1) : LEVEL_BITS;
        reload_blocked<K, KV>(key, val, s_keys, s_vals);
        rank_pass<S1, W1, K, KV>(key, val, nvalid, f, s_keys, s_vals, s_cnt, s_wsum);
    }
    if constexpr (NLEV >= 3) {
        constexpr int S2 = 2 * LEVEL_BITS;
        constexpr int W2 = LOG_B - S2;
        reload_blocked<K, KV>(key, val, s_keys, s_vals);
        rank_pass<S2, W2, K, KV>(key, val, nvalid, f, s_keys, s_vals, s_cnt, s_wsum);
    }

    // Write-out in slot order: consecutive threads write consecutive slots,
    // which are consecutive addresses inside each bucket run.
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t p = (uint32_t)j * kThreads + tid;
        if (p < nvalid) {
            const uint32_t k = s_keys[p];
            const uint32_t gp = s_delta[f(k)] + p;
            if (gp < n) {  // always true when the scanned matrix is consistent
                kout[gp] = k;
                if constexpr (KV) vout[gp] = s_vals[p];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of u32 (reduce, scan of partials, downsweep).
// ---------------------------------------------------------------------------
constexpr int kScanItems = 16;
constexpr uint32_t kScanChunk = kThreads * kScanItems;  // 4096 elements per workgroup

__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                               uint32_t m,
                                                               uint32_t* __restrict__ part);
__global__ __launch_bounds__(1024) void scan_partials_kernel(uint32_t* __restrict__ part,
                                                             uint32_t np);
__global__ __launch_bounds__(kThreads) void scan_down_kernel(const uint32_t* in, uint32_t* out,
                                                             uint32_t m,
                                                             const uint32_t* __restrict__ part);
__global__ void gather_offsets_kernel(const uint32_t* __restrict__ G, uint32_t num_tiles,
                                      uint32_t num_buckets, uint32_t* __restrict__ out);
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, uint32_t n);
__global__ __launch_bounds__(kThreads) void copy_u32_kernel(const uint32_t* __restrict__ src,
                                                            uint32_t* __restrict__ dst, size_t n);

}  // namespace gms
