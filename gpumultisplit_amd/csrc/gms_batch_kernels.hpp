// Batched multisplit: many independent multisplits (instances) in one
// prescan launch, one scan and one postscan launch per (bucket class,
// key-value) group, whatever the number of instances.
//
// Work is a list of (instance, tile) items, processed either one item per
// workgroup or by a resident grid pulling items from per-XCD work queues
// (see xcd_range below).
//
// The histogram matrices of all instances are concatenated (instance i's
// 2^LOG_B x tiles_i block at mbase_i) and scanned as ONE array.  The scan of
// the concatenation at (i, b, t) is obase_i + (scan of instance i alone),
// where obase_i = keys of all earlier instances, so one plain device-wide
// scan serves every instance; the element after instance i's block is
// obase_i + n_i, which closes its last bucket run.
#pragma once

#include "gms_kernels.hpp"

namespace gms {

// One instance of a batch (device table, read with scalar loads).
struct BatchInst {
    const uint32_t* kin;
    const uint32_t* vin;  // NULL in key-only groups
    uint32_t* kout;
    uint32_t* vout;
    uint32_t* offs;  // B bucket starts (device) or NULL
    uint32_t n;
    uint32_t num_tiles;
    uint32_t mbase;  // first word of this instance's matrix block
    uint32_t obase;  // keys of all earlier instances of the batch
    uint32_t nb;     // bucket count B
    AnyF f;
};

struct BatchWork {
    uint32_t inst;
    uint32_t tile;
};

// One instance of elements with a 64-bit key and/or value (KW key words, VW
// value words per element; the group fixes both).  The identifier is f64 for
// 64-bit keys and f for 32-bit ones.
struct BatchWideInst {
    const uint32_t* kin;
    const uint32_t* vin;  // NULL without values
    uint32_t* kout;
    uint32_t* vout;
    uint32_t* offs;
    uint32_t n;
    uint32_t num_tiles;
    uint32_t mbase;
    uint32_t obase;
    uint32_t nb;
    AnyF f;
    AnyF64 f64;
};

// Launch of one 64-bit group (gms_batch_wide.hip), one workgroup per item.
struct BatchWideLaunch {
    const BatchWideInst* inst;
    const BatchWork* work;  // this group's items
    uint32_t num_work;
    uint32_t* H;            // whole concatenated matrix
    uint32_t m_total;
    uint32_t keys_total;
    int log_b;
    uint32_t kw, vw;
    bool vec;
    hipStream_t stream;
};
hipError_t launch_batch_wide(const BatchWideLaunch& c, bool post);

// Work order.  Workgroups go to the 8 XCDs round-robin (workgroup i on XCD
// i mod 8).  Every XCD owns one contiguous range of work items (the split of
// xcd_tile), so neighbouring tiles of an instance -- which share cache lines
// of the histogram matrix -- meet in the same L2.
//  * one workgroup per item: item = xcd_tile(blockIdx.x), no atomics;
//  * resident grid: the workgroups of XCD x take the items of its range in
//    order, first blockIdx.x / 8, then from a counter of that XCD
//    (wg_x + atomicAdd), until the range is exhausted -- every workgroup
//    reaches that exit.
struct XcdRange {
    uint32_t begin, count, wgs;  // items [begin, begin+count), workgroups on this XCD
};

__device__ __forceinline__ XcdRange xcd_range(uint32_t num_work) {
    constexpr uint32_t NX = 8;
    const uint32_t x = blockIdx.x % NX;
    const uint32_t q = num_work / NX, r = num_work % NX;
    const uint32_t g = gridDim.x;
    return XcdRange{x * q + (x < r ? x : r), q + (x < r ? 1u : 0u), g / NX + (x < g % NX ? 1u : 0u)};
}

// Next local item of this workgroup's XCD range (uniform); the barrier also
// separates the previous tile's last LDS reads from the next tile's writes.
__device__ __forceinline__ uint32_t xcd_next(uint32_t* queue, uint32_t wgs, uint32_t* s_next) {
    if (threadIdx.x == 0) *s_next = wgs + atomicAdd(queue + blockIdx.x % 8u, 1u);
    __syncthreads();
    return (uint32_t)GMS_READFIRSTLANE((int)*s_next);
}

template <int LOG_B, int K, bool VEC, bool STREAM>
__device__ __forceinline__ void batch_prescan_item(const BatchInst* __restrict__ inst,
                                                   const BatchWork* __restrict__ work, uint32_t w,
                                                   uint32_t* __restrict__ H, uint32_t* s_h) {
    constexpr uint32_t TILE = kThreads * K;
    const BatchWork it = work[w];
    const BatchInst& I = inst[it.inst];
    const uint32_t nvalid = tile_valid(I.n, it.tile, TILE);
    const size_t base = (size_t)it.tile * TILE;
    uint32_t* Hi = H + I.mbase;
    uint32_t x[K];
    if (nvalid == TILE) {
        load_striped4_full<K, VEC, STREAM>(I.kin, base, x);
        prescan_tile<LOG_B, K, true>(x, TILE, it.tile, I.num_tiles, Hi, I.f, s_h);
    } else {
        load_striped4_guarded<K>(I.kin, base, nvalid, x);
        prescan_tile<LOG_B, K, false>(x, nvalid, it.tile, I.num_tiles, Hi, I.f, s_h);
    }
}

// STREAM: streamed key loads, for batches larger than the caches (whose
// postscans cannot find the keys still cached).
template <int LOG_B, int K, bool VEC, bool PERSIST, bool STREAM = false>
__global__ __launch_bounds__(kThreads) void ms_batch_prescan_kernel(
    const BatchInst* __restrict__ inst, const BatchWork* __restrict__ work, uint32_t num_work,
    uint32_t* __restrict__ H, uint32_t* __restrict__ queue) {
    constexpr int B = 1 << LOG_B;
    __shared__ __attribute__((aligned(16))) uint32_t s_h[kWaves * B];
    if constexpr (!PERSIST) {
        batch_prescan_item<LOG_B, K, VEC, STREAM>(inst, work, xcd_tile(blockIdx.x, num_work), H,
                                                  s_h);
    } else {
        __shared__ uint32_t s_next;
        const XcdRange R = xcd_range(num_work);
        for (uint32_t i = blockIdx.x / 8u; i < R.count; i = xcd_next(queue, R.wgs, &s_next))
            batch_prescan_item<LOG_B, K, VEC, STREAM>(inst, work, R.begin + i, H, s_h);
    }
}

template <int LOG_B, int K, int LEVEL_BITS, bool KV, bool VEC, bool ALIAS>
__device__ __forceinline__ void batch_postscan_item(
    const BatchInst* __restrict__ inst, const BatchWork* __restrict__ work, uint32_t w,
    const uint32_t* __restrict__ G, uint32_t m_total, uint32_t keys_total, uint32_t* s_keys,
    uint32_t* s_vals, uint32_t* s_cnt, uint32_t* s_delta, uint32_t* s_bs, uint32_t* s_ws) {
    constexpr int B = 1 << LOG_B;
    constexpr uint32_t TILE = kThreads * K;
    const uint32_t tid = threadIdx.x;
    const BatchWork it = work[w];
    const BatchInst& I = inst[it.inst];
    const uint32_t t = it.tile;
    const uint32_t nvalid = tile_valid(I.n, t, TILE);
    const size_t base = (size_t)t * TILE;

    // G[b][t] of this instance and the next word in scan order; the word
    // after the whole concatenation is the batch's key total.
    const uint32_t bt = tid < (uint32_t)B ? tid : (uint32_t)B - 1u;
    const uint32_t idx = I.mbase + bt * I.num_tiles + t;
    const uint32_t g0 = G[idx];
    const uint32_t g1 = idx + 1u < m_total ? G[idx + 1u] : keys_total;
    const uint32_t cntb = g1 - g0;
    const uint32_t g0l = g0 - I.obase;  // start inside this instance's output
    if (t == 0 && I.offs && tid < I.nb) I.offs[tid] = g0l;
    const uint32_t g0a[1] = {g0l}, cnta[1] = {tid < (uint32_t)B ? cntb : 0u};

    uint32_t key[K], val[K];
    if (nvalid == TILE) {
        load_blocked_full<K, VEC>(I.kin, base, key);
        if constexpr (KV) load_blocked_full<K, VEC>(I.vin, base, val);
        else {
#pragma unroll
            for (int k = 0; k < K; ++k) val[k] = 0;
        }
        postscan_rank<LOG_B, K, LEVEL_BITS, KV, true, ALIAS>(key, val, g0a, cnta, TILE, I.f, s_keys,
                                                      s_vals, s_cnt, s_delta, s_bs, s_ws);
        postscan_write<K, KV, true>(TILE, I.n, I.kout, I.vout, I.f, s_keys, s_vals, s_delta);
    } else {
        load_blocked_guarded<K>(I.kin, base, nvalid, key);
        if constexpr (KV) load_blocked_guarded<K>(I.vin, base, nvalid, val);
        else {
#pragma unroll
            for (int k = 0; k < K; ++k) val[k] = 0;
        }
        postscan_rank<LOG_B, K, LEVEL_BITS, KV, false, ALIAS>(key, val, g0a, cnta, nvalid, I.f, s_keys,
                                                       s_vals, s_cnt, s_delta, s_bs, s_ws);
        postscan_write<K, KV, false>(nvalid, I.n, I.kout, I.vout, I.f, s_keys, s_vals, s_delta);
    }
}

template <int LOG_B, int K, int LEVEL_BITS, bool KV, bool VEC, bool PERSIST>
__global__ __launch_bounds__(kThreads) void ms_batch_postscan_kernel(
    const BatchInst* __restrict__ inst, const BatchWork* __restrict__ work, uint32_t num_work,
    const uint32_t* __restrict__ G, uint32_t m_total, uint32_t keys_total,
    uint32_t* __restrict__ queue) {
    constexpr int B = 1 << LOG_B;
    constexpr uint32_t TILE = kThreads * K;
    constexpr int NLEV = (LOG_B + LEVEL_BITS - 1) / LEVEL_BITS;
    static_assert(NLEV >= 1 && NLEV <= 3, "at most three ranking passes");
    static_assert(K % 4 == 0, "vector loads");

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

    if constexpr (!PERSIST) {
        batch_postscan_item<LOG_B, K, LEVEL_BITS, KV, VEC, ALIAS>(
            inst, work, xcd_tile(blockIdx.x, num_work), G, m_total, keys_total, s_keys, s_vals,
            s_cnt, s_delta, s_bs, s_ws);
    } else {
        __shared__ uint32_t s_next;
        const XcdRange R = xcd_range(num_work);
        for (uint32_t i = blockIdx.x / 8u; i < R.count; i = xcd_next(queue, R.wgs, &s_next))
            batch_postscan_item<LOG_B, K, LEVEL_BITS, KV, VEC, ALIAS>(inst, work, R.begin + i, G,
                                                                m_total, keys_total, s_keys,
                                                                s_vals, s_cnt, s_delta, s_bs,
                                                                s_ws);
    }
}

// Launchers, one translation unit per tile shape (gms_batch_k*.hip).
struct BatchLaunch {
    const BatchInst* inst;
    const BatchWork* work;  // this group's items
    uint32_t num_work;
    uint32_t grid;          // workgroups (num_work, or the resident grid)
    uint32_t* H;            // whole concatenated matrix
    uint32_t m_total;
    uint32_t keys_total;
    uint32_t* queue;        // this launch's 8 per-XCD counters (zero on entry)
    int log_b;
    bool kv;
    bool vec;               // every instance of the group 16-byte aligned
    bool persistent;
    hipStream_t stream;
    bool stream_loads = false;  // the step is larger than the caches: prescan loads stream
};

hipError_t launch_batch_prescan_k8(const BatchLaunch& c);
hipError_t launch_batch_prescan_k16(const BatchLaunch& c);
hipError_t launch_batch_prescan_k24(const BatchLaunch& c);
hipError_t launch_batch_prescan_k32(const BatchLaunch& c);
hipError_t launch_batch_postscan_k8(const BatchLaunch& c);
hipError_t launch_batch_postscan_k16(const BatchLaunch& c);
hipError_t launch_batch_postscan_k24(const BatchLaunch& c);
hipError_t launch_batch_postscan_k32(const BatchLaunch& c);
// Resident workgroups per CU of the persistent batch kernels.
int batch_occupancy_k8(int log_b, bool kv, bool postscan);
int batch_occupancy_k16(int log_b, bool kv, bool postscan);
int batch_occupancy_k24(int log_b, bool kv, bool postscan);
int batch_occupancy_k32(int log_b, bool kv, bool postscan);

template <int K>
struct BatchKernels {
    template <int LOG_B, bool VEC, bool P, bool S = false>
    static hipError_t prescan_v(const BatchLaunch& c) {
        hipLaunchKernelGGL((ms_batch_prescan_kernel<LOG_B, K, VEC, P, S>), dim3(c.grid),
                           dim3(kThreads), 0, c.stream, c.inst, c.work, c.num_work, c.H, c.queue);
        return hipGetLastError();
    }
    template <int LOG_B>
    static hipError_t prescan(const BatchLaunch& c) {
        if (c.persistent) {
            if (c.vec && c.stream_loads) return prescan_v<LOG_B, true, true, true>(c);
            return c.vec ? prescan_v<LOG_B, true, true>(c) : prescan_v<LOG_B, false, true>(c);
        }
        if (c.vec && c.stream_loads) return prescan_v<LOG_B, true, false, true>(c);
        return c.vec ? prescan_v<LOG_B, true, false>(c) : prescan_v<LOG_B, false, false>(c);
    }
    template <int LOG_B, bool KV, bool VEC, bool P>
    static hipError_t postscan_v(const BatchLaunch& c) {
        constexpr int LB = tune_level_bits(LOG_B);
        hipLaunchKernelGGL((ms_batch_postscan_kernel<LOG_B, K, LB, KV, VEC, P>), dim3(c.grid),
                           dim3(kThreads), 0, c.stream, c.inst, c.work, c.num_work,
                           (const uint32_t*)c.H, c.m_total, c.keys_total, c.queue);
        return hipGetLastError();
    }
    template <int LOG_B, bool KV>
    static hipError_t postscan(const BatchLaunch& c) {
        if (c.persistent)
            return c.vec ? postscan_v<LOG_B, KV, true, true>(c) : postscan_v<LOG_B, KV, false, true>(c);
        return c.vec ? postscan_v<LOG_B, KV, true, false>(c) : postscan_v<LOG_B, KV, false, false>(c);
    }
    template <int LOG_B, bool KV>
    static int occupancy(bool post) {
        int occ = 0;
        hipError_t e;
        if (post) {
            constexpr int LB = tune_level_bits(LOG_B);
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, ms_batch_postscan_kernel<LOG_B, K, LB, KV, true, true>, kThreads, 0);
        } else {
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, ms_batch_prescan_kernel<LOG_B, K, true, true>, kThreads, 0);
        }
        return (e == hipSuccess && occ > 0) ? occ : 1;
    }
};

#define GMS_BATCH_SWITCH(FN, ARGS, log_b, kv)      \
    switch ((log_b) * 2 + ((kv) ? 1 : 0)) {         \
        case 2: return FN<1, false> ARGS;           \
        case 3: return FN<1, true> ARGS;            \
        case 4: return FN<2, false> ARGS;           \
        case 5: return FN<2, true> ARGS;            \
        case 6: return FN<3, false> ARGS;           \
        case 7: return FN<3, true> ARGS;            \
        case 8: return FN<4, false> ARGS;           \
        case 9: return FN<4, true> ARGS;            \
        case 10: return FN<5, false> ARGS;          \
        case 11: return FN<5, true> ARGS;           \
        case 12: return FN<6, false> ARGS;          \
        case 13: return FN<6, true> ARGS;           \
        case 14: return FN<7, false> ARGS;          \
        case 15: return FN<7, true> ARGS;           \
        case 16: return FN<8, false> ARGS;          \
        case 17: return FN<8, true> ARGS;           \
        default: break;                             \
    }

#define GMS_DEFINE_BATCH_LAUNCHERS(KK)                                                   \
    hipError_t launch_batch_prescan_k##KK(const BatchLaunch& c) {                        \
        switch (c.log_b) {                                                               \
            case 1: return BatchKernels<KK>::prescan<1>(c);                              \
            case 2: return BatchKernels<KK>::prescan<2>(c);                              \
            case 3: return BatchKernels<KK>::prescan<3>(c);                              \
            case 4: return BatchKernels<KK>::prescan<4>(c);                              \
            case 5: return BatchKernels<KK>::prescan<5>(c);                              \
            case 6: return BatchKernels<KK>::prescan<6>(c);                              \
            case 7: return BatchKernels<KK>::prescan<7>(c);                              \
            case 8: return BatchKernels<KK>::prescan<8>(c);                              \
            default: return hipErrorInvalidValue;                                        \
        }                                                                                \
    }                                                                                    \
    hipError_t launch_batch_postscan_k##KK(const BatchLaunch& c) {                       \
        GMS_BATCH_SWITCH(BatchKernels<KK>::postscan, (c), c.log_b, c.kv)                      \
        return hipErrorInvalidValue;                                                     \
    }                                                                                    \
    int batch_occupancy_k##KK(int log_b, bool kv, bool post) {                           \
        GMS_BATCH_SWITCH(BatchKernels<KK>::occupancy, (post), log_b, kv)                         \
        return 1;                                                                        \
    }

}  // namespace gms
