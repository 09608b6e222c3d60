// Wave-level multisplit (WMS) for up to 64 buckets: every wave of 64 lanes
// owns its own subproblem of kWmsItems x 64 consecutive keys.
//
// The reference's warp-wide variant (src/kernels/wms/wms_prescan.cuh:21-149,
// wms_postscan.cuh:20-505, wms_postscan_pairs.cuh:29-538; sizes wms_api.h:
// 25-83, driven by src/main/main_wms.cu) splits the input into per-warp
// subproblems and ranks keys with warp ballots, so no block-wide barrier is
// needed.  Here a wave64 ranks item k of its subproblem (keys k*64 + lane) at
// once: log2(B) ballots of the bucket-id bits give each lane the mask of the
// lanes with its own bucket, its rank among them is the count of that mask
// below it (v_mbcnt), and the lowest of them advances the wave's running
// count of that bucket in LDS.  With 64 lanes a wave covers 64 buckets, one
// lane per bucket for the per-wave counts and starts.
//
// Same three stages as the block-wide path on the same bucket-major matrix,
// with one column per wave subproblem; a block of four waves takes four
// consecutive subproblems in the XCD-contiguous tile order.
//
// DMS, the reference's direct variant (DMS_postscan_4rolls_protected,
// wms_postscan.cuh:328-505, and its pairs twin wms_postscan_pairs.cuh:357),
// is the same postscan without the LDS reorder: every key goes from its
// register straight to global start + running count + rank, so lanes of one
// store instruction write up to B separate runs.
#include "gms_internal.hpp"
#include "gms_kernels.hpp"

namespace gms {

constexpr uint32_t kWmsItems = 16;                       // keys per lane
constexpr uint32_t kWmsSub = 64u * kWmsItems;            // keys per wave subproblem

uint32_t wms_subproblem_keys() { return kWmsSub; }

template <int LOG_B, class F>
__global__ __launch_bounds__(kThreads) void wms_prescan_kernel(const uint32_t* __restrict__ kin,
                                                               uint32_t n, uint32_t num_sub,
                                                               uint32_t* __restrict__ H, F f) {
    constexpr uint32_t B = 1u << LOG_B;
    __shared__ uint32_t s_cnt[kThreads];  // 64 counters per wave
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t blk = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t sub = blk * (kThreads / 64) + w;
    if (sub >= num_sub) return;  // whole wave; no block-wide barrier follows
    const uint32_t nvalid = tile_valid(n, sub, kWmsSub);
    const size_t base = (size_t)sub * kWmsSub;
    uint32_t* cnt = s_cnt + w * 64u;
    cnt[lane] = 0;
    uint32_t key[kWmsItems];
#pragma unroll
    for (uint32_t k = 0; k < kWmsItems; ++k) {
        const uint32_t i = k * 64u + lane;
        key[k] = i < nvalid ? kin[base + i] : 0u;
    }
    GMS_WAVE_SYNC();
#pragma unroll
    for (uint32_t k = 0; k < kWmsItems; ++k) {
        const bool valid = k * 64u + lane < nvalid;
        const uint32_t d = f(key[k]);
        const uint64_t peers = wave_peers<LOG_B>(d, valid);
        if (valid && lanes_below(peers) == 0u) cnt[d] += (uint32_t)__builtin_popcountll(peers);
        GMS_WAVE_SYNC();
    }
    if (lane < B) H[(size_t)lane * num_sub + sub] = cnt[lane];
}

template <int LOG_B, bool KV, bool DIRECT, class F>
__global__ __launch_bounds__(kThreads) void wms_postscan_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, uint32_t num_sub,
    const uint32_t* __restrict__ G, F f) {
    constexpr uint32_t B = 1u << LOG_B;
    constexpr uint32_t NW = kThreads / 64;
    __shared__ __attribute__((aligned(16))) uint32_t s_k[DIRECT ? 1 : NW * kWmsSub];
    __shared__ __attribute__((aligned(16))) uint32_t s_v[(KV && !DIRECT) ? NW * kWmsSub : 1];
    __shared__ uint32_t s_run[kThreads];  // per wave: next local slot of every bucket
    __shared__ uint32_t s_dl[kThreads];   // per wave: global start - local start
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t blk = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t sub = blk * NW + w;
    if (sub >= num_sub) return;
    const uint32_t nvalid = tile_valid(n, sub, kWmsSub);
    const size_t base = (size_t)sub * kWmsSub;
    uint32_t* sk = s_k + w * kWmsSub;
    uint32_t* sv = KV ? s_v + w * kWmsSub : s_v;
    uint32_t* run = s_run + w * 64u;
    uint32_t* dl = s_dl + w * 64u;

    uint32_t key[kWmsItems], val[kWmsItems];
#pragma unroll
    for (uint32_t k = 0; k < kWmsItems; ++k) {
        const uint32_t i = k * 64u + lane;
        key[k] = i < nvalid ? kin[base + i] : 0u;
        val[k] = (KV && i < nvalid) ? vin[base + i] : 0u;
    }
    // Lane b < B: this subproblem's global start of bucket b and its count
    // (the next entry in scan order minus this one; n after the last).
    const size_t mtot = (size_t)B * num_sub;
    const size_t idx = (size_t)(lane < B ? lane : 0u) * num_sub + sub;
    uint32_t g0 = 0, cnt = 0;
    if (lane < B) {
        g0 = G[idx];
        cnt = (idx + 1 < mtot ? G[idx + 1] : n) - g0;
    }
    const uint32_t ls = wave_incl_scan(cnt) - cnt;  // local start of bucket `lane`
    run[lane] = ls;
    dl[lane] = g0 - ls;
    GMS_WAVE_SYNC();
#pragma unroll
    for (uint32_t k = 0; k < kWmsItems; ++k) {
        const bool valid = k * 64u + lane < nvalid;
        const uint32_t d = f(key[k]);
        const uint64_t peers = wave_peers<LOG_B>(d, valid);
        const uint32_t r = lanes_below(peers);
        if (valid) {
            if constexpr (DIRECT) {
                uint32_t gp = dl[d] + run[d] + r;
                gp = gp < n ? gp : n - 1u;  // in bounds even if the counts were not
                kout[gp] = key[k];
                if constexpr (KV) vout[gp] = val[k];
            } else {
                uint32_t pos = run[d] + r;
                pos = pos < kWmsSub ? pos : kWmsSub - 1u;  // in bounds even if the counts were not
                sk[pos] = key[k];
                if constexpr (KV) sv[pos] = val[k];
            }
        }
        GMS_WAVE_SYNC();  // every lane has read run[] before the leaders advance it
        if (valid && r == 0u) run[d] += (uint32_t)__builtin_popcountll(peers);
        GMS_WAVE_SYNC();
    }
    if constexpr (DIRECT) return;
    // Slot-order write-out: consecutive lanes, consecutive slots of a bucket run.
#pragma unroll
    for (uint32_t j = 0; j < kWmsItems; ++j) {
        const uint32_t p = j * 64u + lane;
        if (p < nvalid) {
            const uint32_t k = sk[p];
            uint32_t gp = dl[f(k)] + p;
            gp = gp < n ? gp : n - 1u;
            kout[gp] = k;
            if constexpr (KV) vout[gp] = sv[p];
        }
    }
}

// ---------------------------------------------------------------- launches
template <int LOG_B, bool KV, bool DIRECT>
static void wms_launch_post(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                            uint32_t* vout, uint32_t n, uint32_t num_sub, uint32_t* H,
                            const AnyF& f, uint32_t blocks, hipStream_t stream) {
    hipLaunchKernelGGL((wms_postscan_kernel<LOG_B, KV, DIRECT, AnyF>), dim3(blocks),
                       dim3(kThreads), 0, stream, kin, vin, kout, vout, n, num_sub, H, f);
}

template <int LOG_B>
static hipError_t wms_launch_b(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                               uint32_t* vout, uint32_t n, uint32_t num_sub, uint32_t* H,
                               const AnyF& f, bool prescan, bool direct, hipStream_t stream) {
    const uint32_t blocks = (num_sub + kThreads / 64 - 1) / (kThreads / 64);
    if (prescan)
        hipLaunchKernelGGL((wms_prescan_kernel<LOG_B, AnyF>), dim3(blocks), dim3(kThreads), 0,
                           stream, kin, n, num_sub, H, f);
    else if (vin && direct)
        wms_launch_post<LOG_B, true, true>(kin, vin, kout, vout, n, num_sub, H, f, blocks, stream);
    else if (vin)
        wms_launch_post<LOG_B, true, false>(kin, vin, kout, vout, n, num_sub, H, f, blocks, stream);
    else if (direct)
        wms_launch_post<LOG_B, false, true>(kin, vin, kout, vout, n, num_sub, H, f, blocks, stream);
    else
        wms_launch_post<LOG_B, false, false>(kin, vin, kout, vout, n, num_sub, H, f, blocks, stream);
    return hipGetLastError();
}

static hipError_t wms_launch(int log_b, const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                             uint32_t* vout, uint32_t n, uint32_t num_sub, uint32_t* H,
                             const AnyF& f, bool prescan, bool direct, hipStream_t stream) {
    switch (log_b) {
        case 1: return wms_launch_b<1>(kin, vin, kout, vout, n, num_sub, H, f, prescan, direct, stream);
        case 2: return wms_launch_b<2>(kin, vin, kout, vout, n, num_sub, H, f, prescan, direct, stream);
        case 3: return wms_launch_b<3>(kin, vin, kout, vout, n, num_sub, H, f, prescan, direct, stream);
        case 4: return wms_launch_b<4>(kin, vin, kout, vout, n, num_sub, H, f, prescan, direct, stream);
        case 5: return wms_launch_b<5>(kin, vin, kout, vout, n, num_sub, H, f, prescan, direct, stream);
        case 6: return wms_launch_b<6>(kin, vin, kout, vout, n, num_sub, H, f, prescan, direct, stream);
        default: return hipErrorInvalidValue;
    }
}

gms_status run_wms_multisplit(Scratch& s, const uint32_t* kin, const uint32_t* vin,
                              uint32_t* kout, uint32_t* vout, uint32_t n, uint32_t B,
                              gms_bucket_desc d, uint32_t* d_offsets, bool direct,
                              hipStream_t stream) {
    if (B < 2 || B > kWmsMaxBuckets) return GMS_ERR_UNSUPPORTED_BUCKETS;
    if (n == 0) return d_offsets ? hip_status(launch_fill(d_offsets, 0u, B, stream)) : GMS_SUCCESS;
    const int lb = ceil_log2_u32(B);
    const uint32_t ns = (uint32_t)(((size_t)n + kWmsSub - 1) / kWmsSub);
    const size_t m = ((size_t)1 << lb) * ns;
    if (m > s.hist_cap || scan_partials_needed(m) > s.part_cap) return GMS_ERR_CAPACITY;
    const AnyF f = AnyF::make(d.kind, d.param, d.width, B);
    hipError_t e = wms_launch(lb, kin, vin, kout, vout, n, ns, s.hist, f, true, direct, stream);
    if (e == hipSuccess) e = launch_exclusive_scan(s.hist, s.hist, (uint32_t)m, s.part, stream);
    if (e == hipSuccess)
        e = wms_launch(lb, kin, vin, kout, vout, n, ns, s.hist, f, false, direct, stream);
    if (e == hipSuccess && d_offsets) e = launch_gather_offsets(s.hist, ns, B, d_offsets, stream);
    return hip_status(e);
}

}  // namespace gms
