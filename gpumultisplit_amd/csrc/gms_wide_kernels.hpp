// Kernels of the multisplit of elements with 64-bit keys and/or 64-bit values
// (the host engine is gms_wide.hip).
//
// The reference's API is templated on KeyT and ValueT (src/api/multisplit.cuh:
// 81-82, :199-200; src/api/multisplit_sort.cuh:366-527;
// multisplit_sort_pairs.cuh:67-228) although its kernels move 32-bit words.
// Here an element is KW key words and VW value words (KW, VW in {1, 2}, 0 <=
// VW), stored as the caller's arrays of 32- or 64-bit integers.
//
// Same three stages as the 32-bit path (prescan, device scan, postscan) on the
// same histogram matrix layout.  The postscan does not move the element
// through its LDS ranking passes: it ranks one packed word per element,
// bucket << 16 | position in the tile, with the 32-bit ranking code
// (postscan_rank), while the element's words wait in LDS in input order in
// the area the ranking counters used.  The write-out reads the packed word of
// each slot and fetches the element's words through its position, so a
// 16-byte element costs the ranking no more LDS traffic than a 4-byte key.
#pragma once

#include "gms_internal.hpp"
#include "gms_kernels.hpp"

namespace gms {

// Elements per thread of a tile: TILE * (KW + VW) words of payload share LDS
// with the packed words, so wider elements take shorter tiles: 2 words per
// element K = 20 (60 KB, 2 workgroups per CU; 256-bucket key-only 64-bit keys
// at 2^27: 0.761 ms against 0.773 at K = 16 with 3 per CU, 0.760-0.766 at 24),
// 3 words K = 12 (3 per CU), 4 words K = 12 (2 per CU; K = 16 would leave one).
// K is a multiple of 4 (16-byte LDS row reads).
constexpr int wide_items(int kw, int vw) { return kw + vw <= 2 ? 20 : 12; }

// Packed word -> bucket id (the ranking's identifier).
struct PackedBidF {
    GMS_HD uint32_t operator()(uint32_t w) const { return w >> 16; }
};

template <int KW>
__device__ __forceinline__ uint64_t key_word_of(const uint32_t* w, int k) {
    if constexpr (KW == 2) return (uint64_t)w[2 * k] | ((uint64_t)w[2 * k + 1] << 32);
    else return w[k];
}

// Words of the thread's K consecutive elements (blocked: thread t owns
// elements [t*K, t*K+K) of the tile).  Unguarded 16-byte loads for full tiles
// (K*NW is a multiple of 4 and the caller checked the alignment).
template <int NW, int K, bool VEC, bool FULL>
__device__ __forceinline__ void load_elems(const uint32_t* __restrict__ p, size_t base,
                                           uint32_t nvalid, uint32_t (&x)[K * NW]) {
    const uint32_t e0 = threadIdx.x * (uint32_t)K;
    const uint32_t* q = p + (base + e0) * NW;
    if constexpr (FULL && VEC) {
        static_assert((K * NW) % 4 == 0, "vector loads");
#pragma unroll
        for (int i = 0; i < K * NW / 4; ++i) {
            const uint4 v = reinterpret_cast<const uint4*>(q)[i];
            x[4 * i + 0] = v.x; x[4 * i + 1] = v.y; x[4 * i + 2] = v.z; x[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int w = 0; w < NW; ++w)
                x[k * NW + w] = (FULL || e0 + (uint32_t)k < nvalid) ? q[k * NW + w] : 0u;
    }
}

// Prescan of tile t: the tile's key words are read as one flat array in
// 16-byte stripes (the 32-bit prescan's lane-contiguous loads); word
// 4*(j*256+t)+i belongs to element (4*(j*256+t)+i) / KW, and a 64-bit key
// never straddles a stripe.  Writes column t of the num_tiles-column matrix H.
// STREAM: streamed loads, for key arrays larger than the caches (as in the
// 32-bit prescan).
template <int LOG_B, int K, int KW, bool VEC, bool STREAM = false, class F>
__device__ __forceinline__ void wide_prescan_tile(const uint32_t* __restrict__ kin, uint32_t n,
                                                  uint32_t t, uint32_t num_tiles,
                                                  uint32_t* __restrict__ H, F f, uint32_t* s_h) {
    constexpr int B = 1 << LOG_B;
    constexpr int WPT = K * KW;  // key words per thread
    constexpr uint32_t TILE = kThreads * K;
    static_assert(WPT % 4 == 0, "16-byte stripes");
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t nvalid = tile_valid(n, t, TILE);
    const size_t base = (size_t)t * TILE * KW;
    uint32_t x[WPT];
    if (nvalid == TILE) load_striped4_full<WPT, VEC, STREAM>(kin, base, x);
    else load_striped4_guarded<WPT>(kin, base, nvalid * KW, x);
    uint32_t* my = s_h + wid * B;
    for (uint32_t i = lane; i < (uint32_t)B; i += 64) my[i] = 0;
    GMS_WAVE_SYNC();
    // Probed on the wave's first element slot (wave_ids_coincide).
    const bool coincide = wave_ids_coincide(f(key_word_of<KW>(x, 0)), (4u * tid) / KW < nvalid);
#pragma unroll
    for (int j = 0; j < WPT / 4; ++j) {
        const uint32_t e4 = (4u * ((uint32_t)j * kThreads + tid)) / KW;  // first element of the stripe
#pragma unroll
        for (int q = 0; q < 4 / KW; ++q) {
            const uint32_t d = f(key_word_of<KW>(x + 4 * j, q));
            const bool v = e4 + (uint32_t)q < nvalid;
            if (coincide) wave_count_add(my, d, v);
            else if (v) atomicAdd(&my[d], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = tid; b < (uint32_t)B; b += kThreads) {
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q) s += s_h[q * B + b];
        H[(size_t)b * num_tiles + t] = s;
    }
}

// `clear` (nclear 64-bit words, or NULL): the status words of the
// single-pass scan that follows, zeroed by workgroup 0.
template <int LOG_B, int K, int KW, bool VEC, bool STREAM, class F>
__global__ __launch_bounds__(kThreads) void ms_wide_prescan_kernel(const uint32_t* __restrict__ kin,
                                                                   uint32_t n, uint32_t num_tiles,
                                                                   uint32_t* __restrict__ H, F f,
                                                                   uint64_t* __restrict__ clear,
                                                                   uint32_t nclear) {
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < nclear; i += kThreads) store_device_u64(&clear[i], 0ull);
    __shared__ __attribute__((aligned(16))) uint32_t s_h[kWaves * (1 << LOG_B)];
    wide_prescan_tile<LOG_B, K, KW, VEC, STREAM>(kin, n, xcd_tile(blockIdx.x, num_tiles),
                                                 num_tiles, H, f, s_h);
}

template <int LOG_B, int K, int LEVEL_BITS, int KW, int VW, bool FULL, class F>
__device__ __forceinline__ void wide_postscan_tile(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, size_t base,
    uint32_t nvalid, uint32_t g0, uint32_t cntb, F f, uint32_t* s_keys, uint32_t* s_u,
    uint32_t* s_delta, uint32_t* s_bs, uint32_t* s_ws, const uint32_t (&kx)[K * KW],
    const uint32_t (&vx)[K * (VW ? VW : 1)]) {
    constexpr int NW = KW + VW;
    constexpr uint32_t TILE = kThreads * K;
    const uint32_t tid = threadIdx.x;
    const uint32_t e0 = tid * (uint32_t)K;
    uint32_t pk[K], unused[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        pk[k] = (f(key_word_of<KW>(kx, k)) << 16) | (e0 + (uint32_t)k);
        unused[k] = 0;
    }
    constexpr uint32_t B = 1u << LOG_B;
    const uint32_t g0a[1] = {g0}, cnta[1] = {tid < B ? cntb : 0u};
    postscan_rank<LOG_B, K, LEVEL_BITS, false, FULL, false>(pk, unused, g0a, cnta, nvalid,
                                                            PackedBidF{}, s_keys, nullptr, s_u,
                                                            s_delta, s_bs, s_ws);
    // The ranking ended with a barrier: its counters in s_u are dead.
#pragma unroll
    for (int k = 0; k < K; ++k) {
        uint32_t* e = s_u + (e0 + (uint32_t)k) * NW;
#pragma unroll
        for (int w = 0; w < KW; ++w) e[w] = kx[k * KW + w];
#pragma unroll
        for (int w = 0; w < VW; ++w) e[KW + w] = vx[k * VW + w];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t p = (uint32_t)j * kThreads + tid;
        if (FULL || p < nvalid) {
            const uint32_t pw = s_keys[p];
            uint32_t idx = pw & 0xFFFFu;
            idx = idx < TILE ? idx : TILE - 1u;
            uint32_t gp = s_delta[pw >> 16] + p;
            gp = gp < n ? gp : n - 1u;
            const uint32_t* e = s_u + idx * NW;
            if constexpr (KW == 2)
                *reinterpret_cast<uint2*>(kout + 2 * (size_t)gp) = make_uint2(e[0], e[1]);
            else
                kout[gp] = e[0];
            if constexpr (VW == 2)
                *reinterpret_cast<uint2*>(vout + 2 * (size_t)gp) = make_uint2(e[KW], e[KW + 1]);
            else if constexpr (VW == 1)
                vout[gp] = e[KW];
        }
    }
}

// LDS of the 64-bit postscan: packed words, then the element words (which
// reuse the ranking counters' area), bucket deltas and scan scratch.
template <int LOG_B, int K, int LEVEL_BITS, int KW, int VW>
struct WideSmem {
    static constexpr uint32_t TILE = kThreads * K;
    static constexpr uint32_t CNT = cnt_words(LEVEL_BITS);
    static constexpr uint32_t PAY = TILE * (KW + VW);
    static_assert(TILE <= 0x10000u && LOG_B <= 8, "packed word: 8-bit bucket, 16-bit position");
    uint32_t keys[TILE];
    uint32_t u[PAY > CNT ? PAY : CNT];
    uint32_t delta[1 << LOG_B];
    uint32_t bs[kWaves];
    uint32_t ws[kWaves];
};

// Postscan of tile t of an input of n elements, whose bucket b starts at
// output element g0 (thread b's value) and holds cntb of the tile's elements.
template <int LOG_B, int K, int LEVEL_BITS, int KW, int VW, bool VEC, class F>
__device__ __forceinline__ void wide_postscan_item(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, uint32_t t,
    uint32_t g0, uint32_t cntb, F f, WideSmem<LOG_B, K, LEVEL_BITS, KW, VW>& sm) {
    constexpr uint32_t TILE = kThreads * K;
    const uint32_t nvalid = tile_valid(n, t, TILE);
    const size_t base = (size_t)t * TILE;
    uint32_t kx[K * KW], vx[K * (VW ? VW : 1)];
    if (nvalid == TILE) {
        load_elems<KW, K, VEC, true>(kin, base, TILE, kx);
        if constexpr (VW) load_elems<VW, K, VEC, true>(vin, base, TILE, vx);
        wide_postscan_tile<LOG_B, K, LEVEL_BITS, KW, VW, true>(kin, vin, kout, vout, n, base, TILE,
                                                              g0, cntb, f, sm.keys, sm.u, sm.delta,
                                                              sm.bs, sm.ws, kx, vx);
    } else {
        load_elems<KW, K, false, false>(kin, base, nvalid, kx);
        if constexpr (VW) load_elems<VW, K, false, false>(vin, base, nvalid, vx);
        wide_postscan_tile<LOG_B, K, LEVEL_BITS, KW, VW, false>(kin, vin, kout, vout, n, base,
                                                               nvalid, g0, cntb, f, sm.keys, sm.u,
                                                               sm.delta, sm.bs, sm.ws, kx, vx);
    }
}

template <int LOG_B, int K, int LEVEL_BITS, int KW, int VW, bool VEC, class F>
__global__ __launch_bounds__(kThreads) void ms_wide_postscan_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, uint32_t num_tiles,
    const uint32_t* __restrict__ G, F f) {
    constexpr int B = 1 << LOG_B;
    __shared__ __attribute__((aligned(16))) WideSmem<LOG_B, K, LEVEL_BITS, KW, VW> sm;
    const uint32_t tid = threadIdx.x;
    const uint32_t t = xcd_tile(blockIdx.x, num_tiles);
    // G[b][t] and the next entry in scan order (n past the end), as in the
    // 32-bit postscan: their difference is the tile's count in bucket b.
    const uint32_t bt = tid < (uint32_t)B ? tid : (uint32_t)B - 1u;
    const size_t idx = (size_t)bt * num_tiles + t;
    const size_t mtot = (size_t)B * num_tiles;
    const uint32_t g0 = G[idx];
    const uint32_t g1raw = G[idx + 1 < mtot ? idx + 1 : idx];
    const uint32_t cntb = (idx + 1 < mtot ? g1raw : n) - g0;
    wide_postscan_item<LOG_B, K, LEVEL_BITS, KW, VW, VEC>(kin, vin, kout, vout, n, t, g0, cntb, f,
                                                          sm);
}

// ---------------------------------------------------------------- launches
struct WideCall {
    const uint32_t* kin;
    const uint32_t* vin;
    uint32_t* kout;
    uint32_t* vout;
    uint32_t n;
    uint32_t num_tiles;
    uint32_t* H;
    bool vec;
    hipStream_t stream;
    // Status words of the single-pass scan that follows, cleared by the
    // prescan (NULL: the scan that follows is the two-launch one).
    uint64_t* scan_status = nullptr;
    uint32_t scan_chunks = 0;
};

template <int LOG_B, int KW, int VW, bool VEC, class F>
inline hipError_t wide_prescan_v(const WideCall& c, const F& f) {
    constexpr int K = wide_items(KW, VW);
    // Key arrays of at least kStreamLoadMin words stream their loads.
    if constexpr (VEC) {
        if ((size_t)c.n * KW >= kStreamLoadMin) {
            hipLaunchKernelGGL((ms_wide_prescan_kernel<LOG_B, K, KW, true, true, F>),
                               dim3(c.num_tiles), dim3(kThreads), 0, c.stream, c.kin, c.n,
                               c.num_tiles, c.H, f, c.scan_status, c.scan_chunks);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((ms_wide_prescan_kernel<LOG_B, K, KW, VEC, false, F>), dim3(c.num_tiles),
                       dim3(kThreads), 0, c.stream, c.kin, c.n, c.num_tiles, c.H, f,
                       c.scan_status, c.scan_chunks);
    return hipGetLastError();
}

template <int LOG_B, int KW, int VW, bool VEC, class F>
inline hipError_t wide_postscan_v(const WideCall& c, const F& f) {
    constexpr int K = wide_items(KW, VW);
    constexpr int LB = tune_level_bits(LOG_B);
    hipLaunchKernelGGL((ms_wide_postscan_kernel<LOG_B, K, LB, KW, VW, VEC, F>), dim3(c.num_tiles),
                       dim3(kThreads), 0, c.stream, c.kin, c.vin, c.kout, c.vout, c.n, c.num_tiles,
                       c.H, f);
    return hipGetLastError();
}

template <int LOG_B, int KW, int VW, class F>
inline hipError_t wide_stages_b(const WideCall& c, const F& f, bool prescan) {
    if (prescan) return c.vec ? wide_prescan_v<LOG_B, KW, VW, true>(c, f)
                              : wide_prescan_v<LOG_B, KW, VW, false>(c, f);
    return c.vec ? wide_postscan_v<LOG_B, KW, VW, true>(c, f)
                 : wide_postscan_v<LOG_B, KW, VW, false>(c, f);
}

template <int KW, int VW, class F>
inline hipError_t wide_stage(const WideCall& c, int log_b, const F& f, bool prescan) {
    switch (log_b) {
        case 1: return wide_stages_b<1, KW, VW>(c, f, prescan);
        case 2: return wide_stages_b<2, KW, VW>(c, f, prescan);
        case 3: return wide_stages_b<3, KW, VW>(c, f, prescan);
        case 4: return wide_stages_b<4, KW, VW>(c, f, prescan);
        case 5: return wide_stages_b<5, KW, VW>(c, f, prescan);
        case 6: return wide_stages_b<6, KW, VW>(c, f, prescan);
        case 7: return wide_stages_b<7, KW, VW>(c, f, prescan);
        case 8: return wide_stages_b<8, KW, VW>(c, f, prescan);
        default: return hipErrorInvalidValue;
    }
}

// Stage launch for the element shape; mk64 / mk32 build the identifier of a
// 64-bit / 32-bit key (only the one the shape uses is built).
template <class Mk64, class Mk32>
inline hipError_t wide_dispatch(const WideCall& c, int log_b, uint32_t kw, uint32_t vw,
                                const Mk64& mk64, const Mk32& mk32, bool prescan) {
    if (kw == 2) {
        const auto f = mk64();
        switch (vw) {
            case 0: return wide_stage<2, 0>(c, log_b, f, prescan);
            case 1: return wide_stage<2, 1>(c, log_b, f, prescan);
            case 2: return wide_stage<2, 2>(c, log_b, f, prescan);
            default: return hipErrorInvalidValue;
        }
    }
    if (vw != 2) return hipErrorInvalidValue;
    return wide_stage<1, 2>(c, log_b, mk32(), prescan);
}

}  // namespace gms
