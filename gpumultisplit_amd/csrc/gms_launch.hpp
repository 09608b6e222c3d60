// Launch helpers instantiated once per bucket-identifier family (inst_*.hip).
#pragma once

#include "gms_internal.hpp"
#include "gms_kernels.hpp"

namespace gms {

// Inputs of at least kStreamLoadMin keys (always long-tile ones) stream
// their key loads in the prescan (STREAM).
template <int LOG_B, bool KV, bool VEC, bool LARGE, bool STREAM, class F>
hipError_t launch_prescan_v(const MsCall& c, const F& f) {
    constexpr int K = tune_items(LOG_B, KV, LARGE);
    if constexpr (LARGE && LOG_B >= 9) {
        // Groups of 8 tiles from 8192 tiles on, else of 4 (8 would leave fewer
        // groups than CUs: 1024 buckets at 2^24 key-value pairs 24 us against
        // 17.5 in groups of 4).
        if (c.num_tiles >= 8192u) {
            hipLaunchKernelGGL((ms_prescan_group_kernel<LOG_B, K, VEC, STREAM, 8, F>),
                               dim3((c.num_tiles + 7) / 8), dim3(kThreads), 0, c.stream, c.kin,
                               c.n, c.num_tiles, c.H, f, c.scan_status, c.scan_chunks);
        } else {
            hipLaunchKernelGGL((ms_prescan_group_kernel<LOG_B, K, VEC, STREAM, 4, F>),
                               dim3((c.num_tiles + 3) / 4), dim3(kThreads), 0, c.stream, c.kin,
                               c.n, c.num_tiles, c.H, f, c.scan_status, c.scan_chunks);
        }
    } else {
        hipLaunchKernelGGL((ms_prescan_kernel<LOG_B, K, VEC, STREAM, F>), dim3(c.num_tiles),
                           dim3(kThreads), 0, c.stream, c.kin, c.n, c.num_tiles, c.H, f,
                           c.scan_status, c.scan_chunks);
    }
    return hipGetLastError();
}

template <int LOG_B, bool KV, class F>
hipError_t launch_prescan_t(const MsCall& c, const F& f) {
    static_assert(kStreamLoadMin >= kLongTileMin, "streamed inputs take the long tiles");
    if (c.large) {
        if constexpr (long_tile_min(LOG_B, KV) < kStreamLoadMin) {
            if (c.n < kStreamLoadMin)
                return c.vec_ok ? launch_prescan_v<LOG_B, KV, true, true, false>(c, f)
                                : launch_prescan_v<LOG_B, KV, false, true, false>(c, f);
        }
        return c.vec_ok ? launch_prescan_v<LOG_B, KV, true, true, true>(c, f)
                        : launch_prescan_v<LOG_B, KV, false, true, true>(c, f);
    }
    return c.vec_ok ? launch_prescan_v<LOG_B, KV, true, false, false>(c, f)
                    : launch_prescan_v<LOG_B, KV, false, false, false>(c, f);
}

template <int LOG_B, bool KV, bool VEC, bool LARGE, class F>
hipError_t launch_postscan_v(const MsCall& c, const F& f) {
    constexpr int K = tune_items(LOG_B, KV, LARGE);
    constexpr int LB = tune_level_bits(LOG_B);
    hipLaunchKernelGGL((ms_postscan_kernel<LOG_B, K, LB, KV, VEC, F>), dim3(c.num_tiles),
                       dim3(kThreads), 0, c.stream, c.kin, c.vin, c.kout, c.vout, c.n,
                       c.num_tiles, c.H, f);
    return hipGetLastError();
}

template <int LOG_B, bool KV, class F>
hipError_t launch_postscan_t(const MsCall& c, const F& f) {
    if constexpr (tune_items(LOG_B, KV, true) != tune_items(LOG_B, KV, false)) {
        if (c.large)
            return c.vec_ok ? launch_postscan_v<LOG_B, KV, true, true>(c, f)
                            : launch_postscan_v<LOG_B, KV, false, true>(c, f);
    }
    return c.vec_ok ? launch_postscan_v<LOG_B, KV, true, false>(c, f)
                    : launch_postscan_v<LOG_B, KV, false, false>(c, f);
}

#define GMS_SWITCH_LOGB_KV(FN, c, f)                                   \
    switch ((c).log_b * 2 + ((c).kv ? 1 : 0)) {                        \
        case 2: return FN<1, false>(c, f);                             \
        case 3: return FN<1, true>(c, f);                              \
        case 4: return FN<2, false>(c, f);                             \
        case 5: return FN<2, true>(c, f);                              \
        case 6: return FN<3, false>(c, f);                             \
        case 7: return FN<3, true>(c, f);                              \
        case 8: return FN<4, false>(c, f);                             \
        case 9: return FN<4, true>(c, f);                              \
        case 10: return FN<5, false>(c, f);                            \
        case 11: return FN<5, true>(c, f);                             \
        case 12: return FN<6, false>(c, f);                            \
        case 13: return FN<6, true>(c, f);                             \
        case 14: return FN<7, false>(c, f);                            \
        case 15: return FN<7, true>(c, f);                             \
        case 16: return FN<8, false>(c, f);                            \
        case 17: return FN<8, true>(c, f);                             \
        case 18: return FN<9, false>(c, f);                            \
        case 19: return FN<9, true>(c, f);                             \
        case 20: return FN<10, false>(c, f);                           \
        case 21: return FN<10, true>(c, f);                            \
        default: return hipErrorInvalidValue;                          \
    }

#define GMS_DEFINE_LAUNCHERS(FT)                                                         \
    hipError_t launch_prescan(const MsCall& c, const FT& f) {                            \
        GMS_SWITCH_LOGB_KV(launch_prescan_t, c, f)                                       \
    }                                                                                    \
    hipError_t launch_postscan(const MsCall& c, const FT& f) {                           \
        GMS_SWITCH_LOGB_KV(launch_postscan_t, c, f)                                      \
    }

}  // namespace gms
