// Batched multisplit of elements with 64-bit keys and/or values: the
// (instance, tile) work items of the 64-bit groups of a batch, one workgroup
// per item, on the 64-bit tile code of gms_wide_kernels.hpp.  Their matrix
// blocks are part of the batch's one concatenated matrix, so the batch's one
// scan serves them with the 32-bit groups.
#include "gms_batch.hpp"
#include "gms_wide_kernels.hpp"

namespace gms {

template <int LOG_B, int KW, int VW, bool VEC>
__global__ __launch_bounds__(kThreads) void ms_batch_wide_prescan_kernel(
    const BatchWideInst* __restrict__ inst, const BatchWork* __restrict__ work, uint32_t num_work,
    uint32_t* __restrict__ H) {
    constexpr int K = wide_items(KW, VW);
    __shared__ __attribute__((aligned(16))) uint32_t s_h[kWaves * (1 << LOG_B)];
    const BatchWork it = work[xcd_tile(blockIdx.x, num_work)];
    const BatchWideInst& I = inst[it.inst];
    if constexpr (KW == 2)
        wide_prescan_tile<LOG_B, K, KW, VEC>(I.kin, I.n, it.tile, I.num_tiles, H + I.mbase, I.f64,
                                             s_h);
    else
        wide_prescan_tile<LOG_B, K, KW, VEC>(I.kin, I.n, it.tile, I.num_tiles, H + I.mbase, I.f,
                                             s_h);
}

template <int LOG_B, int KW, int VW, bool VEC>
__global__ __launch_bounds__(kThreads) void ms_batch_wide_postscan_kernel(
    const BatchWideInst* __restrict__ inst, const BatchWork* __restrict__ work, uint32_t num_work,
    const uint32_t* __restrict__ G, uint32_t m_total, uint32_t keys_total) {
    constexpr int K = wide_items(KW, VW);
    constexpr int LB = tune_level_bits(LOG_B);
    constexpr int B = 1 << LOG_B;
    __shared__ __attribute__((aligned(16))) WideSmem<LOG_B, K, LB, KW, VW> sm;
    const uint32_t tid = threadIdx.x;
    const BatchWork it = work[xcd_tile(blockIdx.x, num_work)];
    const BatchWideInst& I = inst[it.inst];
    const uint32_t t = it.tile;
    // G[b][t] of this instance and the next word in scan order; the word
    // after the whole concatenation is the batch's element total.
    const uint32_t bt = tid < (uint32_t)B ? tid : (uint32_t)B - 1u;
    const uint32_t idx = I.mbase + bt * I.num_tiles + t;
    const uint32_t g0 = G[idx];
    const uint32_t g1 = idx + 1u < m_total ? G[idx + 1u] : keys_total;
    const uint32_t g0l = g0 - I.obase;  // start inside this instance's output
    if (t == 0 && I.offs && tid < I.nb) I.offs[tid] = g0l;
    if constexpr (KW == 2)
        wide_postscan_item<LOG_B, K, LB, KW, VW, VEC>(I.kin, I.vin, I.kout, I.vout, I.n, t, g0l,
                                                      g1 - g0, I.f64, sm);
    else
        wide_postscan_item<LOG_B, K, LB, KW, VW, VEC>(I.kin, I.vin, I.kout, I.vout, I.n, t, g0l,
                                                      g1 - g0, I.f, sm);
}

template <int LOG_B, int KW, int VW>
static hipError_t batch_wide_b(const BatchWideLaunch& c, bool post) {
    const dim3 grid(c.num_work), block(kThreads);
    if (post) {
        if (c.vec)
            hipLaunchKernelGGL((ms_batch_wide_postscan_kernel<LOG_B, KW, VW, true>), grid, block, 0,
                               c.stream, c.inst, c.work, c.num_work, (const uint32_t*)c.H,
                               c.m_total, c.keys_total);
        else
            hipLaunchKernelGGL((ms_batch_wide_postscan_kernel<LOG_B, KW, VW, false>), grid, block,
                               0, c.stream, c.inst, c.work, c.num_work, (const uint32_t*)c.H,
                               c.m_total, c.keys_total);
    } else {
        if (c.vec)
            hipLaunchKernelGGL((ms_batch_wide_prescan_kernel<LOG_B, KW, VW, true>), grid, block, 0,
                               c.stream, c.inst, c.work, c.num_work, c.H);
        else
            hipLaunchKernelGGL((ms_batch_wide_prescan_kernel<LOG_B, KW, VW, false>), grid, block,
                               0, c.stream, c.inst, c.work, c.num_work, c.H);
    }
    return hipGetLastError();
}

template <int KW, int VW>
static hipError_t batch_wide_shape(const BatchWideLaunch& c, bool post) {
    switch (c.log_b) {
        case 1: return batch_wide_b<1, KW, VW>(c, post);
        case 2: return batch_wide_b<2, KW, VW>(c, post);
        case 3: return batch_wide_b<3, KW, VW>(c, post);
        case 4: return batch_wide_b<4, KW, VW>(c, post);
        case 5: return batch_wide_b<5, KW, VW>(c, post);
        case 6: return batch_wide_b<6, KW, VW>(c, post);
        case 7: return batch_wide_b<7, KW, VW>(c, post);
        case 8: return batch_wide_b<8, KW, VW>(c, post);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_wide(const BatchWideLaunch& c, bool post) {
    if (c.num_work == 0) return hipSuccess;
    if (c.kw == 2) {
        switch (c.vw) {
            case 0: return batch_wide_shape<2, 0>(c, post);
            case 1: return batch_wide_shape<2, 1>(c, post);
            case 2: return batch_wide_shape<2, 2>(c, post);
            default: return hipErrorInvalidValue;
        }
    }
    if (c.kw == 1 && c.vw == 2) return batch_wide_shape<1, 2>(c, post);
    return hipErrorInvalidValue;
}

}  // namespace gms
