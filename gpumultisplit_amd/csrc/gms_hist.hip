// Device-wide histograms of up to 256 bins (the reference's multisplit
// histogram, src/kernels/histogram/multisplit2_histograms.cuh:30-392).
//
// Persistent grid over tiles of 4096 keys; each wave counts into its own LDS
// sub-histogram with ds_add, the block folds its sub-histograms and adds its
// non-zero bins into the global histogram once at the end.
#include "gms_internal.hpp"
#include "gms_kernels.hpp"

namespace gms {

constexpr uint32_t kHistTile = kThreads * 16;

template <class BinOf>
__device__ __forceinline__ void hist_body(uint32_t n, uint32_t nbins, uint32_t* __restrict__ d_hist,
                                          BinOf bin_of) {
    __shared__ uint32_t s_h[kWaves * 256];
    const uint32_t tid = threadIdx.x, wid = tid >> 6;
    for (uint32_t i = tid; i < kWaves * 256u; i += kThreads) s_h[i] = 0;
    __syncthreads();
    uint32_t* my = s_h + wid * 256u;
    const size_t ntiles = ((size_t)n + kHistTile - 1) / kHistTile;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const size_t base = t * kHistTile;
#pragma unroll 4
        for (uint32_t j = 0; j < 16; ++j) {
            const size_t i = base + (size_t)j * kThreads + tid;
            if (i < n) {
                const uint32_t b = bin_of(i);
                if (b < nbins) atomicAdd(&my[b], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t b = tid; b < nbins; b += kThreads) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += s_h[w * 256 + b];
        if (s) atomicAdd(&d_hist[b], s);
    }
}

template <class F>
__global__ __launch_bounds__(kThreads) void hist_u32_kernel(const uint32_t* __restrict__ kin,
                                                            uint32_t n, uint32_t nbins,
                                                            uint32_t* __restrict__ d_hist, F f) {
    hist_body(n, nbins, d_hist, [&](size_t i) { return f(kin[i]); });
}

__global__ __launch_bounds__(kThreads) void hist_even_f32_kernel(const float* __restrict__ x,
                                                                 uint32_t n, uint32_t nbins,
                                                                 uint32_t* __restrict__ d_hist,
                                                                 float lower, float inv_delta) {
    hist_body(n, nbins, d_hist, [&](size_t i) {
        const float d = x[i] - lower;
        // The product is rounded on its own (no contraction with a later add);
        // negative, NaN and >= nbins samples map to an out-of-range bin.
        const float s = __fmul_rn(d, inv_delta);
        return (d >= 0.0f && s < (float)nbins) ? (uint32_t)s : 0xFFFFFFFFu;
    });
}

static uint32_t hist_grid(uint32_t n) {
    uint32_t tiles = (uint32_t)(((size_t)n + kHistTile - 1) / kHistTile);
    return tiles < 2048u ? (tiles ? tiles : 1u) : 2048u;
}

template <class F>
static hipError_t launch_hist(const uint32_t* kin, uint32_t n, uint32_t* d_hist, uint32_t nbins,
                              const F& f, hipStream_t stream) {
    hipLaunchKernelGGL((hist_u32_kernel<F>), dim3(hist_grid(n)), dim3(kThreads), 0, stream, kin,
                       n, nbins, d_hist, f);
    return hipGetLastError();
}

hipError_t launch_histogram_u32(const uint32_t* kin, uint32_t n, uint32_t* d_hist, uint32_t nbins,
                                gms_bucket_desc d, hipStream_t stream) {
    switch (d.kind) {
        case GMS_BUCKET_RANGE: return launch_hist(kin, n, d_hist, nbins, RangeF{nbins}, stream);
        case GMS_BUCKET_DELTA:
            return launch_hist(kin, n, d_hist, nbins, DeltaF::make(d.param, nbins), stream);
        case GMS_BUCKET_DIGIT:
            return launch_hist(kin, n, d_hist, nbins,
                               DigitF{d.param, d.width ? (1u << d.width) - 1u : nbins - 1u}, stream);
        case GMS_BUCKET_LOWBITS:
            return launch_hist(kin, n, d_hist, nbins, DigitF{0u, (1u << d.param) - 1u}, stream);
        case GMS_BUCKET_IDENTITY: return launch_hist(kin, n, d_hist, nbins, IdentF{nbins - 1u}, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_histogram_even_f32(const float* x, uint32_t n, int32_t* d_hist, uint32_t nbins,
                                     float lower, float inv_delta, hipStream_t stream) {
    hipLaunchKernelGGL(hist_even_f32_kernel, dim3(hist_grid(n)), dim3(kThreads), 0, stream, x, n,
                       nbins, reinterpret_cast<uint32_t*>(d_hist), lower, inv_delta);
    return hipGetLastError();
}

}  // namespace gms
