// Device-wide histograms of up to 256 bins (the reference's multisplit
// histogram, src/kernels/histogram/multisplit2_histograms.cuh:30-392).
//
// A resident grid (two workgroups per CU) walks full tiles of 256 x 16 words
// with unconditional 16-byte loads, issuing the next tile's loads before it
// counts the current one, so every workgroup keeps a tile in flight while its
// waves add bins into their own LDS sub-histograms (ds_add).  The ragged tail
// (< one tile) is counted by workgroup 0 with guarded loads.  Each workgroup
// folds its sub-histograms and adds its non-zero bins to the global
// histogram once at the end.
#include "gms_internal.hpp"
#include "gms_kernels.hpp"

namespace gms {

constexpr int kHistItems = 16;
constexpr uint32_t kHistTile = kThreads * kHistItems;
constexpr uint32_t kHistGrid = 512;

template <bool VEC, class BinOf>
__device__ __forceinline__ void hist_body(const uint32_t* __restrict__ p, uint32_t n, uint32_t nbins,
                                          uint32_t* __restrict__ d_hist, BinOf bin_of) {
    __shared__ uint32_t s_h[kWaves * 256];
    const uint32_t tid = threadIdx.x, wid = tid >> 6;
    for (uint32_t i = tid; i < kWaves * 256u; i += kThreads) s_h[i] = 0;
    __syncthreads();
    uint32_t* my = s_h + wid * 256u;
    const uint32_t nfull = n / kHistTile;
    uint32_t t = blockIdx.x;
    uint32_t x[kHistItems];
    if (t < nfull) load_striped4_full<kHistItems, VEC>(p, (size_t)t * kHistTile, x);
    while (t < nfull) {
        const uint32_t tn = t + gridDim.x;
        uint32_t y[kHistItems];
        if (tn < nfull) load_striped4_full<kHistItems, VEC>(p, (size_t)tn * kHistTile, y);
#pragma unroll
        for (int k = 0; k < kHistItems; ++k) {
            const uint32_t b = bin_of(x[k]);
            if (b < nbins) atomicAdd(&my[b], 1u);
        }
#pragma unroll
        for (int k = 0; k < kHistItems; ++k) x[k] = y[k];
        t = tn;
    }
    if (blockIdx.x == 0) {
        for (uint32_t i = nfull * kHistTile + tid; i < n; i += kThreads) {
            const uint32_t b = bin_of(p[i]);
            if (b < nbins) atomicAdd(&my[b], 1u);
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

template <bool VEC, class F>
__global__ __launch_bounds__(kThreads) void hist_u32_kernel(const uint32_t* __restrict__ kin,
                                                            uint32_t n, uint32_t nbins,
                                                            uint32_t* __restrict__ d_hist, F f) {
    hist_body<VEC>(kin, n, nbins, d_hist, [&](uint32_t k) { return f(k); });
}

template <bool VEC>
__global__ __launch_bounds__(kThreads) void hist_even_f32_kernel(const uint32_t* __restrict__ xbits,
                                                                 uint32_t n, uint32_t nbins,
                                                                 uint32_t* __restrict__ d_hist,
                                                                 float lower, float inv_delta) {
    hist_body<VEC>(xbits, n, nbins, d_hist, [&](uint32_t w) {
        const float d = __uint_as_float(w) - lower;
        // The product is rounded on its own (no contraction with a later add);
        // negative, NaN and >= nbins samples map to an out-of-range bin.
        const float s = __fmul_rn(d, inv_delta);
        return (d >= 0.0f && s < (float)nbins) ? (uint32_t)s : 0xFFFFFFFFu;
    });
}

static uint32_t hist_grid(uint32_t n) {
    const uint32_t full = n / kHistTile;
    return full < kHistGrid ? (full ? full : 1u) : kHistGrid;
}

static bool aligned16(const void* p) { return (((uintptr_t)p) & 15u) == 0; }

template <class F>
static hipError_t launch_hist(const uint32_t* kin, uint32_t n, uint32_t* d_hist, uint32_t nbins,
                              const F& f, hipStream_t stream) {
    if (aligned16(kin))
        hipLaunchKernelGGL((hist_u32_kernel<true, F>), dim3(hist_grid(n)), dim3(kThreads), 0,
                           stream, kin, n, nbins, d_hist, f);
    else
        hipLaunchKernelGGL((hist_u32_kernel<false, F>), dim3(hist_grid(n)), dim3(kThreads), 0,
                           stream, kin, n, nbins, d_hist, f);
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
    const uint32_t* xb = reinterpret_cast<const uint32_t*>(x);
    uint32_t* h = reinterpret_cast<uint32_t*>(d_hist);
    if (aligned16(x))
        hipLaunchKernelGGL((hist_even_f32_kernel<true>), dim3(hist_grid(n)), dim3(kThreads), 0,
                           stream, xb, n, nbins, h, lower, inv_delta);
    else
        hipLaunchKernelGGL((hist_even_f32_kernel<false>), dim3(hist_grid(n)), dim3(kThreads), 0,
                           stream, xb, n, nbins, h, lower, inv_delta);
    return hipGetLastError();
}

}  // namespace gms
