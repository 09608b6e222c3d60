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
constexpr uint32_t kRangeGrid = 1024;
constexpr int kHistCopies = 8;
constexpr int kRangeCopies = 4;

// bins_of(x, b) maps a register array of words to bins (>= nbins: not
// counted); it sees a whole tile row at once so that a kernel can issue the
// lookups of all its items before it waits for any.  Elements are W words
// (1, or 2 for 64-bit keys): a 16-byte vector of a striped row holds whole
// elements, and bins_of reports an element's bin at its first word and an
// out-of-range bin at its others.  n counts elements.
//
// Every wave counts into C copies of its sub-histogram, lane l into copy
// l % C, with rows of 257 words so that the copies of one bin lie in
// different banks.  Keys that share a bin across a wave (sorted runs, one
// dominant bin) then add to C addresses instead of one, which the LDS
// serialises per address; evenly spread keys see the same bank spread as
// with one copy.
template <bool VEC, int W = 1, int C = kHistCopies, bool STREAM = false, class BinsOf>
__device__ __forceinline__ void hist_body(const uint32_t* __restrict__ p, uint32_t n, uint32_t nbins,
                                          uint32_t* __restrict__ d_hist, BinsOf bins_of) {
    static_assert(W == 1 || W == 2, "elements of 1 or 2 words");
    static_assert(C >= 1 && C <= 64 && (C & (C - 1)) == 0, "copies: a power of two");
    constexpr uint32_t ROW = C > 1 ? 257u : 256u;
    constexpr uint32_t NH = (uint32_t)(kWaves * C);
    __shared__ uint32_t s_h[NH * ROW];
    const uint32_t tid = threadIdx.x, wid = tid >> 6;
    for (uint32_t i = tid; i < NH * ROW; i += kThreads) s_h[i] = 0;
    __syncthreads();
    uint32_t* my = s_h + (wid * (uint32_t)C + (tid & (uint32_t)(C - 1))) * ROW;
    const uint32_t nfull = (uint32_t)((size_t)n * W / kHistTile);
    uint32_t t = blockIdx.x;
    uint32_t x[kHistItems];
    if (t < nfull) load_striped4_full<kHistItems, VEC, STREAM>(p, (size_t)t * kHistTile, x);
    while (t < nfull) {
        const uint32_t tn = t + gridDim.x;
        uint32_t y[kHistItems];
        if (tn < nfull) load_striped4_full<kHistItems, VEC, STREAM>(p, (size_t)tn * kHistTile, y);
        uint32_t b[kHistItems];
        bins_of(x, b);
#pragma unroll
        for (int k = 0; k < kHistItems; ++k)
            if (b[k] < nbins) atomicAdd(&my[b[k]], 1u);
#pragma unroll
        for (int k = 0; k < kHistItems; ++k) x[k] = y[k];
        t = tn;
    }
    if (blockIdx.x == 0) {
        for (size_t e = (size_t)nfull * (kHistTile / W) + tid; e < n; e += kThreads) {
            uint32_t xt[W], bt[W];
#pragma unroll
            for (int w = 0; w < W; ++w) xt[w] = p[e * W + w];
            bins_of(xt, bt);
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (bt[w] < nbins) atomicAdd(&my[bt[w]], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = tid; b < nbins; b += kThreads) {
        uint32_t s = 0;
#pragma unroll
        for (uint32_t q = 0; q < NH; ++q) s += s_h[q * ROW + b];
        if (s) atomicAdd(&d_hist[b], s);
    }
}

// Adapts a per-word bin function to hist_body.
template <class F>
__device__ __forceinline__ auto per_word(F f) {
    return [f](auto& x, auto& b) {
        constexpr int N = (int)(sizeof(x) / sizeof(x[0]));
#pragma unroll
        for (int k = 0; k < N; ++k) b[k] = f(x[k]);
    };
}

// STREAM: streamed loads, for inputs of at least kStreamLoadMin words (each
// line is read once and the input is larger than the caches).
template <bool VEC, bool STREAM, class F>
__global__ __launch_bounds__(kThreads) void hist_u32_kernel(const uint32_t* __restrict__ kin,
                                                            uint32_t n, uint32_t nbins,
                                                            uint32_t* __restrict__ d_hist, F f) {
    hist_body<VEC, 1, kHistCopies, STREAM>(kin, n, nbins, d_hist,
                                           per_word([&](uint32_t k) { return f(k); }));
}

template <bool VEC, bool STREAM>
__global__ __launch_bounds__(kThreads) void hist_even_f32_kernel(const uint32_t* __restrict__ xbits,
                                                                 uint32_t n, uint32_t nbins,
                                                                 uint32_t* __restrict__ d_hist,
                                                                 float lower, float inv_delta) {
    hist_body<VEC, 1, kHistCopies, STREAM>(xbits, n, nbins, d_hist, per_word([&](uint32_t w) {
        const float d = __uint_as_float(w) - lower;
        // The product is rounded on its own (no contraction with a later add);
        // negative, NaN and >= nbins samples map to an out-of-range bin.
        const float s = __fmul_rn(d, inv_delta);
        return (d >= 0.0f && s < (float)nbins) ? (uint32_t)s : 0xFFFFFFFFu;
    }));
}

// Custom bin edges (the reference's multisplit2_histogram_range_*,
// multisplit2_histograms.cuh:397-699, compared there with CUB's
// HistogramRange): bin i = [levels[i], levels[i+1]); samples outside
// [levels[0], levels[nl-1]) and NaN are not counted.
//
// Every workgroup first builds a table over kRangeCells equal cells of
// [levels[0], levels[nl-1]): cell c holds lo_c = #{interior edges j with
// cell(levels[j]) < c} and cnt_c = #{j : cell(levels[j]) == c}.  cell() is
// non-decreasing and the same function is applied to samples and edges, so
// for a sample x in cell c every edge in a lower cell is < x and every edge in
// a higher one is > x: its bin lies in [lo_c, lo_c + cnt_c].  Most cells hold
// no edge and the sample's bin is lo_c after one LDS read; the others finish
// with a binary search of floor(log2(cnt_c)) + 1 steps over the edges, which
// are staged in LDS padded with the type's largest value.  The table reads of
// a thread's 16 samples are issued together, and the search runs the same
// floor(log2(max_c cnt_c)) + 1 steps for every sample, a sample joining at
// the first step h <= cnt_c; samples outside it read a broadcast word.
constexpr uint32_t kRangeCells = 4096;
constexpr uint32_t kCellsPerThread = kRangeCells / kThreads;

template <class T>
struct RangeCellF;

// float: t = (x - lower) * (cells / (upper - lower)), each operation rounded
// on its own and each monotone, truncated and clamped to a cell.
template <>
struct RangeCellF<float> {
    float lower, scale;
    __device__ RangeCellF(float lo, float hi)
        : lower(lo), scale(__fdiv_rn((float)kRangeCells, __fsub_rn(hi, lo))) {}
    __device__ uint32_t operator()(float x) const {
        const float t = __fmul_rn(__fsub_rn(x, lower), scale);
        return t > 0.0f ? (t < (float)kRangeCells ? (uint32_t)t : kRangeCells - 1u) : 0u;
    }
    __device__ static float pad() { return __builtin_inff(); }
};

// uint32: cells of 2^shift keys from lower.
template <>
struct RangeCellF<uint32_t> {
    uint32_t lower, shift;
    __device__ RangeCellF(uint32_t lo, uint32_t hi) : lower(lo), shift(0) {
        const uint32_t r1 = hi - lo - 1u;  // largest x - lower of an in-range x
        const uint32_t bits = r1 ? 32u - (uint32_t)__builtin_clz(r1) : 0u;
        shift = bits > 12u ? bits - 12u : 0u;
    }
    __device__ uint32_t operator()(uint32_t x) const {
        const uint32_t c = (x - lower) >> shift;
        return c < kRangeCells ? c : kRangeCells - 1u;
    }
    __device__ static uint32_t pad() { return 0xFFFFFFFFu; }
};
static_assert(kRangeCells == (1u << 12), "the integer cell shift assumes 2^12 cells");

template <class T, bool VEC>
__global__ __launch_bounds__(kThreads) void hist_range_kernel(const uint32_t* __restrict__ xbits,
                                                              uint32_t n, uint32_t nl,
                                                              const T* __restrict__ d_levels,
                                                              uint32_t* __restrict__ d_hist) {
    __shared__ __attribute__((aligned(16))) uint32_t s_tab[kRangeCells];  // lo_c | cnt_c << 16
    __shared__ T s_lv[512];
    __shared__ uint32_t s_ws[kWaves];
    __shared__ uint32_t s_maxcnt;
    const uint32_t tid = threadIdx.x, nb = nl - 1u;
    const T lo_edge = d_levels[0], hi_edge = d_levels[nb];
    const RangeCellF<T> cell(lo_edge, hi_edge);
    // s_lv[i] = levels[i] for i < nb, then the pad (index <= lo_c + 2^steps - 1 <= 510)
    for (uint32_t i = tid; i < 512u; i += kThreads) s_lv[i] = i < nb ? d_levels[i] : RangeCellF<T>::pad();
    uint32_t* mine = s_tab + tid * kCellsPerThread;
#pragma unroll
    for (uint32_t i = 0; i < kCellsPerThread; i += 4)
        *reinterpret_cast<uint4*>(mine + i) = make_uint4(0u, 0u, 0u, 0u);
    if (tid == 0) s_maxcnt = 0;
    __syncthreads();
    for (uint32_t j = tid + 1u; j < nb; j += kThreads) atomicAdd(&s_tab[cell(d_levels[j])], 1u);
    __syncthreads();
    uint32_t c[kCellsPerThread], sum = 0, mx = 0;
#pragma unroll
    for (uint32_t i = 0; i < kCellsPerThread; i += 4) {
        const uint4 v = *reinterpret_cast<const uint4*>(mine + i);
        c[i] = v.x; c[i + 1] = v.y; c[i + 2] = v.z; c[i + 3] = v.w;
        sum += v.x + v.y + v.z + v.w;
        mx = std::max(mx, std::max(std::max(v.x, v.y), std::max(v.z, v.w)));
    }
    if (mx) atomicMax(&s_maxcnt, mx);
    uint32_t total;
    uint32_t run = block_excl_scan<kWaves>(sum, s_ws, total);  // its barrier also orders s_maxcnt
    const uint32_t maxcnt = GMS_READFIRSTLANE(s_maxcnt);
    const uint32_t h0 = maxcnt ? 1u << (31 - __builtin_clz(maxcnt)) : 0u;  // first search step
#pragma unroll
    for (uint32_t i = 0; i < kCellsPerThread; ++i) {
        const uint32_t e = run | (c[i] << 16);
        run += c[i];
        c[i] = e;
    }
#pragma unroll
    for (uint32_t i = 0; i < kCellsPerThread; i += 4)
        *reinterpret_cast<uint4*>(mine + i) = make_uint4(c[i], c[i + 1], c[i + 2], c[i + 3]);
    // (hist_body's first barrier orders the table before any lookup)
    // Four sub-histogram copies per wave (35 KB of LDS per workgroup with the
    // cell table: the grid of kRangeGrid = 4 workgroups per CU is resident):
    // sorted samples, whose lanes share a bin, contend on 4 addresses instead
    // of one.
    hist_body<VEC, 1, kRangeCopies>(xbits, n, nb, d_hist, [&](auto& x, auto& b) {
        constexpr int N = (int)(sizeof(x) / sizeof(x[0]));
        T v[N];
        uint32_t cnt[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            __builtin_memcpy(&v[k], &x[k], sizeof(T));
            const uint32_t e = s_tab[cell(v[k])];  // any word maps to a cell
            b[k] = e & 0xFFFFu;
            cnt[k] = e >> 16;
        }
        for (uint32_t h = h0; h; h >>= 1) {
#pragma unroll
            for (int k = 0; k < N; ++k) {
                // inactive samples add 0 whatever they read
                const uint32_t hk = h <= cnt[k] ? h : 0u;
                const uint32_t idx = hk ? b[k] + hk : 0u;
                b[k] += v[k] >= s_lv[idx] ? hk : 0u;
            }
        }
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (!(v[k] >= lo_edge && v[k] < hi_edge)) b[k] = 0xFFFFFFFFu;  // outside, or NaN
    });
}

// 64-bit keys through a 64-bit identifier: the same prefetching resident
// grid over the key array viewed as words, each (low, high) word pair one key.
template <bool VEC, bool STREAM>
__global__ __launch_bounds__(kThreads) void hist_u64_kernel(const uint32_t* __restrict__ words,
                                                            uint32_t n, uint32_t nbins,
                                                            uint32_t* __restrict__ d_hist, AnyF64 f) {
    hist_body<VEC, 2, kHistCopies, STREAM>(words, n, nbins, d_hist, [&](auto& x, auto& b) {
        constexpr int N = (int)(sizeof(x) / sizeof(x[0]));
#pragma unroll
        for (int k = 0; k < N; k += 2) {
            b[k] = f((uint64_t)x[k] | ((uint64_t)x[k + 1] << 32));
            b[k + 1] = 0xFFFFFFFFu;
        }
    });
}

static uint32_t hist_grid(uint32_t n, uint32_t max_grid = kHistGrid) {
    const uint32_t full = n / kHistTile;
    return full < max_grid ? (full ? full : 1u) : max_grid;
}

static bool aligned16(const void* p) { return (((uintptr_t)p) & 15u) == 0; }
static bool stream_words(size_t words) { return words >= kStreamLoadMin; }

hipError_t launch_histogram_u64(const uint64_t* keys, uint32_t n, uint32_t* d_hist,
                                uint32_t nbins, const AnyF64& f, hipStream_t stream) {
    const size_t full = (size_t)n * 2 / kHistTile;
    const uint32_t grid = full < kHistGrid ? (full ? (uint32_t)full : 1u) : kHistGrid;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(keys);
    if (aligned16(keys) && stream_words((size_t)n * 2))
        hipLaunchKernelGGL((hist_u64_kernel<true, true>), dim3(grid), dim3(kThreads), 0, stream,
                           w, n, nbins, d_hist, f);
    else if (aligned16(keys))
        hipLaunchKernelGGL((hist_u64_kernel<true, false>), dim3(grid), dim3(kThreads), 0, stream,
                           w, n, nbins, d_hist, f);
    else
        hipLaunchKernelGGL((hist_u64_kernel<false, false>), dim3(grid), dim3(kThreads), 0, stream,
                           w, n, nbins, d_hist, f);
    return hipGetLastError();
}

template <class F>
static hipError_t launch_hist(const uint32_t* kin, uint32_t n, uint32_t* d_hist, uint32_t nbins,
                              const F& f, hipStream_t stream) {
    if (aligned16(kin) && stream_words(n))
        hipLaunchKernelGGL((hist_u32_kernel<true, true, F>), dim3(hist_grid(n)), dim3(kThreads),
                           0, stream, kin, n, nbins, d_hist, f);
    else if (aligned16(kin))
        hipLaunchKernelGGL((hist_u32_kernel<true, false, F>), dim3(hist_grid(n)), dim3(kThreads),
                           0, stream, kin, n, nbins, d_hist, f);
    else
        hipLaunchKernelGGL((hist_u32_kernel<false, false, F>), dim3(hist_grid(n)), dim3(kThreads),
                           0, stream, kin, n, nbins, d_hist, f);
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

template <class T>
static hipError_t launch_range(const void* x, uint32_t n, int32_t* d_hist, uint32_t num_levels,
                               const T* d_levels, hipStream_t stream) {
    const uint32_t* xb = reinterpret_cast<const uint32_t*>(x);
    uint32_t* h = reinterpret_cast<uint32_t*>(d_hist);
    if (aligned16(x))
        hipLaunchKernelGGL((hist_range_kernel<T, true>), dim3(hist_grid(n, kRangeGrid)),
                           dim3(kThreads), 0, stream, xb, n, num_levels, d_levels, h);
    else
        hipLaunchKernelGGL((hist_range_kernel<T, false>), dim3(hist_grid(n, kRangeGrid)),
                           dim3(kThreads), 0, stream, xb, n, num_levels, d_levels, h);
    return hipGetLastError();
}

hipError_t launch_histogram_range_f32(const float* x, uint32_t n, int32_t* d_hist,
                                      uint32_t num_levels, const float* d_levels,
                                      hipStream_t stream) {
    return launch_range<float>(x, n, d_hist, num_levels, d_levels, stream);
}

hipError_t launch_histogram_range_u32(const uint32_t* x, uint32_t n, int32_t* d_hist,
                                      uint32_t num_levels, const uint32_t* d_levels,
                                      hipStream_t stream) {
    return launch_range<uint32_t>(x, n, d_hist, num_levels, d_levels, stream);
}

hipError_t launch_histogram_even_f32(const float* x, uint32_t n, int32_t* d_hist, uint32_t nbins,
                                     float lower, float inv_delta, hipStream_t stream) {
    const uint32_t* xb = reinterpret_cast<const uint32_t*>(x);
    uint32_t* h = reinterpret_cast<uint32_t*>(d_hist);
    if (aligned16(x) && stream_words(n))
        hipLaunchKernelGGL((hist_even_f32_kernel<true, true>), dim3(hist_grid(n)), dim3(kThreads),
                           0, stream, xb, n, nbins, h, lower, inv_delta);
    else if (aligned16(x))
        hipLaunchKernelGGL((hist_even_f32_kernel<true, false>), dim3(hist_grid(n)),
                           dim3(kThreads), 0, stream, xb, n, nbins, h, lower, inv_delta);
    else
        hipLaunchKernelGGL((hist_even_f32_kernel<false, false>), dim3(hist_grid(n)),
                           dim3(kThreads), 0, stream, xb, n, nbins, h, lower, inv_delta);
    return hipGetLastError();
}

}  // namespace gms
