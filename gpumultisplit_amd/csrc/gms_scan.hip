// Device-wide exclusive scan of u32 and small utility kernels.
//
// The scan replaces cub::DeviceScan::ExclusiveSum of the reference's scan
// stage (src/api/multisplit.cuh:141, :258).  Reduce-then-scan over chunks of
// 4096 elements: (1) each workgroup sums its chunk, (2) one workgroup scans the
// chunk sums, (3) each workgroup rescans its chunk with its offset.  Up to
// kScanFusedParts chunks, step (2) is folded into (3): every workgroup adds
// the sums of the chunks before its own.  Chunks are read with coalesced
// 16-byte loads.
#include "gms_internal.hpp"
#include "gms_kernels.hpp"

namespace gms {

__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                               uint32_t m,
                                                               uint32_t* __restrict__ part) {
    __shared__ uint32_t s_w[kWaves];
    const size_t base = (size_t)blockIdx.x * kScanChunk;
    const uint32_t tid = threadIdx.x;
    uint32_t sum = 0;
    const bool al = (((uintptr_t)in) & 15u) == 0;
    if (al && base + kScanChunk <= m) {
#pragma unroll
        for (int j = 0; j < kScanItems / 4; ++j) {
            uint4 v = *reinterpret_cast<const uint4*>(in + base + 4u * ((uint32_t)j * kThreads + tid));
            sum += v.x + v.y + v.z + v.w;
        }
    } else {
        for (uint32_t i = tid; base + i < m && i < kScanChunk; i += kThreads) sum += in[base + i];
    }
    uint32_t tot;
    block_excl_scan<kWaves>(sum, s_w, tot);
    if (tid == 0) part[blockIdx.x] = tot;
}

// ---------------------------------------------------------------------------
// Single-pass exclusive scan: one launch, each word read once and written
// once.  Chunk c (workgroup c) publishes its total in status[c] (bit 62 set,
// total in bits 0-31) as soon as it has loaded its chunk, then its threads
// read the statuses of all chunks before it in parallel -- waiting for
// unpublished ones -- and add them up.
//  * Chunks are numbered by workgroup index, not by a ticket counter: every
//    XCD starts its workgroups in index order, so the lowest unfinished chunk
//    is always running or next to start, and it waits only for lower chunks.
//    On MI355X a ticket (one device-scope atomic per workgroup on one
//    address) and a finish counter took the scan of 1170 chunks from 11 to
//    36-39 us.  The spin bound ends every wait even if a status were lost.
//  * Summing every predecessor's total (the chunk count is small next to the
//    matrix: 1170 statuses for 2^27 keys in 256 buckets) replaces the
//    walk-back to the nearest inclusive prefix of decoupled look-back
//    (Merrill and Garland, 2016), whose prefixes would form a chain of
//    device-coherent rounds when every chunk's workgroup is resident at once.
//  * The statuses must be zero on entry: the prescan of the same call clears
//    them, so no clearing launch is needed and a captured graph replays
//    correctly.
// ---------------------------------------------------------------------------
constexpr uint64_t kStatusTotal = 1ull << 62;
constexpr uint32_t kStatusSpins = 1u << 20;

__global__ __launch_bounds__(kThreads) void scan_single_kernel(const uint32_t* in, uint32_t* out,
                                                               uint32_t m,
                                                               uint64_t* __restrict__ status) {
    __shared__ uint32_t s_w[kWaves];
    const uint32_t tid = threadIdx.x;
    const uint32_t c = blockIdx.x;
    const size_t base = (size_t)c * kScanChunk;
    const uint32_t off = tid * (uint32_t)kScanItems;
    const bool full = base + kScanChunk <= m && ((((uintptr_t)in) | ((uintptr_t)out)) & 15u) == 0;
    uint32_t x[kScanItems];
    if (full) {
#pragma unroll
        for (int i = 0; i < kScanItems / 4; ++i) {
            uint4 v = *reinterpret_cast<const uint4*>(in + base + off + 4 * i);
            x[4 * i + 0] = v.x; x[4 * i + 1] = v.y; x[4 * i + 2] = v.z; x[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) x[i] = (base + off + i < m) ? in[base + off + i] : 0u;
    }
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) s += x[i];
    uint32_t tot;
    const uint32_t pre = block_excl_scan<kWaves>(s, s_w, tot);
    if (tid == 0) store_device_u64(&status[c], kStatusTotal | tot);
    // Totals of the chunks before this one: thread t adds chunks t, t+256,
    // ..., issuing up to 8 status loads at once so that their device-coherent
    // latency is paid once, then waiting only on unpublished ones.
    uint32_t ps = 0;
    for (uint32_t i0 = 0; i0 < c; i0 += 8u * kThreads) {
        uint64_t w[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t i = i0 + (uint32_t)j * kThreads + tid;
            w[j] = i < c ? load_device_u64(&status[i]) : kStatusTotal;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t i = i0 + (uint32_t)j * kThreads + tid;
            for (uint32_t spins = 0; w[j] == 0 && spins < kStatusSpins; ++spins) {
                wait_pause();
                w[j] = load_device_u64(&status[i]);
            }
            ps += (uint32_t)w[j];
        }
    }
    uint32_t excl;
    __syncthreads();  // s_w is reused
    block_excl_scan<kWaves>(ps, s_w, excl);
    uint32_t run = excl + pre;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const uint32_t v = x[i];
        x[i] = run;
        run += v;
    }
    if (full) {
#pragma unroll
        for (int i = 0; i < kScanItems / 4; ++i)
            *reinterpret_cast<uint4*>(out + base + off + 4 * i) =
                make_uint4(x[4 * i + 0], x[4 * i + 1], x[4 * i + 2], x[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i)
            if (base + off + i < m) out[base + off + i] = x[i];
    }
}

// One workgroup of 1024 threads scans np chunk sums in place (exclusive).
__global__ __launch_bounds__(1024) void scan_partials_kernel(uint32_t* __restrict__ part,
                                                             uint32_t np) {
    __shared__ uint32_t s_w[16];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < np; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t x = i < np ? part[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<16>(x, s_w, tot);
        if (i < np) part[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
}

// Downsweep of chunk blockIdx.x.  SUM_PARTS: `part` holds the unscanned
// chunk sums and the workgroup adds up those before its own (used when there
// are at most kScanFusedParts chunks, which saves the partials launch);
// otherwise `part` is already an exclusive scan of them.
template <bool SUM_PARTS>
__global__ __launch_bounds__(kThreads) void scan_down_kernel(const uint32_t* in, uint32_t* out,
                                                             uint32_t m,
                                                             const uint32_t* __restrict__ part) {
    __shared__ uint32_t s_w[kWaves];
    const size_t base = (size_t)blockIdx.x * kScanChunk;
    const uint32_t tid = threadIdx.x;
    const uint32_t off = tid * (uint32_t)kScanItems;
    const bool full = base + kScanChunk <= m && ((((uintptr_t)in) | ((uintptr_t)out)) & 15u) == 0;
    uint32_t x[kScanItems];
    if (full) {
#pragma unroll
        for (int i = 0; i < kScanItems / 4; ++i) {
            uint4 v = *reinterpret_cast<const uint4*>(in + base + off + 4 * i);
            x[4 * i + 0] = v.x; x[4 * i + 1] = v.y; x[4 * i + 2] = v.z; x[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) x[i] = (base + off + i < m) ? in[base + off + i] : 0u;
    }
    uint32_t base_sum;
    if constexpr (SUM_PARTS) {
        __shared__ uint32_t s_p[kWaves];
        uint32_t ps = 0;
        for (uint32_t i = tid; i < blockIdx.x; i += kThreads) ps += part[i];
        uint32_t ptot;
        block_excl_scan<kWaves>(ps, s_p, ptot);
        base_sum = ptot;
    } else {
        base_sum = part[blockIdx.x];
    }
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) s += x[i];
    uint32_t tot;
    uint32_t run = block_excl_scan<kWaves>(s, s_w, tot) + base_sum;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const uint32_t v = x[i];
        x[i] = run;
        run += v;
    }
    if (full) {
#pragma unroll
        for (int i = 0; i < kScanItems / 4; ++i)
            *reinterpret_cast<uint4*>(out + base + off + 4 * i) =
                make_uint4(x[4 * i + 0], x[4 * i + 1], x[4 * i + 2], x[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i)
            if (base + off + i < m) out[base + off + i] = x[i];
    }
}

// out[b] = G[b * num_tiles]: the first word of each bucket's row of the
// scanned bucket-major matrix is the bucket's start.
__global__ void gather_offsets_kernel(const uint32_t* __restrict__ G, uint32_t num_tiles,
                                      uint32_t num_buckets, uint32_t* __restrict__ out) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < num_buckets) out[b] = G[(size_t)b * num_tiles];
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Grid-stride streaming copy with 16-byte accesses where aligned.
__global__ __launch_bounds__(kThreads) void copy_u32_kernel(const uint32_t* __restrict__ src,
                                                            uint32_t* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * kThreads;
    const size_t t0 = (size_t)blockIdx.x * kThreads + threadIdx.x;
    const bool aligned = (((uintptr_t)src | (uintptr_t)dst) & 15u) == 0;
    if (aligned) {
        const size_t n4 = n / 4;
        const uint4* s4 = reinterpret_cast<const uint4*>(src);
        uint4* d4 = reinterpret_cast<uint4*>(dst);
        for (size_t i = t0; i < n4; i += stride) d4[i] = s4[i];
        for (size_t i = n4 * 4 + t0; i < n; i += stride) dst[i] = src[i];
    } else {
        for (size_t i = t0; i < n; i += stride) dst[i] = src[i];
    }
}

size_t scan_partials_needed(size_t m) { return (m + kScanChunk - 1) / kScanChunk + 1; }

uint32_t scan_chunks(size_t m) { return (uint32_t)((m + kScanChunk - 1) / kScanChunk); }

hipError_t launch_scan_single(const uint32_t* in, uint32_t* out, uint32_t m, uint64_t* status,
                              hipStream_t stream) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(scan_single_kernel, dim3(scan_chunks(m)), dim3(kThreads), 0, stream, in, out,
                       m, status);
    return hipGetLastError();
}

hipError_t launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t m, uint32_t* part,
                                 hipStream_t stream) {
    if (m == 0) return hipSuccess;
    const uint32_t nb = (uint32_t)((m + kScanChunk - 1) / kScanChunk);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(kThreads), 0, stream, in, m, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (nb <= kScanFusedParts) {
        hipLaunchKernelGGL((scan_down_kernel<true>), dim3(nb), dim3(kThreads), 0, stream, in, out, m,
                           (const uint32_t*)part);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(1024), 0, stream, part, nb);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((scan_down_kernel<false>), dim3(nb), dim3(kThreads), 0, stream, in, out, m,
                       (const uint32_t*)part);
    return hipGetLastError();
}

hipError_t launch_gather_offsets(const uint32_t* G, uint32_t num_tiles, uint32_t num_buckets,
                                 uint32_t* out, hipStream_t stream) {
    hipLaunchKernelGGL(gather_offsets_kernel, dim3((num_buckets + 255) / 256), dim3(256), 0, stream,
                       G, num_tiles, num_buckets, out);
    return hipGetLastError();
}

hipError_t launch_fill(uint32_t* p, uint32_t v, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, v, n);
    return hipGetLastError();
}

hipError_t launch_copy(const uint32_t* src, uint32_t* dst, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n / 4 + kThreads - 1) / kThreads;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(copy_u32_kernel, dim3((uint32_t)blocks), dim3(kThreads), 0, stream, src,
                       dst, n);
    return hipGetLastError();
}

}  // namespace gms
