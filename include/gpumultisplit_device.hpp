// gpumultisplit_device.hpp -- multisplit and compaction with the caller's own
// bucket functor, compiled into the caller's code object.
//
// The reference's bucket function "is provided by the programmer"
// (README.md:4): multisplit_key_only / multisplit_key_value
// (src/api/multisplit.cuh:81-88, :199-208) and compaction_key_only /
// compaction_key_value (src/kernels/compaction/multisplit2_compaction.cuh:
// 525-573) take any functor type with
//     __device__ uint32_t operator()(KeyT key) const
// and compile it into their kernels.  A functor cannot cross the C ABI of
// libgpumultisplit.so, so this header does what the reference's header does:
// it instantiates the library's own prescan and postscan kernel templates
// (gpumultisplit_amd/csrc, the same code the library's built-in identifiers
// run) with the caller's functor, in the caller's translation unit, and lets
// the library do the rest through gms_multisplit_stage_plan / gms_stage_scan /
// gms_stage_offsets (the scratch, the scan stage, the bucket starts).
//
// Use: include this header instead of gpumultisplit.hpp and compile with
//     hipcc --offload-arch=gfx950 -std=c++17 -I<repo>/include
//           -Xclang -target-feature -Xclang -packed-fp32-ops ...
// linking libgpumultisplit.so.  Then the C++ API's multisplit_key_only,
// multisplit_key_value, compaction_key_only and compaction_key_value accept
// any functor without a gms::bucket_traits descriptor; functors with one
// (the built-in identifiers) still take the library's precompiled kernels.
// Ids the functor returns at or above the bucket count go to the last bucket
// (the built-in identifiers clamp the same way; the reference leaves it
// undefined).  Limits: up to 1024 buckets for 32-bit keys and values, up to
// 256 when a key or value is 64-bit (the reference's API stops at 256).
#ifndef GPUMULTISPLIT_DEVICE_HPP_
#define GPUMULTISPLIT_DEVICE_HPP_

#ifdef GPUMULTISPLIT_HPP_
#error "include gpumultisplit_device.hpp before (instead of) gpumultisplit.hpp"
#endif
#define GMS_DEVICE_FUNCTORS 1

#ifndef GMS_HOST_EMULATION  // the CPU emulation of the kernels brings its own runtime
#include <hip/hip_runtime.h>
#endif

#include <type_traits>

#include "gpumultisplit.h"
#include "../gpumultisplit_amd/csrc/gms_launch.hpp"
#include "../gpumultisplit_amd/csrc/gms_wide_kernels.hpp"

namespace gms {
namespace device {

// The caller's functor with its ids clamped into [0, B): an id past the
// bucket count would otherwise index the kernels' per-bucket tables.
template <class F, class K>
struct clamped {
    F f;
    uint32_t bmax;
    __host__ __device__ __forceinline__ uint32_t operator()(K k) const {
        const uint32_t b = (uint32_t)f(k);
        return b < bmax ? b : bmax;
    }
};

template <class F>
inline hipError_t plain_prescan(const MsCall& c, const F& f) {
    GMS_SWITCH_LOGB_KV(launch_prescan_t, c, f)
}
template <class F>
inline hipError_t plain_postscan(const MsCall& c, const F& f) {
    GMS_SWITCH_LOGB_KV(launch_postscan_t, c, f)
}

// The stages of a planned call with functor f: 32-bit elements on the
// kernels of gms_kernels.hpp, others on the 64-bit ones (gms_wide_kernels.hpp).
template <class F>
inline gms_status run_stages(const gms_stage_plan& p, const void* kin, const void* vin, void* kout,
                             void* vout, F f, hipStream_t stream) {
    if (p.done) return GMS_SUCCESS;
    hipError_t e;
    const uint32_t bmax = p.num_buckets - 1u;
    const uint32_t* ki = static_cast<const uint32_t*>(kin);
    const uint32_t* vi = static_cast<const uint32_t*>(vin);
    uint32_t* ko = static_cast<uint32_t*>(kout);
    uint32_t* vo = static_cast<uint32_t*>(vout);
    if (p.key_words == 1 && p.value_words <= 1) {
        const clamped<F, uint32_t> g{f, bmax};
        const MsCall c{ki, vi, ko, vo, p.n, (int)p.log_b, p.value_words != 0, p.large != 0,
                       p.num_tiles, p.d_matrix, (int)p.vec, stream};
        e = plain_prescan(c, g);
        if (e != hipSuccess) return GMS_ERR_HIP;
        const gms_status st = gms_stage_scan(&p, stream);
        if (st != GMS_SUCCESS) return st;
        e = plain_postscan(c, g);
    } else {
        const WideCall c{ki, vi, ko, vo, p.n, p.num_tiles, p.d_matrix, p.vec != 0, stream};
        const int lb = (int)p.log_b;
        auto stage = [&](bool pre) -> hipError_t {
            if (p.key_words == 2) {
                const clamped<F, uint64_t> g{f, bmax};
                if (p.value_words == 0) return wide_stage<2, 0>(c, lb, g, pre);
                if (p.value_words == 1) return wide_stage<2, 1>(c, lb, g, pre);
                return wide_stage<2, 2>(c, lb, g, pre);
            }
            const clamped<F, uint32_t> g{f, bmax};
            return wide_stage<1, 2>(c, lb, g, pre);
        };
        e = stage(true);
        if (e != hipSuccess) return GMS_ERR_HIP;
        const gms_status st = gms_stage_scan(&p, stream);
        if (st != GMS_SUCCESS) return st;
        e = stage(false);
    }
    return e == hipSuccess ? GMS_SUCCESS : GMS_ERR_HIP;
}

template <class T>
constexpr uint32_t bytes_of() {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "keys and values must be 32- or 64-bit");
    return (uint32_t)sizeof(T);
}

// Multisplit (d_value_in == nullptr: key-only).  h_offsets: the num_buckets
// bucket starts on the host (the call then blocks), as the reference.
template <class KeyT, class ValueT, class F>
inline gms_status multisplit(gms_multisplit_ctx* ctx, const KeyT* kin, const ValueT* vin, KeyT* kout,
                             ValueT* vout, uint32_t n, F f, uint32_t* h_offsets, void* stream) {
    const uint32_t vb = vin || vout ? bytes_of<ValueT>() : 0u;
    hipStream_t s = (hipStream_t)stream;
    gms_stage_plan p;
    gms_status st = gms_multisplit_stage_plan(ctx, kin, vin, kout, vout, n, bytes_of<KeyT>(), vb,
                                              stream, &p);
    if (st != GMS_SUCCESS) return st;
    st = run_stages(p, kin, vin, kout, vout, f, s);
    if (st != GMS_SUCCESS || !h_offsets) return st;
    st = gms_stage_offsets(&p, p.d_staging, stream);
    if (st != GMS_SUCCESS) return st;
    if (hipMemcpyAsync(h_offsets, p.d_staging, (size_t)p.num_buckets * sizeof(uint32_t),
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return GMS_ERR_HIP;
    return GMS_SUCCESS;
}

// Compaction: the stable 2-bucket multisplit, bucket 0 first; *count gets the
// bucket-0 count (blocking, as the reference).
template <class KeyT, class ValueT, class F>
inline gms_status compaction(gms_compaction_ctx* ctx, const KeyT* kin, const ValueT* vin, KeyT* kout,
                             ValueT* vout, uint32_t n, F f, unsigned* count, void* stream) {
    const uint32_t vb = vin || vout ? bytes_of<ValueT>() : 0u;
    hipStream_t s = (hipStream_t)stream;
    gms_stage_plan p;
    gms_status st = gms_compaction_stage_plan(ctx, kin, vin, kout, vout, n, bytes_of<KeyT>(), vb,
                                              stream, &p);
    if (st != GMS_SUCCESS) return st;
    st = run_stages(p, kin, vin, kout, vout, f, s);
    if (st != GMS_SUCCESS) return st;
    st = gms_stage_offsets(&p, p.d_staging, stream);
    if (st != GMS_SUCCESS) return st;
    uint32_t c = 0;
    if (hipMemcpyAsync(&c, p.d_staging + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return GMS_ERR_HIP;
    *count = c;
    return GMS_SUCCESS;
}

}  // namespace device
}  // namespace gms

#include "gpumultisplit.hpp"

#endif  // GPUMULTISPLIT_DEVICE_HPP_
