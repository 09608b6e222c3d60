// Multisplit into more than 256 buckets (up to kMaxBuckets).
//
// The reference's answer beyond 256 buckets is the reduced-bit sort (README,
// "Multisplit" item; src/kernels/RBsort/reduced_bit_sort.cuh:24-57 and its use
// in src/main/main_wms.cu:534-755): rank the elements by the bits of their
// bucket id with a stable radix sort, which equals the multisplit.  Here the
// id is never stored: LSD passes of this library's own multisplit rank by
// 8-bit digits of f(key) (DigitOfF, computed from the key in every pass),
// carrying the keys and values themselves, ping-ponging through scratch so
// that the last pass writes the output and the input is never written.  The
// bucket starts, when asked, are read off the output: start[b] = first i with
// f(kout[i]) >= b.
#include "gms_internal.hpp"
#include "gms_kernels.hpp"

namespace gms {

gms_status large_scratch_alloc(LargeScratch& L, uint32_t n_capacity, uint32_t num_buckets,
                               bool temporaries) {
    large_scratch_free(L);
    const size_t words = std::max<size_t>(n_capacity, 1);
    if (temporaries) {
        if (hipMalloc((void**)&L.ktmp, words * sizeof(uint32_t)) != hipSuccess) L.ktmp = nullptr;
        if (hipMalloc((void**)&L.vtmp, words * sizeof(uint32_t)) != hipSuccess) L.vtmp = nullptr;
    }
    if (hipMalloc((void**)&L.offs, (size_t)num_buckets * sizeof(uint32_t)) != hipSuccess)
        L.offs = nullptr;
    if ((temporaries && (!L.ktmp || !L.vtmp)) || !L.offs) {
        large_scratch_free(L);
        return GMS_ERR_ALLOC;
    }
    L.cap = n_capacity;
    return GMS_SUCCESS;
}

void large_scratch_free(LargeScratch& L) {
    uint32_t** bufs[3] = {&L.ktmp, &L.vtmp, &L.offs};
    for (uint32_t** b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    L.cap = 0;
}

gms_status run_large_multisplit(LargeScratch& L, Scratch& s, const uint32_t* kin,
                                const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint32_t n,
                                uint32_t B, gms_bucket_desc d, uint32_t* d_offsets,
                                hipStream_t stream) {
    if (n == 0) return d_offsets ? hip_status(launch_fill(d_offsets, 0u, B, stream)) : GMS_SUCCESS;
    if (n > L.cap || !L.ktmp || !L.vtmp) return GMS_ERR_CAPACITY;
    const AnyF base = AnyF::make(d.kind, d.param, d.width, B);
    const bool kv = vin != nullptr;
    const int bits = ceil_log2_u32(B);
    const int passes = (bits + 7) / 8;
    const uint32_t* sk = kin;
    const uint32_t* sv = vin;
    for (int p = 0; p < passes; ++p) {
        const uint32_t shift = 8u * (uint32_t)p;
        const uint32_t w = (uint32_t)(bits - (int)shift < 8 ? bits - (int)shift : 8);
        const bool to_out = ((passes - 1 - p) % 2) == 0;  // the last pass writes the output
        uint32_t* dk = to_out ? kout : L.ktmp;
        uint32_t* dv = kv ? (to_out ? vout : L.vtmp) : nullptr;
        const DigitOfF f{base, shift, (1u << w) - 1u};
        gms_status st = run_multisplit_digit(s, sk, sv, dk, dv, n, 1u << w, f, stream);
        if (st != GMS_SUCCESS) return st;
        sk = dk;
        sv = dv;
    }
    if (!d_offsets) return GMS_SUCCESS;
    return hip_status(launch_bucket_starts<1>(kout, n, B, base, d_offsets, stream));
}

}  // namespace gms
