// Multisplit and radix sort of elements with 64-bit keys and/or 64-bit values.
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
#include <algorithm>

#include "gms_wide_kernels.hpp"

namespace gms {

uint32_t wide_tile_elems(uint32_t key_words, uint32_t value_words) {
    return kThreads * (uint32_t)wide_items((int)key_words, (int)value_words);
}

uint32_t wide_min_tile_elems() {
    uint32_t m = wide_tile_elems(2, 0);
    for (uint32_t kw = 1; kw <= 2; ++kw)
        for (uint32_t vw = 1; vw <= 2; ++vw)
            if (kw == 2 || vw == 2) m = std::min(m, wide_tile_elems(kw, vw));
    return m;
}

template <class Mk64, class Mk32>
static gms_status wide_stages(Scratch& s, const uint32_t* kin, const uint32_t* vin,
                              uint32_t* kout, uint32_t* vout, uint32_t n, uint32_t kw,
                              uint32_t vw, uint32_t B, const Mk64& mk64, const Mk32& mk32,
                              uint32_t* d_offsets, hipStream_t stream) {
    if (n == 0) return d_offsets ? hip_status(launch_fill(d_offsets, 0u, B, stream)) : GMS_SUCCESS;
    if (B == 1) {
        hipError_t e = launch_copy(kin, kout, (size_t)n * kw, stream);
        if (e == hipSuccess && vw) e = launch_copy(vin, vout, (size_t)n * vw, stream);
        if (e == hipSuccess && d_offsets) e = launch_fill(d_offsets, 0u, 1u, stream);
        return hip_status(e);
    }
    const int lb = ceil_log2_u32(B);
    const uint32_t tile = wide_tile_elems(kw, vw);
    const uint32_t nt = (uint32_t)(((size_t)n + tile - 1) / tile);
    const size_t m = ((size_t)1 << lb) * nt;
    // One-launch scan when the scratch has its status words (the prescan
    // clears them), else the two-launch one.
    const bool one = s.status && scan_chunks(m) <= s.status_cap;
    if (m > s.hist_cap || (!one && scan_partials_needed(m) > s.part_cap)) return GMS_ERR_CAPACITY;
    auto al16 = [](const void* p) { return p == nullptr || (((uintptr_t)p) & 15u) == 0; };
    WideCall c{kin, vin, kout, vout, n, nt, s.hist, al16(kin) && al16(vin), stream};
    if (one) {
        c.scan_status = s.status;
        c.scan_chunks = scan_chunks(m);
    }
    hipError_t e = wide_dispatch(c, lb, kw, vw, mk64, mk32, true);
    if (e == hipSuccess)
        e = one ? launch_scan_single(s.hist, s.hist, (uint32_t)m, s.status, stream)
                : launch_exclusive_scan(s.hist, s.hist, (uint32_t)m, s.part, stream);
    if (e == hipSuccess) e = wide_dispatch(c, lb, kw, vw, mk64, mk32, false);
    if (e == hipSuccess && d_offsets) e = launch_gather_offsets(s.hist, nt, B, d_offsets, stream);
    return hip_status(e);
}

gms_status run_wide_multisplit(Scratch& s, const uint32_t* kin, const uint32_t* vin,
                               uint32_t* kout, uint32_t* vout, uint32_t n, uint32_t kw,
                               uint32_t vw, uint32_t B, const gms_bucket_desc64& d,
                               uint32_t* d_offsets, hipStream_t stream) {
    return wide_stages(
        s, kin, vin, kout, vout, n, kw, vw, B,
        [&] { return AnyF64::make(d.kind, d.param, d.width, B); },
        [&] { return AnyF::make(d.kind, (uint32_t)d.param, d.width, B); }, d_offsets, stream);
}

gms_status run_wide_large(Scratch& s, uint32_t* ktmp, uint32_t* vtmp, const uint32_t* kin,
                          const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint32_t n,
                          uint32_t kw, uint32_t vw, uint32_t B, const gms_bucket_desc64& d,
                          uint32_t* d_offsets, hipStream_t stream) {
    if (n == 0) return d_offsets ? hip_status(launch_fill(d_offsets, 0u, B, stream)) : GMS_SUCCESS;
    const int bits = ceil_log2_u32(B);
    const int passes = (bits + 7) / 8;
    const uint32_t* sk = kin;
    const uint32_t* sv = vin;
    for (int p = 0; p < passes; ++p) {
        const uint32_t shift = 8u * (uint32_t)p;
        const uint32_t w = (uint32_t)(bits - (int)shift < 8 ? bits - (int)shift : 8);
        const bool to_out = ((passes - 1 - p) % 2) == 0;  // the last pass writes the output
        uint32_t* dk = to_out ? kout : ktmp;
        uint32_t* dv = vw ? (to_out ? vout : vtmp) : nullptr;
        const uint32_t mask = (1u << w) - 1u;
        gms_status st = wide_stages(
            s, sk, sv, dk, dv, n, kw, vw, 1u << w,
            [&] { return DigitOf64F{AnyF64::make(d.kind, d.param, d.width, B), shift, mask}; },
            [&] { return DigitOfF{AnyF::make(d.kind, (uint32_t)d.param, d.width, B), shift, mask}; },
            nullptr, stream);
        if (st != GMS_SUCCESS) return st;
        sk = dk;
        sv = dv;
    }
    if (!d_offsets) return GMS_SUCCESS;
    if (kw == 2)
        return hip_status(launch_bucket_starts<2>(kout, n, B, AnyF64::make(d.kind, d.param, d.width, B),
                                                  d_offsets, stream));
    return hip_status(launch_bucket_starts<1>(kout, n, B,
                                              AnyF::make(d.kind, (uint32_t)d.param, d.width, B),
                                              d_offsets, stream));
}

int wide_sort_passes(uint32_t key_words, uint32_t digit_bits, SortPass* p) {
    if (key_words == 1) return sort_schedule(digit_bits, p);
    if (digit_bits < 4 || digit_bits > 8) return 0;
    int k = 0;
    for (uint32_t s = 0; s < 64; s += digit_bits)
        p[k++] = {s, 64 - s < digit_bits ? 64 - s : digit_bits};
    return k;
}

gms_status run_wide_sort(Scratch& s, uint32_t* ktmp, uint32_t* vtmp, const uint32_t* kin,
                         const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint32_t n,
                         uint32_t kw, uint32_t vw, uint32_t digit_bits, hipStream_t stream) {
    SortPass p[16];
    const int np = wide_sort_passes(kw, digit_bits, p);
    if (np == 0) return GMS_ERR_INVALID_ARG;
    const uint32_t* sk = kin;
    const uint32_t* sv = vin;
    for (int i = 0; i < np; ++i) {
        // the last pass writes the output; the input is never written
        const bool to_out = ((np - 1 - i) % 2) == 0;
        uint32_t* dk = to_out ? kout : ktmp;
        uint32_t* dv = vw ? (to_out ? vout : vtmp) : nullptr;
        const gms_bucket_desc64 d{GMS_BUCKET_DIGIT, p[i].bits, p[i].shift};
        gms_status st = run_wide_multisplit(s, sk, sv, dk, dv, n, kw, vw, 1u << p[i].bits, d,
                                            nullptr, stream);
        if (st != GMS_SUCCESS) return st;
        sk = dk;
        sv = dv;
    }
    return GMS_SUCCESS;
}

}  // namespace gms
