// Host-side engine shared by the C-ABI entry points.
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/gpumultisplit.h"
#include "gms_device.hpp"
#include "gms_tune.hpp"

namespace gms {

// Arguments of one plain multisplit launch sequence.
struct MsCall {
    const uint32_t* kin;
    const uint32_t* vin;  // NULL for key-only
    uint32_t* kout;
    uint32_t* vout;
    uint32_t n;
    int log_b;
    bool kv;
    bool large;  // n >= long_tile_min(log_b, kv): the long-tile shape of the class
    uint32_t num_tiles;
    uint32_t* H;  // histogram matrix, num_buckets(pow2 of log_b) * num_tiles
    int vec_ok;
    hipStream_t stream;
    // Status words of the single-pass scan that follows, cleared by the
    // prescan (NULL: the scan that follows is the two-launch one).
    uint64_t* scan_status = nullptr;
    uint32_t scan_chunks = 0;
};

// Functor-family launchers (one translation unit per family).
hipError_t launch_prescan(const MsCall& c, const RangeF& f);
hipError_t launch_prescan(const MsCall& c, const DeltaF& f);
hipError_t launch_prescan(const MsCall& c, const DigitF& f);
hipError_t launch_prescan(const MsCall& c, const IdentF& f);
hipError_t launch_prescan(const MsCall& c, const DigitOfF& f);
hipError_t launch_postscan(const MsCall& c, const RangeF& f);
hipError_t launch_postscan(const MsCall& c, const DeltaF& f);
hipError_t launch_postscan(const MsCall& c, const DigitF& f);
hipError_t launch_postscan(const MsCall& c, const IdentF& f);
hipError_t launch_postscan(const MsCall& c, const DigitOfF& f);

// Exclusive scan of m u32 (in may equal out); part holds scan_partials_needed(m).
hipError_t launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t m, uint32_t* part,
                                 hipStream_t stream);
size_t scan_partials_needed(size_t m);
// The same scan in one launch.  `status` holds scan_chunks(m) 64-bit words
// that must be zero when it starts (the plain path's prescan clears them);
// scans sharing them must not overlap in time.
uint32_t scan_chunks(size_t m);
hipError_t launch_scan_single(const uint32_t* in, uint32_t* out, uint32_t m, uint64_t* status,
                              hipStream_t stream);
hipError_t launch_gather_offsets(const uint32_t* G, uint32_t num_tiles, uint32_t num_buckets,
                                 uint32_t* out, hipStream_t stream);
hipError_t launch_fill(uint32_t* p, uint32_t v, uint32_t n, hipStream_t stream);
hipError_t launch_copy(const uint32_t* src, uint32_t* dst, size_t n, hipStream_t stream);
hipError_t launch_histogram_u32(const uint32_t* kin, uint32_t n, uint32_t* d_hist, uint32_t nbins,
                                gms_bucket_desc d, hipStream_t stream);
hipError_t launch_histogram_even_f32(const float* x, uint32_t n, int32_t* d_hist, uint32_t nbins,
                                     float lower, float inv_delta, hipStream_t stream);
hipError_t launch_histogram_u64(const uint64_t* keys, uint32_t n, uint32_t* d_hist,
                                uint32_t nbins, const AnyF64& f, hipStream_t stream);
hipError_t launch_histogram_range_f32(const float* x, uint32_t n, int32_t* d_hist,
                                      uint32_t num_levels, const float* d_levels,
                                      hipStream_t stream);
hipError_t launch_histogram_range_u32(const uint32_t* x, uint32_t n, int32_t* d_hist,
                                      uint32_t num_levels, const uint32_t* d_levels,
                                      hipStream_t stream);

// Device scratch of a multisplit: the histogram matrix and the scan partials.
struct Scratch {
    uint32_t* hist = nullptr;
    size_t hist_cap = 0;  // u32 elements
    uint32_t* part = nullptr;
    size_t part_cap = 0;
    uint32_t* offs = nullptr;  // 256 bucket starts (staging for host offsets)
    uint64_t* status = nullptr;  // single-pass scan statuses of the plain path
    size_t status_cap = 0;       // 64-bit words
    int device = -1;
};

// Multisplits into more than 256 buckets run as a stable radix sort of the
// bucket ids (gms_large.hip); up to this many buckets.
constexpr uint32_t kMaxBuckets = 1u << 20;

struct LargeScratch {
    uint32_t* ktmp = nullptr;  // keys between passes
    uint32_t* vtmp = nullptr;  // values between passes
    uint32_t* offs = nullptr;  // num_buckets bucket starts (staging for host offsets)
    uint32_t cap = 0;
};
// offs (num_buckets words) always; the two capacity-word pass temporaries
// only with `temporaries` (bucket counts beyond one pass).
gms_status large_scratch_alloc(LargeScratch& L, uint32_t n_capacity, uint32_t num_buckets,
                               bool temporaries);
void large_scratch_free(LargeScratch& L);
gms_status run_large_multisplit(LargeScratch& L, Scratch& s, const uint32_t* kin,
                                const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint32_t n,
                                uint32_t num_buckets, gms_bucket_desc d, uint32_t* d_offsets,
                                hipStream_t stream);

gms_status scratch_alloc(Scratch& s, size_t hist_elems);
void scratch_free(Scratch& s);
// Matrix elements needed for n keys in B buckets (key-only and key-value tiles).
size_t matrix_elems(uint32_t n, uint32_t num_buckets);

// Per-call stage events (prescan, scan, postscan) when profiling is on.
struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> ev;  // 4 per recorded call
    size_t used = 0;             // events in use
    double ms[3] = {0, 0, 0};
    uint32_t calls = 0;
    hipEvent_t* next(hipStream_t s);  // 4 fresh events (flushes when full)
    gms_status flush();
    void destroy();
};

gms_status check_desc(uint32_t num_buckets, gms_bucket_desc d);

// Full prescan -> scan -> postscan on the given scratch.  d_offsets (device,
// num_buckets entries) optional.  Validation is the caller's job.
gms_status run_multisplit(Scratch& s, const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                          uint32_t* vout, uint32_t n, uint32_t num_buckets, gms_bucket_desc d,
                          uint32_t* d_offsets, hipStream_t stream, Profiler* prof);

// The same with one digit of a wide bucket id as the identifier (B = 2^width
// of the digit, at most 256).
gms_status run_multisplit_digit(Scratch& s, const uint32_t* kin, const uint32_t* vin,
                                uint32_t* kout, uint32_t* vout, uint32_t n, uint32_t num_buckets,
                                const DigitOfF& f, hipStream_t stream);

// Prescan only (for the stage-level check).
gms_status run_prescan_only(Scratch& s, const uint32_t* kin, uint32_t n, uint32_t num_buckets,
                            gms_bucket_desc d, bool kv, hipStream_t stream, uint32_t* num_tiles);

// LSD digit schedule of the radix sort for digit_bits in {4..8}: the
// reference's passes (multisplit_sort.cuh:320-559), least significant first.
// Returns the pass count (0 for an unsupported width); p holds up to 8.
struct SortPass {
    uint32_t shift, bits;
};
int sort_schedule(uint32_t digit_bits, SortPass* p);

// Elements with 64-bit keys and/or values (gms_wide.hip): key_words 1 or 2,
// value_words 0, 1 or 2, not both below 2; up to 256 buckets.
uint32_t wide_tile_elems(uint32_t key_words, uint32_t value_words);
// The shortest of those tiles over every supported element shape.
uint32_t wide_min_tile_elems();
gms_status run_wide_multisplit(Scratch& s, const uint32_t* kin, const uint32_t* vin,
                               uint32_t* kout, uint32_t* vout, uint32_t n, uint32_t key_words,
                               uint32_t value_words, uint32_t num_buckets,
                               const gms_bucket_desc64& d, uint32_t* d_offsets,
                               hipStream_t stream);
// The same into more than 256 buckets (up to kMaxBuckets): LSD passes over
// 8-bit digits of the bucket id, ping-ponging through ktmp/vtmp (n x key_words
// and n x value_words words).
gms_status run_wide_large(Scratch& s, uint32_t* ktmp, uint32_t* vtmp, const uint32_t* kin,
                          const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint32_t n,
                          uint32_t key_words, uint32_t value_words, uint32_t num_buckets,
                          const gms_bucket_desc64& d, uint32_t* d_offsets, hipStream_t stream);
// Passes of the radix sort of key_words-word keys (p holds up to 16).
int wide_sort_passes(uint32_t key_words, uint32_t digit_bits, SortPass* p);
// Descriptor check for keys of key_bits (32 or 64) bits.
gms_status check_desc64(uint32_t num_buckets, const gms_bucket_desc64& d, uint32_t key_bits);
gms_status run_wide_sort(Scratch& s, uint32_t* ktmp, uint32_t* vtmp, const uint32_t* kin,
                         const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint32_t n,
                         uint32_t key_words, uint32_t value_words, uint32_t digit_bits,
                         hipStream_t stream);

// Wave-level multisplit (gms_wms.hip): 2..64 buckets, one subproblem of
// wms_subproblem_keys() keys per wave.
constexpr uint32_t kWmsMaxBuckets = 64;
uint32_t wms_subproblem_keys();
// direct: the DMS postscan (keys scattered from registers, no LDS reorder).
gms_status run_wms_multisplit(Scratch& s, const uint32_t* kin, const uint32_t* vin,
                              uint32_t* kout, uint32_t* vout, uint32_t n, uint32_t num_buckets,
                              gms_bucket_desc d, uint32_t* d_offsets, bool direct,
                              hipStream_t stream);

gms_status hip_status(hipError_t e);
bool ranges_overlap(const void* a, size_t an, const void* b, size_t bn);

}  // namespace gms
