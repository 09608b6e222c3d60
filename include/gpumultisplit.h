/*
 * gpumultisplit.h -- C ABI of the MI355X (gfx950) multisplit library.
 *
 * libgpumultisplit.so exports exactly the functions declared here.  They carry
 * plain pointers and sizes only: device pointers are hipMalloc'd memory owned
 * by the caller, `stream` is a hipStream_t passed as void* (NULL = the default
 * stream), host pointers are ordinary host memory.  Every function returns a
 * gms_status; none throws.  The C++ header gpumultisplit.hpp rebuilds the
 * reference's C++ API (multisplit_context, multisplit_key_value, ...) on top.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * the reference repository).
 */
#ifndef GPUMULTISPLIT_H_
#define GPUMULTISPLIT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMS_ABI_VERSION 1

typedef enum gms_status {
    GMS_SUCCESS = 0,
    GMS_ERR_INVALID_ARG = 1,        /* null/aliased pointer, bad descriptor     */
    GMS_ERR_UNSUPPORTED_BUCKETS = 2,/* bucket count outside the supported range */
    GMS_ERR_CAPACITY = 3,           /* n larger than the allocated capacity     */
    GMS_ERR_NOT_ALLOCATED = 4,      /* call before *_allocate                   */
    GMS_ERR_ALLOC = 5,              /* device allocation failed                 */
    GMS_ERR_HIP = 6,                /* a HIP runtime call failed                */
    GMS_ERR_ALIAS = 7,              /* input and output ranges overlap          */
    GMS_ERR_NO_DEVICE = 8           /* no HIP device visible                    */
} gms_status;

/* Bucket identifiers.  A functor object cannot cross a C ABI, so the built-in
 * identifiers of the reference are enumerated; `param` is their one runtime
 * parameter.  Ids that fall outside [0, B) -- undefined behaviour in the
 * reference -- are clamped into bucket B-1. */
typedef enum gms_bucket_kind {
    GMS_BUCKET_RANGE = 0,     /* floor(key * B / 2^32); param unused            */
    GMS_BUCKET_DELTA = 1,     /* key / param           (delta_func,
                                 src/main/main_multisplit.cu:41-47)           */
    GMS_BUCKET_DIGIT = 2,     /* (key >> param) & (2^width - 1); width 0 means
                                 log2(B) for a power-of-two B; 2^width <= B,
                                 width <= 20
                                 (ms_bucket_sort_Nbit<x>,
                                 src/api/multisplit_sort.cuh:63-98)           */
    GMS_BUCKET_IDENTITY = 3,  /* key                   (identity_bucket,
                                 src/main/main_multisplit.cu:34-39)           */
    GMS_BUCKET_LOWBITS = 4    /* key & ((1<<param)-1)  (test_bucket<flag>,
                                 src/main/main_compaction.cu:27-32)           */
} gms_bucket_kind;

typedef struct gms_bucket_desc {
    uint32_t kind;    /* gms_bucket_kind */
    uint32_t param;   /* delta, shift or flag */
    uint32_t width;   /* digit width in bits (GMS_BUCKET_DIGIT only) */
} gms_bucket_desc;

/* Identifier of a 64-bit key (or of a 32-bit key whose values are 64-bit):
 * the same kinds, with a 64-bit delta / shift / flag.  On 64-bit keys
 *   RANGE    floor(key * B / 2^64)
 *   DELTA    key / param
 *   DIGIT    (key >> param) & (2^width - 1), or & (B - 1) when width is 0
 *   IDENTITY key
 *   LOWBITS  key & (2^param - 1)
 * each clamped into [0, B), as the 32-bit forms (the reference's functors
 * written for KeyT = uint64_t, src/main/main_multisplit.cu:34-47). */
typedef struct gms_bucket_desc64 {
    uint32_t kind;    /* gms_bucket_kind */
    uint32_t width;   /* digit width in bits (GMS_BUCKET_DIGIT only) */
    uint64_t param;   /* delta, shift or flag */
} gms_bucket_desc64;

/* Short description of a status code (static storage). */
const char *gms_status_string(gms_status s);

/* ABI version of the loaded library (GMS_ABI_VERSION it was built with). */
uint32_t gms_abi_version(void);

/* Bucket ids evaluated on the host with the same integer code the kernels run
 * (used to check the magic-number division of GMS_BUCKET_DELTA on the CPU). */
gms_status gms_bucket_ids_host(uint32_t num_buckets, gms_bucket_desc desc,
                               const uint32_t *keys, uint32_t *ids, size_t n);
/* The same for 64-bit keys and gms_bucket_desc64 (up to 256 buckets). */
gms_status gms_bucket_ids64_host(uint32_t num_buckets, gms_bucket_desc64 desc,
                                 const uint64_t *keys, uint32_t *ids, size_t n);

/* ------------------------------------------------------------------------ */
/* Multisplit                                                                */
/* ------------------------------------------------------------------------ */

typedef struct gms_multisplit_ctx gms_multisplit_ctx;

/* multisplit_context(uint32_t n_buckets) (src/api/multisplit.cuh:29-47).
 * 1 <= num_buckets <= 2^20.  Up to 1024 buckets the multisplit of 32-bit
 * elements runs as one prescan -> scan -> postscan (the reference's API
 * accepts powers of two 2..256); above that as a stable radix sort of the
 * bucket ids, the reduced-bit sort the reference points to for more buckets
 * (src/kernels/RBsort/reduced_bit_sort.cuh:24-57, src/main/main_wms.cu:
 * 534-755). */
gms_status gms_multisplit_create(uint32_t num_buckets, gms_multisplit_ctx **out);

/* multisplit_allocate_key_only(n, ctx)  (src/api/multisplit.cuh:50-61) with
 * key_value = 0, multisplit_allocate_key_value(n, ctx) (:63-74) with 1.
 * Allocates the per-tile histogram matrix and scan scratch for any later call
 * with n <= n_capacity (the reference freezes the grid at this n). */
gms_status gms_multisplit_allocate(gms_multisplit_ctx *ctx, uint32_t n_capacity,
                                   int key_value);

/* multisplit_release_memory(ctx) (src/api/multisplit.cuh:76-79). */
gms_status gms_multisplit_release(gms_multisplit_ctx *ctx);

/* Frees the context object itself (the reference's destructor is a no-op). */
gms_status gms_multisplit_destroy(gms_multisplit_ctx *ctx);

/* multisplit_key_only(d_key_in, d_key_out, n, ctx, f, h_offsets)
 * (src/api/multisplit.cuh:81-197).  Stable.  If h_offsets != NULL it receives
 * the B global bucket starts and the call blocks until the result is ready;
 * otherwise the call is asynchronous on `stream`. */
gms_status gms_multisplit_key_only(gms_multisplit_ctx *ctx, const uint32_t *d_key_in,
                                   uint32_t *d_key_out, uint32_t n,
                                   gms_bucket_desc desc, uint32_t *h_offsets,
                                   void *stream);

/* multisplit_key_value(d_key_in, d_value_in, d_key_out, d_value_out, n, ctx, f,
 * h_offsets) (src/api/multisplit.cuh:199-313).  Values are 32-bit payloads. */
gms_status gms_multisplit_key_value(gms_multisplit_ctx *ctx, const uint32_t *d_key_in,
                                    const uint32_t *d_value_in, uint32_t *d_key_out,
                                    uint32_t *d_value_out, uint32_t n,
                                    gms_bucket_desc desc, uint32_t *h_offsets,
                                    void *stream);

/* Same as the two above (d_value_in/out NULL = key-only) but writes the B
 * bucket starts to device memory d_offsets without blocking.  Replaces the
 * reference's B-1 blocking 4-byte copies (src/api/multisplit.cuh:144-149). */
gms_status gms_multisplit(gms_multisplit_ctx *ctx, const uint32_t *d_key_in,
                          const uint32_t *d_value_in, uint32_t *d_key_out,
                          uint32_t *d_value_out, uint32_t n, gms_bucket_desc desc,
                          uint32_t *d_offsets, void *stream);

/* Size of the device scratch of gms_multisplit_noalloc, one copy per GPU. */
#define GMS_NOALLOC_SCRATCH_BYTES (32u << 20)

/* No counterpart in the reference.  gms_multisplit without a context, for
 * callers that cannot allocate (code under hipGraph capture, serving loops,
 * harnesses that forbid hipMalloc): it allocates nothing, on the device or
 * the host.  Its scratch (the per-tile count matrix and its scan state) is
 * a file-scope device array of GMS_NOALLOC_SCRATCH_BYTES in the library, one
 * copy on each GPU, held until the process ends.  Up to 1024 buckets; the
 * call needs 2^ceil(log2 B) * ceil(n / gms_tile_keys(B, key_value, n)) words
 * of matrix, and returns GMS_ERR_CAPACITY beyond the array less 16.5 KiB of
 * scan state (2^27 key-value elements in 256 buckets take 19 MiB).  Same results as gms_multisplit.  Every call on one GPU
 * shares that GPU's array: calls that overlap in time on one GPU (two
 * streams, two host threads) race and give wrong results; calls on one
 * stream are ordered and safe. */
gms_status gms_multisplit_noalloc(const uint32_t *d_key_in, const uint32_t *d_value_in,
                                  uint32_t *d_key_out, uint32_t *d_value_out, uint32_t n,
                                  uint32_t num_buckets, gms_bucket_desc desc,
                                  uint32_t *d_offsets, void *stream);

/* Multisplit of 32- or 64-bit keys carrying 32- or 64-bit values: the
 * reference's multisplit_key_only<KeyT> / multisplit_key_value<KeyT, ValueT>
 * (src/api/multisplit.cuh:81-197, :199-313) for KeyT, ValueT of 4 or 8 bytes.
 * key_bytes 4 or 8; value_bytes 0 (key-only), 4 or 8.  Arrays of 8-byte
 * elements must be 8-byte aligned.  With 4-byte keys and values (or none) this
 * is gms_multisplit / gms_multisplit_key_value; any 8-byte operand takes the
 * 64-bit kernels: one pass up to 256 buckets, LSD passes over 8-bit digits of
 * the bucket id above (up to 2^20 buckets; the first such call on a context
 * allocates its pass temporaries, 2 x capacity x 8 bytes).  Bucket starts go to
 * d_offsets (device, asynchronous) and/or h_offsets (host; the call then
 * blocks), num_buckets entries each.  The context's allocation serves both. */
gms_status gms_multisplit_wide(gms_multisplit_ctx *ctx, const void *d_key_in,
                               const void *d_value_in, void *d_key_out, void *d_value_out,
                               uint32_t n, uint32_t key_bytes, uint32_t value_bytes,
                               gms_bucket_desc64 desc, uint32_t *d_offsets, uint32_t *h_offsets,
                               void *stream);

/* Kernel family of a context's multisplits (up to 64 buckets).  BMS: the
 * block-wide kernels, one tile per workgroup (the reference's BMS path,
 * src/kernels/bms/).  WMS: the wave-level kernels, one subproblem per wave
 * ranked with wave ballots (the reference's warp-wide path,
 * src/kernels/wms/wms_prescan.cuh:21-149, wms_postscan.cuh:20-327,
 * wms_postscan_pairs.cuh:29-356, which its driver main_wms.cu selects by
 * hand).  DMS: the wave-level kernels whose postscan writes every key from
 * its register to its output slot, without the LDS reorder (the reference's
 * DMS_postscan_4rolls_protected, wms_postscan.cuh:328-505, and
 * wms_postscan_pairs.cuh:357-538).  AUTO (the default) is BMS: on MI355X it
 * measured as fast or faster than WMS at every bucket count and size (WMS at
 * most 5 % ahead, at 2 buckets and up to 2^20 keys). */
typedef enum gms_variant {
    GMS_VARIANT_AUTO = 0,
    GMS_VARIANT_BMS = 1,
    GMS_VARIANT_WMS = 2,
    GMS_VARIANT_DMS = 3
} gms_variant;
gms_status gms_multisplit_set_variant(gms_multisplit_ctx *ctx, int variant);

/* subproblem_size_wms_key_only / _key_value(num_buckets, size_block&)
 * (src/api/wms_api.h:25-83): keys per wave subproblem of the wave-level
 * variant, and keys per workgroup in *size_block; 0 outside 2..64 buckets. */
uint32_t gms_subproblem_size_wms(uint32_t num_buckets, int key_value, uint32_t *size_block);

/* subproblem_size_bms_key_only / _key_value (src/api/bms_api.h:23-89): keys
 * per tile for a bucket count (the tile of short inputs, also the shortest
 * tile of any input).  Returns 0 for an unsupported count. */
uint32_t gms_subproblem_size(uint32_t num_buckets, int key_value);

/* Keys per tile of an n-key call (inputs of 2^26 keys and more, 2^24 for
 * 1024-bucket key-value, may use a longer tile); no counterpart in the
 * reference, whose tile depends on the bucket count only.  Returns 0 for an
 * unsupported count. */
uint32_t gms_tile_keys(uint32_t num_buckets, int key_value, uint32_t n);

/* Stage timing for measurement: when enabled, every multisplit on this context
 * records HIP events on its stream around the prescan, scan and postscan
 * kernels.  gms_multisplit_profile synchronises and returns the summed
 * milliseconds and the number of recorded calls since the last reset. */
gms_status gms_multisplit_set_profiling(gms_multisplit_ctx *ctx, int enable);
gms_status gms_multisplit_profile(gms_multisplit_ctx *ctx, double *ms_prescan,
                                  double *ms_scan, double *ms_postscan,
                                  uint32_t *calls);

/* Stage-level check (up to 1024 buckets): runs only the prescan on d_key_in and copies the per-tile
 * bucket counts, bucket-major (H[b * tiles + t]), to d_matrix (B * tiles u32,
 * tiles = ceil(n / gms_tile_keys(B, key_value, n))).  Same quantity the reference's
 * cpu_prescan_test checks (src/cpu_functions.h:55-103). */
gms_status gms_multisplit_prescan_matrix(gms_multisplit_ctx *ctx, const uint32_t *d_key_in,
                                         uint32_t n, gms_bucket_desc desc, int key_value,
                                         uint32_t *d_matrix, void *stream);

/* Device-wide exclusive prefix sum of n u32 (d_in == d_out allowed), the scan
 * stage that replaces cub::DeviceScan::ExclusiveSum (src/api/multisplit.cuh:141,
 * :258).  Uses the context's scratch; n <= the context's matrix capacity. */
gms_status gms_exclusive_scan_u32(gms_multisplit_ctx *ctx, const uint32_t *d_in,
                                  uint32_t *d_out, uint32_t n, void *stream);

/* Stages of a multisplit whose bucket identifier is the caller's own functor.
 * The reference takes any `bucket_t` functor object (src/api/multisplit.cuh:
 * 81-88, :199-208, compaction.cuh:525-573; "provided by the programmer",
 * README.md:4), and a functor cannot cross a C ABI: the header-only
 * gpumultisplit_device.hpp, compiled by the caller's hipcc, instantiates this
 * library's prescan and postscan kernel templates with the functor and
 * launches them on the scratch a plan describes; these calls do the rest.
 *   gms_multisplit_stage_plan  validate the call as gms_multisplit_wide does
 *                              and describe its stages (or finish it: n == 0
 *                              or one bucket, `done` = 1);
 *   (caller: prescan)          gms_stage_scan  (caller: postscan)
 *   gms_stage_offsets          bucket starts of the finished call.
 * 32-bit elements up to 1024 buckets, elements with a 64-bit key or value up
 * to 256 (GMS_ERR_UNSUPPORTED_BUCKETS above). */
typedef struct gms_stage_plan {
    uint32_t num_buckets;
    uint32_t log_b;          /* the count matrices have 2^log_b rows */
    uint32_t key_words;      /* 32-bit words per key (1 or 2) */
    uint32_t value_words;    /* 32-bit words per value (0, 1 or 2) */
    uint32_t n;
    uint32_t num_tiles;      /* tiles of the kernels' tile shape */
    uint32_t large;          /* n takes the long-tile shape (32-bit elements) */
    uint32_t vec;            /* the inputs are 16-byte aligned */
    uint32_t done;           /* the library finished the call: nothing to launch */
    uint32_t *d_matrix;      /* per-tile bucket counts, scanned in place */
    uint32_t *d_partials;    /* scan scratch */
    uint32_t *d_staging;     /* num_buckets words of the context (bucket starts) */
    uint64_t scan_words;     /* words the scan stage covers */
} gms_stage_plan;

gms_status gms_multisplit_stage_plan(gms_multisplit_ctx *ctx, const void *d_key_in,
                                     const void *d_value_in, void *d_key_out,
                                     void *d_value_out, uint32_t n, uint32_t key_bytes,
                                     uint32_t value_bytes, void *stream, gms_stage_plan *plan);

/* The scan stage between the caller's prescan and postscan. */
gms_status gms_stage_scan(const gms_stage_plan *plan, void *stream);

/* After the postscan: the num_buckets bucket starts to device memory d_out
 * (asynchronous; plan->d_staging is a valid d_out). */
gms_status gms_stage_offsets(const gms_stage_plan *plan, uint32_t *d_out, void *stream);

/* ------------------------------------------------------------------------ */
/* Multisplit radix sort                                                     */
/* ------------------------------------------------------------------------ */

typedef struct gms_sort_ctx gms_sort_ctx;

/* ms_sort_context (src/api/multisplit_sort.cuh:100-116). */
gms_status gms_sort_create(gms_sort_ctx **out);

/* multisplit_sort_{4,5,6,7,8}bit_allocate[_pairs](n, ctx)
 * (src/api/multisplit_sort.cuh:118-309).  digit_bits in {4,5,6,7,8}. */
gms_status gms_sort_allocate(gms_sort_ctx *ctx, uint32_t n_capacity, uint32_t digit_bits,
                             int key_value);

/* multisplit_sort_release_memory (src/api/multisplit_sort.cuh:312-318). */
gms_status gms_sort_release(gms_sort_ctx *ctx);
gms_status gms_sort_destroy(gms_sort_ctx *ctx);

/* multisplit_sort_{4..8}bit(d_key_in, d_key_out, n, ctx)
 * (src/api/multisplit_sort.cuh:320-559) and
 * multisplit_sort_{4..8}bit_pairs(d_key_in, d_key_out, d_value_in, d_value_out, n, ctx)
 * (src/api/multisplit_sort_pairs.cuh:21-260) -- note the reference's argument
 * order: keys in, keys out, values in, values out.  Values NULL = key-only.
 * LSD radix sort over the reference's digit schedule; d_key_in is not written. */
gms_status gms_sort(gms_sort_ctx *ctx, const uint32_t *d_key_in, uint32_t *d_key_out,
                    const uint32_t *d_value_in, uint32_t *d_value_out, uint32_t n,
                    uint32_t digit_bits, void *stream);

/* Radix sort of 64-bit keys and/or 64-bit values (the reference's sort
 * templates for KeyT = uint64_t / ValueT of 8 bytes, multisplit_sort.cuh:
 * 366-527, multisplit_sort_pairs.cuh:67-228).  gms_sort_allocate_wide sizes
 * the context for key_bytes (4 or 8) and value_bytes (0, 4 or 8) with at least
 * one of 8 bytes; gms_sort_wide then sorts with those widths or narrower.
 * 64-bit keys take ceil(64 / digit_bits) passes of digit_bits (4..8) bits from
 * bit 0 up, 32-bit keys the reference's 32-bit schedule.  Stable; d_key_in is
 * not written; 8-byte arrays must be 8-byte aligned. */
gms_status gms_sort_allocate_wide(gms_sort_ctx *ctx, uint32_t n_capacity, uint32_t digit_bits,
                                  uint32_t key_bytes, uint32_t value_bytes);
gms_status gms_sort_wide(gms_sort_ctx *ctx, const void *d_key_in, void *d_key_out,
                         const void *d_value_in, void *d_value_out, uint32_t n,
                         uint32_t key_bytes, uint32_t value_bytes, void *stream);

/* ------------------------------------------------------------------------ */
/* Compaction (stable 2-bucket multisplit returning the bucket-0 count)      */
/* ------------------------------------------------------------------------ */

typedef struct gms_compaction_ctx gms_compaction_ctx;

/* compaction_context (src/kernels/compaction/multisplit2_compaction.cuh:477-489). */
gms_status gms_compaction_create(gms_compaction_ctx **out);

/* compaction_allocate_key_only / _key_value (multisplit2_compaction.cuh:491-517). */
gms_status gms_compaction_allocate(gms_compaction_ctx *ctx, uint32_t n_capacity, int key_value);

/* compaction_release_memory (multisplit2_compaction.cuh:519-523). */
gms_status gms_compaction_release(gms_compaction_ctx *ctx);
gms_status gms_compaction_destroy(gms_compaction_ctx *ctx);

/* compaction_key_only(kin, kout, n, ctx, f) -> unsigned bucket-0 count and
 * compaction_key_value(kin, vin, kout, vout, n, ctx, f)
 * (multisplit2_compaction.cuh:525-573).  The 2-bucket descriptor must give
 * ids in {0,1} (others clamp to 1).  If h_count != NULL the call blocks and
 * stores the count there (the reference's behaviour); if d_count != NULL the
 * count is written to device memory asynchronously.  Values NULL = key-only. */
gms_status gms_compaction(gms_compaction_ctx *ctx, const uint32_t *d_key_in,
                          const uint32_t *d_value_in, uint32_t *d_key_out,
                          uint32_t *d_value_out, uint32_t n, gms_bucket_desc desc,
                          uint32_t *h_count, uint32_t *d_count, void *stream);

/* The same for 32- or 64-bit keys and 32- or 64-bit values (the reference's
 * compaction_key_*<key_t, value_t> templates for 8-byte types,
 * multisplit2_compaction.cuh:525-573): key_bytes 4 or 8, value_bytes 0, 4 or
 * 8; 8-byte arrays 8-byte aligned; the identifier as a gms_bucket_desc64. */
gms_status gms_compaction_wide(gms_compaction_ctx *ctx, const void *d_key_in,
                               const void *d_value_in, void *d_key_out, void *d_value_out,
                               uint32_t n, uint32_t key_bytes, uint32_t value_bytes,
                               gms_bucket_desc64 desc, uint32_t *h_count, uint32_t *d_count,
                               void *stream);

/* gms_multisplit_stage_plan for a compaction with the caller's own 2-bucket
 * functor (compaction_key_only / _key_value with any bucket_t,
 * multisplit2_compaction.cuh:525-573): after the postscan, the bucket-0
 * count is the start of bucket 1 (gms_stage_offsets). */
gms_status gms_compaction_stage_plan(gms_compaction_ctx *ctx, const void *d_key_in,
                                     const void *d_value_in, void *d_key_out,
                                     void *d_value_out, uint32_t n, uint32_t key_bytes,
                                     uint32_t value_bytes, void *stream, gms_stage_plan *plan);

/* ------------------------------------------------------------------------ */
/* Histogram                                                                 */
/* ------------------------------------------------------------------------ */

/* Adds the counts of the bucket ids of n keys into d_hist[num_bins] (the
 * caller zeroes it, as with the reference's histogram kernels).  Replaces
 * multisplit2_histogram_even_{128,256} and _{64,128,256}bin_128 over integer
 * bins (src/kernels/histogram/multisplit2_histograms.cuh:30-392). */
gms_status gms_histogram_u32(const uint32_t *d_key_in, uint32_t n, uint32_t *d_hist,
                             uint32_t num_bins, gms_bucket_desc desc, void *stream);

/* Even-bin histogram of float samples, bin = (uint32)((x - lower) * inv_delta)
 * in float32 (multisplit2_histogram_even_*(KeyT*, int n, int* d_hist, lower,
 * upper, inv_delta), multisplit2_histograms.cuh:36-392, bin formula :60).
 * Adds into d_hist[num_bins]; samples outside [0, num_bins) are ignored. */
gms_status gms_histogram_even_f32(const float *d_samples, uint32_t n, int32_t *d_hist,
                                  uint32_t num_bins, float lower, float upper,
                                  float inv_delta, void *stream);

/* The same for 32- or 64-bit keys (key_bytes 4 or 8; a 64-bit array 8-byte
 * aligned) with a gms_bucket_desc64 identifier: the reference's histogram
 * kernels templated on KeyT (multisplit2_histograms.cuh:30-392). */
gms_status gms_histogram_wide(const void *d_keys, uint32_t n, uint32_t key_bytes,
                              uint32_t *d_hist, uint32_t num_bins, gms_bucket_desc64 desc,
                              void *stream);

/* Histogram with custom bin edges: levels (device, num_levels ascending
 * values, 2 <= num_levels <= 257) define num_levels - 1 bins [levels[i],
 * levels[i+1]); samples outside [levels[0], levels[num_levels-1]) and NaN are
 * not counted.  Adds into d_hist[num_levels - 1].  Replaces
 * multisplit2_histogram_range_{128,64bin_128,128bin_128,256bin_128}(KeyT* kin,
 * int n, int* d_hist, KeyT* d_levels) (src/kernels/histogram/
 * multisplit2_histograms.cuh:397-699), with the bin semantics of CUB's
 * HistogramRange that the reference's driver compares them with
 * (src/main/main_histogram.cu:224-310). */
gms_status gms_histogram_range_f32(const float *d_samples, uint32_t n, int32_t *d_hist,
                                   uint32_t num_levels, const float *d_levels, void *stream);
gms_status gms_histogram_range_u32(const uint32_t *d_keys, uint32_t n, int32_t *d_hist,
                                   uint32_t num_levels, const uint32_t *d_levels, void *stream);

/* ------------------------------------------------------------------------ */
/* Batched multisplit (no counterpart in the reference)                      */
/* ------------------------------------------------------------------------ */

/* One independent multisplit of a batch.  Same meaning as the arguments of
 * gms_multisplit: values NULL = key-only; d_offsets (device, num_buckets
 * entries) receives the bucket starts or is NULL.  Outputs of different items
 * must not overlap any other item's input or output. */
typedef struct gms_batch_item {
    const uint32_t *d_key_in;
    const uint32_t *d_value_in;
    uint32_t *d_key_out;
    uint32_t *d_value_out;
    uint32_t n;
    uint32_t num_buckets;   /* 1..256 */
    gms_bucket_desc desc;
    uint32_t *d_offsets;
} gms_batch_item;

/* Tile shape and grid of the batched kernels (picked by gms_batch_autotune). */
typedef struct gms_batch_tuning {
    uint32_t items_per_thread;  /* 8, 16, 24 or 32: tile = 256 x items_per_thread keys */
    uint32_t persistent;        /* 0: one workgroup per tile; 1: resident grid
                                   pulling tiles from a device work queue */
} gms_batch_tuning;

typedef struct gms_batch gms_batch;

gms_status gms_batch_create(gms_batch **out);
gms_status gms_batch_destroy(gms_batch *b);

/* Validates and plans `count` items: builds the (instance, tile) work list and
 * the device tables, and sizes the concatenated histogram matrix.  Blocks
 * until the tables are on the device.  The sum of n and the matrix size must
 * stay below 2^32.  Drops any captured graph.  Buffers are not touched. */
gms_status gms_batch_plan(gms_batch *b, const gms_batch_item *items, uint32_t count,
                          void *stream);

/* One independent radix sort of a batch: multisplit_sort_{4..8}bit[_pairs]
 * (src/api/multisplit_sort.cuh:320-559, multisplit_sort_pairs.cuh:21-260)
 * with the same digit schedule, as gms_sort (as gms_sort_wide for 64-bit
 * keys); values NULL = keys only.  The outputs double as a pass buffer and
 * the input is never written. */
typedef struct gms_batch_sort_item {
    const void *d_key_in;
    const void *d_value_in;
    void *d_key_out;
    void *d_value_out;
    uint32_t n;
    uint32_t digit_bits;    /* 4, 5, 6, 7 or 8 */
    uint32_t key_bytes;     /* 4 or 8 (0 = 4) */
    uint32_t value_bytes;   /* 4 or 8 (0 = 4); ignored without values */
} gms_batch_sort_item;

/* One independent multisplit of elements with 32- or 64-bit keys and 32- or
 * 64-bit values, as gms_multisplit_wide (the reference's KeyT/ValueT
 * templates, src/api/multisplit.cuh:81-313): values NULL = key-only;
 * 1..256 buckets; 64-bit arrays 8-byte aligned.  Shapes with 32-bit keys and
 * values run as gms_batch_item (the descriptor's parameter then fits 32
 * bits). */
typedef struct gms_batch_wide_item {
    const void *d_key_in;
    const void *d_value_in;
    void *d_key_out;
    void *d_value_out;
    uint32_t n;
    uint32_t num_buckets;
    uint32_t key_bytes;     /* 4 or 8 */
    uint32_t value_bytes;   /* 4 or 8; ignored without values */
    gms_bucket_desc64 desc;
    uint32_t *d_offsets;
} gms_batch_wide_item;

/* Plans multisplits and radix sorts as one batch: step 0 runs every multisplit
 * and the first digit pass of every sort as one batched multisplit, step p
 * the p-th pass of every sort that has one, in stream order.  The batch owns
 * the sorts' other pass buffer (the size of the sort's keys and values).
 * gms_batch_run, _capture, _launch_graph, _autotune and _stats cover every
 * step.  gms_batch_plan and gms_batch_plan_mixed are this call without the
 * 64-bit items (and sorts). */
gms_status gms_batch_plan_wide(gms_batch *b, const gms_batch_item *items, uint32_t count,
                               const gms_batch_wide_item *wide, uint32_t num_wide,
                               const gms_batch_sort_item *sorts, uint32_t num_sorts,
                               void *stream);
gms_status gms_batch_plan_mixed(gms_batch *b, const gms_batch_item *items, uint32_t count,
                                const gms_batch_sort_item *sorts, uint32_t num_sorts,
                                void *stream);

/* Runs every planned multisplit: one prescan and one postscan launch per
 * (bucket class, key-value) group and one scan, whatever the item count. */
gms_status gms_batch_run(gms_batch *b, void *stream);

/* Captures the launch sequence of the current plan into a hipGraph (kernel
 * nodes only) and instantiates it; gms_batch_launch_graph replays it. */
gms_status gms_batch_capture(gms_batch *b);
gms_status gms_batch_launch_graph(gms_batch *b, void *stream);

/* Times every tuning candidate on the planned batch (reps runs each, outputs
 * are overwritten with the same results), keeps the fastest, re-plans with it
 * and reports it with its time per run in milliseconds. */
gms_status gms_batch_autotune(gms_batch *b, uint32_t reps, void *stream,
                              gms_batch_tuning *chosen, float *ms_per_run);
gms_status gms_batch_set_tuning(gms_batch *b, gms_batch_tuning t);
gms_status gms_batch_get_tuning(const gms_batch *b, gms_batch_tuning *t);

/* Shape of the current plan: keys, work items (tiles), matrix words, kernel
 * launches per run. */
gms_status gms_batch_stats(const gms_batch *b, uint64_t *keys, uint32_t *work_items,
                           uint64_t *matrix_words, uint32_t *launches);

/* ------------------------------------------------------------------------ */
/* Request server with continuous batching (no counterpart in the reference) */
/* ------------------------------------------------------------------------ */

/* A request: one multisplit (as gms_multisplit, or gms_multisplit_wide when
 * a key or value is 64-bit) or one radix sort (as gms_sort, or
 * gms_sort_wide).  Buffers are device memory the caller keeps alive and does
 * not touch until the request has completed; 64-bit arrays are 8-byte
 * aligned. */
typedef enum gms_request_op {
    GMS_OP_MULTISPLIT = 0,
    GMS_OP_SORT = 1
} gms_request_op;

typedef struct gms_request {
    uint32_t op;                  /* gms_request_op */
    const void *d_key_in;
    const void *d_value_in;       /* NULL = key-only */
    void *d_key_out;
    void *d_value_out;
    uint32_t n;
    uint32_t num_buckets;         /* multisplit: 1..256 */
    gms_bucket_desc desc;         /* multisplit of 32-bit keys and values: identifier */
    uint32_t digit_bits;          /* sort: 4..8 */
    uint32_t *d_offsets;          /* multisplit: B bucket starts (device) or NULL */
    uint32_t key_bytes;           /* 4 or 8 (0 = 4) */
    uint32_t value_bytes;         /* 4 or 8 (0 = 4); ignored without values */
    gms_bucket_desc64 desc64;     /* multisplit with a 64-bit key or value: identifier */
} gms_request;

typedef struct gms_server_config {
    uint64_t scratch_budget_bytes;  /* device scratch arena; admission is bounded by it */
    uint32_t max_inflight_steps;    /* steps queued on the GPU at once (1 or 2) */
    uint32_t max_step_keys;         /* keys per step (0 = no limit below 2^31) */
    uint32_t items_per_thread;      /* batched-kernel tile shape: 8, 16, 24 or 32 */
} gms_server_config;

typedef struct gms_server_stats {
    uint64_t submitted, completed, rejected;
    uint64_t steps;                 /* batched launches sequences issued */
    uint64_t step_items;            /* passes summed over steps */
    uint64_t keys_processed;        /* keys summed over all passes */
    uint64_t peak_scratch_bytes;    /* high-water mark of the arena */
    uint32_t active, pending;       /* requests admitted / waiting now */
} gms_server_stats;

typedef struct gms_server gms_server;

/* Allocates the scratch arena on the current device and starts the
 * scheduler thread.  Every step gathers the next pass of every admitted
 * request (a multisplit has one pass, a sort one per digit) into one batched
 * multisplit; requests that arrive while steps run join the next step. */
gms_status gms_server_create(const gms_server_config *cfg, gms_server **out);

/* Stops accepting requests, finishes every submitted one, frees everything. */
gms_status gms_server_destroy(gms_server *s);

/* Queues a request (thread-safe, does not block on the GPU).  Requests wait
 * in arrival order until their scratch fits in the budget; a request that
 * could never fit is rejected with GMS_ERR_CAPACITY. */
gms_status gms_server_submit(gms_server *s, const gms_request *req, uint64_t *ticket);

/* done = 1 once the request's results are in its output buffers; latency_us
 * = submit to completion as seen by the scheduler. */
gms_status gms_server_poll(gms_server *s, uint64_t ticket, int *done, double *latency_us);

/* Blocks until the request completes, then forgets the ticket. */
gms_status gms_server_wait(gms_server *s, uint64_t ticket, double *latency_us);

gms_status gms_server_get_stats(gms_server *s, gms_server_stats *st);

/* ------------------------------------------------------------------------ */
/* Utilities                                                                 */
/* ------------------------------------------------------------------------ */

/* Plain device copy of n u32 with the library's streaming kernel (measures
 * the copy bandwidth that HBM counters are calibrated against). */
gms_status gms_copy_u32(const uint32_t *d_src, uint32_t *d_dst, size_t n, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GPUMULTISPLIT_H_ */
