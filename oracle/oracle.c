/*
 * CPU oracle for the multisplit hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for cpu_baseline, as the timed CPU restatement).  The product path in
 * gpumultisplit_amd/ never links, loads or calls it.
 *
 * This is a clean-room, single-threaded restatement of the reference's CPU
 * algorithm, written from its description (count -> exclusive scan -> stable
 * scatter), not copied.  Every exported function cites the reference lines it
 * follows (paths relative to the reference repository root).
 *
 * How each function is pinned (checks live in tests/test_oracle.py):
 *   ms_oracle_bucket_ids     Python integer arithmetic for every functor kind,
 *                            incl. edge keys 0, 2^32-1, delta boundaries.
 *   ms_oracle_multisplit     numpy stable argsort of the bucket ids (a merge /
 *                            radix sort, independent of counting), the closed
 *                            form pos(i) = sum_{c<b_i} count(c) + #{j<i: b_j=b_i}
 *                            evaluated by brute force on tiny n, special cases
 *                            (single bucket => identity, pre-sorted => identity).
 *   ms_oracle_sort           numpy sort / stable argsort of the keys.
 *   ms_oracle_histogram      numpy bincount.
 *   ms_oracle_histogram_even numpy float32 arithmetic + bincount.
 *   ms_oracle_histogram_range numpy.searchsorted(side="right") - 1 + bincount,
 *                            incl. samples on the edges, outside, and NaN.
 *   ms_oracle_compaction     Python list partition; count = partition point.
 *   ms_oracle_bucket_ids64   Python integers (k*B >> 64, k // d, ...) for every
 *                            kind, incl. keys 0, 2^64-1 and range boundaries.
 *   ms_oracle_multisplit_wide numpy stable argsort of the 64-bit bucket ids.
 *   ms_oracle_sort_wide      numpy stable argsort of the 64-bit keys.
 * The reference itself ships no golden vectors or known-answer tests; the
 * fixtures under tests/golden/ are produced by this oracle and cross-checked by
 * the independent checks above when they are generated and again in tests.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Bucket-identifier kinds.  Same numbering as include/gpumultisplit.h. */
enum {
    MS_OR_RANGE = 0,    /* floor(key * B / 2^32)  (range identifier)            */
    MS_OR_DELTA = 1,    /* key / delta     (delta_func, src/main/main_multisplit.cu:41-47) */
    MS_OR_DIGIT = 2,    /* (key >> shift) & (B-1) (ms_bucket_sort_Nbit,
                           src/api/multisplit_sort.cuh:63-98)                  */
    MS_OR_IDENTITY = 3, /* key             (identity_bucket, main_multisplit.cu:34-39) */
    MS_OR_LOWBITS = 4   /* key & ((1<<flag)-1) (test_bucket<flag>,
                           src/main/main_compaction.cu:27-32)                 */
};

/* One bucket id.  Ids >= B are undefined behaviour in the reference; the
 * library defines them as clamped into bucket B-1, and the oracle states the
 * same rule so that the library's documented behaviour can be checked. */
static inline uint32_t bucket_of(uint32_t kind, uint32_t B, uint32_t param,
                                 uint32_t key)
{
    uint32_t b;
    switch (kind) {
    case MS_OR_RANGE:    b = (uint32_t)(((uint64_t)key * (uint64_t)B) >> 32); break;
    case MS_OR_DELTA:    b = key / param; break;
    case MS_OR_DIGIT:    b = (key >> param) & (B - 1u); break;
    case MS_OR_IDENTITY: b = key; break;
    case MS_OR_LOWBITS:  b = key & ((param >= 32u) ? 0xFFFFFFFFu : ((1u << param) - 1u)); break;
    default:             b = 0; break;
    }
    return b < B ? b : B - 1u;
}

void ms_oracle_bucket_ids(uint32_t kind, uint32_t B, uint32_t param,
                          const uint32_t *keys, uint64_t n, uint32_t *ids)
{
    for (uint64_t i = 0; i < n; ++i)
        ids[i] = bucket_of(kind, B, param, keys[i]);
}

/*
 * Stable multisplit, key-only (vin/vout NULL) or key-value.
 * Follows cpu_multisplit_general (src/cpu_functions.h:155-220) and
 * cpu_multisplit_pairs_general (src/cpu_functions.h:222-288):
 *   1. per-bucket counts           (cpu_functions.h:167-171, :234-238)
 *   2. exclusive scan of counts    (cpu_functions.h:179-181, :246-248)
 *   3. stable scatter in input order with a running cursor per bucket
 *                                  (cpu_functions.h:191-200, :258-267)
 * offsets (optional, B entries) receives the exclusive scan, i.e. the global
 * start of every bucket; this is what multisplit_key_* returns through its
 * optional host array (src/api/multisplit.cuh:144-149, :261-266).
 * Returns 0, or -1 on allocation failure.
 */
int ms_oracle_multisplit(uint32_t kind, uint32_t B, uint32_t param,
                         const uint32_t *kin, const uint32_t *vin,
                         uint32_t *kout, uint32_t *vout, uint64_t n,
                         uint32_t *offsets)
{
    uint64_t *count = (uint64_t *)calloc(B, sizeof(uint64_t));
    uint64_t *cursor = (uint64_t *)calloc(B, sizeof(uint64_t));
    if (!count || !cursor) { free(count); free(cursor); return -1; }

    for (uint64_t i = 0; i < n; ++i)
        count[bucket_of(kind, B, param, kin[i])]++;

    uint64_t run = 0;
    for (uint32_t b = 0; b < B; ++b) {
        cursor[b] = run;
        run += count[b];
    }
    if (offsets)
        for (uint32_t b = 0; b < B; ++b) offsets[b] = (uint32_t)cursor[b];

    for (uint64_t i = 0; i < n; ++i) {
        uint32_t b = bucket_of(kind, B, param, kin[i]);
        uint64_t dst = cursor[b]++;
        kout[dst] = kin[i];
        if (vin && vout) vout[dst] = vin[i];
    }
    free(count);
    free(cursor);
    return 0;
}

/*
 * Digit schedule of the multisplit radix sorts, as (shift, bits) pairs, least
 * significant digit first:
 *   8-bit: 4 x 8 bits at 0, 8, 16, 24            (multisplit_sort.cuh:529-559)
 *   7-bit: 4 bits at 0, then 7 bits at 4, 11, 18, 25 (:320-364)
 *   6-bit: 4 bits at 0, 4, then 6 bits at 8, 14, 20, 26 (:479-527)
 *   5-bit: 5 bits at 0, 5, 10, 15, 20, 25, then 2 bits at 30 (:423-477)
 *   4-bit: 8 x 4 bits at 0, 4, ..., 28           (:366-421)
 * Key-value variants use the same schedules (multisplit_sort_pairs.cuh:21-260).
 * Returns the number of passes, or 0 for an unsupported width.
 */
int ms_oracle_sort_schedule(uint32_t digit_bits, uint32_t *shifts, uint32_t *bits)
{
    int p = 0;
    switch (digit_bits) {
    case 8:
        for (uint32_t s = 0; s < 32; s += 8) { shifts[p] = s; bits[p++] = 8; }
        break;
    case 7:
        shifts[p] = 0; bits[p++] = 4;
        for (uint32_t s = 4; s < 32; s += 7) { shifts[p] = s; bits[p++] = 7; }
        break;
    case 6:
        shifts[p] = 0; bits[p++] = 4;
        shifts[p] = 4; bits[p++] = 4;
        for (uint32_t s = 8; s < 32; s += 6) { shifts[p] = s; bits[p++] = 6; }
        break;
    case 5:
        for (uint32_t s = 0; s < 30; s += 5) { shifts[p] = s; bits[p++] = 5; }
        shifts[p] = 30; bits[p++] = 2;
        break;
    case 4:
        for (uint32_t s = 0; s < 32; s += 4) { shifts[p] = s; bits[p++] = 4; }
        break;
    default:
        return 0;
    }
    return p;
}

/*
 * LSD radix sort built from repeated stable multisplits on digits
 * (multisplit_sort.cuh:320-559, multisplit_sort_pairs.cuh:21-260).  The input
 * is never written; the result lands in kout/vout.  Returns 0 on success.
 */
int ms_oracle_sort(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                   uint32_t *vout, uint64_t n, uint32_t digit_bits)
{
    uint32_t shifts[8], bits[8];
    int passes = ms_oracle_sort_schedule(digit_bits, shifts, bits);
    if (passes == 0) return -2;
    int kv = (vin != NULL && vout != NULL);
    uint32_t *ka = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *kb = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *va = kv ? (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t)) : NULL;
    uint32_t *vb = kv ? (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t)) : NULL;
    if (!ka || !kb || (kv && (!va || !vb))) {
        free(ka); free(kb); free(va); free(vb);
        return -1;
    }
    memcpy(ka, kin, n * sizeof(uint32_t));
    if (kv) memcpy(va, vin, n * sizeof(uint32_t));
    for (int p = 0; p < passes; ++p) {
        ms_oracle_multisplit(MS_OR_DIGIT, 1u << bits[p], shifts[p], ka,
                             kv ? va : NULL, kb, kv ? vb : NULL, n, NULL);
        uint32_t *t = ka; ka = kb; kb = t;
        if (kv) { t = va; va = vb; vb = t; }
    }
    memcpy(kout, ka, n * sizeof(uint32_t));
    if (kv) memcpy(vout, va, n * sizeof(uint32_t));
    free(ka); free(kb); free(va); free(vb);
    return 0;
}

/*
 * Histogram of bucket ids: hist[b] += #{i : f(key_i) = b}.  Accumulates, as
 * the reference's histogram kernels atomically add into a caller-zeroed array
 * (src/kernels/histogram/multisplit2_histograms.cuh:36-392).
 */
void ms_oracle_histogram(uint32_t kind, uint32_t B, uint32_t param,
                         const uint32_t *keys, uint64_t n, uint32_t *hist)
{
    for (uint64_t i = 0; i < n; ++i)
        hist[bucket_of(kind, B, param, keys[i])]++;
}

/*
 * Even-bin histogram of float samples: bin = (uint32)((x - lower) * inv_delta)
 * evaluated in float32, as multisplit2_histogram_even_* computes it
 * (multisplit2_histograms.cuh:60, :330).  Samples whose bin falls outside
 * [0, nbins) are ignored (the reference does not guard; CUB's HistogramEven,
 * which its driver compares against, ignores them).  Accumulates into hist.
 */
void ms_oracle_histogram_even_f32(const float *x, uint64_t n, int32_t *hist,
                                  uint32_t nbins, float lower, float inv_delta)
{
    for (uint64_t i = 0; i < n; ++i) {
        float d = x[i] - lower;
        if (!(d >= 0.0f)) continue;           /* negative or NaN */
        float s = d * inv_delta;
        if (!(s < (float)nbins)) continue;
        uint32_t b = (uint32_t)s;
        if (b < nbins) hist[b]++;
    }
}

/*
 * Histogram with custom bin edges: levels[0..nl) ascending, bin i counts the
 * samples with levels[i] <= x < levels[i+1]; samples below levels[0], at or
 * above levels[nl-1], and NaN are ignored -- the semantics of CUB's
 * HistogramRange, which the reference's range histograms are compared with
 * (src/main/main_histogram.cu:224-310; kernels multisplit2_histograms.cuh:397-699,
 * bin = UpperBound(levels, x) - 1).  Linear scan of the levels here; the tests
 * pin it against numpy.searchsorted.  Accumulates into hist (nl - 1 bins).
 */
void ms_oracle_histogram_range_f32(const float *x, uint64_t n, int32_t *hist,
                                   const float *levels, uint32_t nl)
{
    if (nl < 2) return;
    for (uint64_t i = 0; i < n; ++i) {
        const float v = x[i];
        if (!(v >= levels[0] && v < levels[nl - 1])) continue;   /* out of range or NaN */
        uint32_t b = 0;
        while (b + 2 < nl && !(v < levels[b + 1])) ++b;
        hist[b]++;
    }
}

void ms_oracle_histogram_range_u32(const uint32_t *x, uint64_t n, int32_t *hist,
                                   const uint32_t *levels, uint32_t nl)
{
    if (nl < 2) return;
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t v = x[i];
        if (!(v >= levels[0] && v < levels[nl - 1])) continue;
        uint32_t b = 0;
        while (b + 2 < nl && !(v < levels[b + 1])) ++b;
        hist[b]++;
    }
}

/*
 * Stable 2-bucket multisplit (bucket 0 first) that returns the bucket-0 count,
 * as compaction_key_only / compaction_key_value do
 * (src/kernels/compaction/multisplit2_compaction.cuh:525-573, return value
 * read at :541-542).  Returns the count, or UINT64_MAX on error.
 */
uint64_t ms_oracle_compaction(uint32_t kind, uint32_t param, const uint32_t *kin,
                              const uint32_t *vin, uint32_t *kout,
                              uint32_t *vout, uint64_t n)
{
    uint32_t offsets[2];
    if (ms_oracle_multisplit(kind, 2u, param, kin, vin, kout, vout, n, offsets) != 0)
        return UINT64_MAX;
    uint64_t zeros = 0;
    for (uint64_t i = 0; i < n; ++i)
        zeros += (bucket_of(kind, 2u, param, kin[i]) == 0u);
    return zeros;
}

/* Per-tile bucket counts in the library's bucket-major layout
 * H[b * num_tiles + t] = #{ keys of tile t in bucket b }, tiles of `tile`
 * consecutive keys.  Same quantity the reference's unused checker
 * cpu_prescan_test computes (src/cpu_functions.h:55-103). */
void ms_oracle_tile_histograms(uint32_t kind, uint32_t B, uint32_t param,
                               const uint32_t *keys, uint64_t n, uint32_t tile,
                               uint32_t *H)
{
    uint64_t nt = (n + tile - 1) / tile;
    memset(H, 0, (size_t)(B * nt) * sizeof(uint32_t));
    for (uint64_t i = 0; i < n; ++i)
        H[(uint64_t)bucket_of(kind, B, param, keys[i]) * nt + i / tile]++;
}

/*
 * 64-bit keys and 64-bit values.  The reference's API is templated on KeyT and
 * ValueT (src/api/multisplit.cuh:81-82, :199-200; multisplit_sort.cuh:366-527;
 * multisplit_sort_pairs.cuh:67-228) and its CPU multisplit is generic over the
 * key type (src/cpu_functions.h:155-288); its bucket identifiers extend to
 * 64-bit keys as written there: key / delta, (key >> shift) & mask, key.  The
 * range identifier is floor(key * B / 2^64), the 64-bit form of the 32-bit one.
 * Elements are key_words (1 or 2) and value_words (0, 1 or 2) little-endian
 * 32-bit words: a 64-bit key is lo | hi << 32.
 */
static inline uint32_t bucket_of64(uint32_t kind, uint32_t B, uint64_t param, uint32_t width,
                                   uint64_t key)
{
    uint64_t b;
    switch (kind) {
    case MS_OR_RANGE:    b = (uint64_t)(((unsigned __int128)key * B) >> 64); break;
    case MS_OR_DELTA:    b = key / param; break;
    case MS_OR_DIGIT:    b = (key >> param) & (width ? ((1ull << width) - 1ull) : (uint64_t)(B - 1u)); break;
    case MS_OR_IDENTITY: b = key; break;
    case MS_OR_LOWBITS:  b = key & (param >= 64u ? ~0ull : ((1ull << param) - 1ull)); break;
    default:             b = 0; break;
    }
    return b < B ? (uint32_t)b : B - 1u;
}

static inline uint64_t key_at(const uint32_t *kin, uint32_t kw, uint64_t i)
{
    return kw == 2 ? ((uint64_t)kin[2 * i] | ((uint64_t)kin[2 * i + 1] << 32)) : kin[i];
}

/* Bucket of element i: the 64-bit identifiers for 64-bit keys, the 32-bit
 * ones (bucket_of, with a digit width) for 32-bit keys. */
static inline uint32_t bucket_at(uint32_t kind, uint32_t B, uint64_t param, uint32_t width,
                                 const uint32_t *kin, uint32_t kw, uint64_t i)
{
    if (kw == 2) return bucket_of64(kind, B, param, width, key_at(kin, 2, i));
    if (kind == MS_OR_DIGIT && width) {
        const uint32_t b = (kin[i] >> param) & (width >= 32 ? ~0u : ((1u << width) - 1u));
        return b < B ? b : B - 1u;
    }
    return bucket_of(kind, B, (uint32_t)param, kin[i]);
}

void ms_oracle_bucket_ids64(uint32_t kind, uint32_t B, uint64_t param, uint32_t width,
                            const uint64_t *keys, uint64_t n, uint32_t *ids)
{
    for (uint64_t i = 0; i < n; ++i)
        ids[i] = bucket_of64(kind, B, param, width, keys[i]);
}

/* Stable multisplit of elements of key_words + value_words words; the same
 * count -> exclusive scan -> stable scatter as ms_oracle_multisplit
 * (cpu_functions.h:155-288).  vin/vout NULL (or value_words 0) = key-only. */
int ms_oracle_multisplit_wide(uint32_t kind, uint32_t B, uint64_t param, uint32_t width,
                              uint32_t key_words, uint32_t value_words,
                              const uint32_t *kin, const uint32_t *vin,
                              uint32_t *kout, uint32_t *vout, uint64_t n, uint32_t *offsets)
{
    if (key_words < 1 || key_words > 2 || value_words > 2) return -2;
    uint64_t *count = (uint64_t *)calloc(B, sizeof(uint64_t));
    uint64_t *cursor = (uint64_t *)calloc(B, sizeof(uint64_t));
    if (!count || !cursor) { free(count); free(cursor); return -1; }
    for (uint64_t i = 0; i < n; ++i)
        count[bucket_at(kind, B, param, width, kin, key_words, i)]++;
    uint64_t run = 0;
    for (uint32_t b = 0; b < B; ++b) {
        cursor[b] = run;
        run += count[b];
    }
    if (offsets)
        for (uint32_t b = 0; b < B; ++b) offsets[b] = (uint32_t)cursor[b];
    const int kv = vin && vout && value_words;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t dst = cursor[bucket_at(kind, B, param, width, kin, key_words, i)]++;
        for (uint32_t w = 0; w < key_words; ++w) kout[dst * key_words + w] = kin[i * key_words + w];
        if (kv)
            for (uint32_t w = 0; w < value_words; ++w)
                vout[dst * value_words + w] = vin[i * value_words + w];
    }
    free(count);
    free(cursor);
    return 0;
}

/* Stable LSD radix sort of 64-bit keys (optionally carrying values of
 * value_words words): passes of digit_bits bits from bit 0 up, the last one
 * narrower, each a stable multisplit by that digit (the schedule of
 * multisplit_sort.cuh:320-559 continued over 64 bits). */
int ms_oracle_sort_wide(const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout,
                        uint64_t n, uint32_t value_words, uint32_t digit_bits)
{
    if (digit_bits < 1 || digit_bits > 8 || value_words > 2) return -2;
    const int kv = vin != NULL && vout != NULL && value_words;
    const size_t kb = (n ? n : 1) * 2 * sizeof(uint32_t), vb = (n ? n : 1) * (value_words ? value_words : 1) * sizeof(uint32_t);
    uint32_t *ka = (uint32_t *)malloc(kb), *kt = (uint32_t *)malloc(kb);
    uint32_t *va = kv ? (uint32_t *)malloc(vb) : NULL, *vt = kv ? (uint32_t *)malloc(vb) : NULL;
    if (!ka || !kt || (kv && (!va || !vt))) {
        free(ka); free(kt); free(va); free(vt);
        return -1;
    }
    memcpy(ka, kin, n * 2 * sizeof(uint32_t));
    if (kv) memcpy(va, vin, n * value_words * sizeof(uint32_t));
    for (uint32_t shift = 0; shift < 64; shift += digit_bits) {
        const uint32_t w = 64 - shift < digit_bits ? 64 - shift : digit_bits;
        ms_oracle_multisplit_wide(MS_OR_DIGIT, 1u << w, shift, w, 2, kv ? value_words : 0, ka,
                                  kv ? va : NULL, kt, kv ? vt : NULL, n, NULL);
        uint32_t *t = ka; ka = kt; kt = t;
        if (kv) { t = va; va = vt; vt = t; }
    }
    memcpy(kout, ka, n * 2 * sizeof(uint32_t));
    if (kv) memcpy(vout, va, n * value_words * sizeof(uint32_t));
    free(ka); free(kt); free(va); free(vt);
    return 0;
}
