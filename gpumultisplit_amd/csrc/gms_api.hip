// C-ABI entry points (include/gpumultisplit.h) and the host-side engine.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>

#include "gms_internal.hpp"
#include "gms_kernels.hpp"

using namespace gms;

struct gms_multisplit_ctx {
    uint32_t num_buckets = 0;
    uint32_t capacity = 0;
    int allocated = 0;
    int variant = GMS_VARIANT_AUTO;  // block-wide or wave-level kernels
    // 64-bit elements into more than 256 buckets: pass temporaries of
    // capacity x 2 words each, allocated by the first such call
    uint32_t* wide_k = nullptr;
    uint32_t* wide_v = nullptr;
    Scratch s;
    LargeScratch large;  // more than 256 buckets only
    Profiler prof;
};

struct gms_sort_ctx {
    uint32_t capacity = 0;
    uint32_t digit_bits = 0;
    int kv = 0;
    uint32_t key_words = 1;    // widths the temporaries hold (gms_sort_allocate_wide)
    uint32_t value_words = 1;
    int allocated = 0;
    Scratch s;
    uint32_t* ktmp = nullptr;
    uint32_t* vtmp = nullptr;
};

struct gms_compaction_ctx {
    uint32_t capacity = 0;
    int allocated = 0;
    Scratch s;
};

namespace gms {

gms_status hip_status(hipError_t e) { return e == hipSuccess ? GMS_SUCCESS : GMS_ERR_HIP; }

bool ranges_overlap(const void* a, size_t an, const void* b, size_t bn) {
    if (!a || !b || an == 0 || bn == 0) return false;
    const uintptr_t a0 = (uintptr_t)a, a1 = a0 + an, b0 = (uintptr_t)b, b1 = b0 + bn;
    return a0 < b1 && b0 < a1;
}

size_t matrix_elems(uint32_t n, uint32_t num_buckets) {
    if (num_buckets <= 1 || n == 0) return 0;
    if (num_buckets > 256) {
        // 64-bit elements, and 32-bit ones above kSinglePassBuckets, take
        // passes of at most 256 buckets; 32-bit ones up to it take one pass.
        size_t m = matrix_elems(n, 256);
        if (num_buckets <= kSinglePassBuckets) {
            const int lb = ceil_log2_u32(num_buckets);
            const uint32_t tko = std::min(tile_keys(lb, false, 0), tile_keys(lb, false, kLongTileMin));
            const uint32_t tkv = std::min(tile_keys(lb, true, 0), tile_keys(lb, true, kLongTileMin));
            const size_t t = std::max(((size_t)n + tko - 1) / tko, ((size_t)n + tkv - 1) / tkv);
            m = std::max(m, ((size_t)1 << lb) * t);
        }
        return m;
    }
    const int lb = ceil_log2_u32(num_buckets);
    // the shortest of the short- and long-input tiles of each class
    const uint32_t tko = std::min(tile_keys(lb, false, 0), tile_keys(lb, false, kLongTileMin));
    const uint32_t tkv = std::min(tile_keys(lb, true, 0), tile_keys(lb, true, kLongTileMin));
    const size_t t_ko = ((size_t)n + tko - 1) / tko;
    const size_t t_kv = ((size_t)n + tkv - 1) / tkv;
    // the shortest tile of the 64-bit kernels (gms_multisplit_wide, gms_sort_wide)
    const size_t t_w = ((size_t)n + wide_min_tile_elems() - 1) / wide_min_tile_elems();
    // one column per wave subproblem of the wave-level variant (up to 64 buckets)
    const size_t t_wms = num_buckets <= kWmsMaxBuckets
                             ? ((size_t)n + wms_subproblem_keys() - 1) / wms_subproblem_keys()
                             : 0;
    return ((size_t)1 << lb) * std::max(std::max(std::max(t_ko, t_kv), t_w), t_wms);
}

gms_status scratch_alloc(Scratch& s, size_t hist_elems) {
    scratch_free(s);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return GMS_ERR_NO_DEVICE;
    s.device = dev;
    const size_t he = std::max<size_t>(hist_elems, 4);
    const size_t pe = scan_partials_needed(he);
    if (hipMalloc((void**)&s.hist, he * sizeof(uint32_t)) != hipSuccess) {
        s.hist = nullptr;
        return GMS_ERR_ALLOC;
    }
    if (hipMalloc((void**)&s.part, pe * sizeof(uint32_t)) != hipSuccess) {
        s.part = nullptr;
        scratch_free(s);
        return GMS_ERR_ALLOC;
    }
    if (hipMalloc((void**)&s.offs, 256 * sizeof(uint32_t)) != hipSuccess) {
        s.offs = nullptr;
        scratch_free(s);
        return GMS_ERR_ALLOC;
    }
    const size_t sw = scan_chunks(he);
    if (hipMalloc((void**)&s.status, sw * sizeof(uint64_t)) != hipSuccess) {
        s.status = nullptr;
        scratch_free(s);
        return GMS_ERR_ALLOC;
    }
    s.hist_cap = he;
    s.part_cap = pe;
    s.status_cap = sw;
    return GMS_SUCCESS;
}

void scratch_free(Scratch& s) {
    if (s.hist) (void)hipFree(s.hist);
    if (s.part) (void)hipFree(s.part);
    if (s.offs) (void)hipFree(s.offs);
    if (s.status) (void)hipFree(s.status);
    s.hist = s.part = s.offs = nullptr;
    s.status = nullptr;
    s.hist_cap = s.part_cap = s.status_cap = 0;
}

hipEvent_t* Profiler::next(hipStream_t) {
    if (used + 4 > ev.size()) {
        if (ev.size() < 4 * 1024) {
            for (int i = 0; i < 4; ++i) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return nullptr;
                ev.push_back(e);
            }
        } else if (flush() != GMS_SUCCESS) {
            return nullptr;
        }
    }
    hipEvent_t* r = &ev[used];
    used += 4;
    return r;
}

gms_status Profiler::flush() {
    for (size_t i = 0; i + 3 < used; i += 4) {
        if (hipEventSynchronize(ev[i + 3]) != hipSuccess) return GMS_ERR_HIP;
        for (int st = 0; st < 3; ++st) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[i + st], ev[i + st + 1]) != hipSuccess) return GMS_ERR_HIP;
            ms[st] += t;
        }
        ++calls;
    }
    used = 0;
    return GMS_SUCCESS;
}

void Profiler::destroy() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
    used = 0;
}

static bool is_pow2(uint32_t x) { return x && !(x & (x - 1)); }

gms_status check_desc(uint32_t num_buckets, gms_bucket_desc d) {
    switch (d.kind) {
        case GMS_BUCKET_RANGE: return GMS_SUCCESS;
        case GMS_BUCKET_DELTA: return d.param >= 1 ? GMS_SUCCESS : GMS_ERR_INVALID_ARG;
        case GMS_BUCKET_DIGIT:
            if (d.param >= 32 || d.width > 20) return GMS_ERR_INVALID_ARG;
            if (d.width == 0) return is_pow2(num_buckets) ? GMS_SUCCESS : GMS_ERR_INVALID_ARG;
            return (1u << d.width) <= num_buckets ? GMS_SUCCESS : GMS_ERR_INVALID_ARG;
        case GMS_BUCKET_IDENTITY: return GMS_SUCCESS;
        case GMS_BUCKET_LOWBITS:
            return (d.param < 32 && (1u << d.param) <= num_buckets) ? GMS_SUCCESS
                                                                     : GMS_ERR_INVALID_ARG;
        default: return GMS_ERR_INVALID_ARG;
    }
}

gms_status check_desc64(uint32_t B, const gms_bucket_desc64& d, uint32_t key_bits) {
    switch (d.kind) {
        case GMS_BUCKET_RANGE:
        case GMS_BUCKET_IDENTITY: return GMS_SUCCESS;
        case GMS_BUCKET_DELTA:
            return (d.param >= 1 && (key_bits == 64 || d.param <= 0xFFFFFFFFull)) ? GMS_SUCCESS
                                                                                  : GMS_ERR_INVALID_ARG;
        case GMS_BUCKET_DIGIT:
            if (d.param >= key_bits || d.width > 20) return GMS_ERR_INVALID_ARG;
            if (d.width == 0) return is_pow2(B) ? GMS_SUCCESS : GMS_ERR_INVALID_ARG;
            return (1u << d.width) <= B ? GMS_SUCCESS : GMS_ERR_INVALID_ARG;
        case GMS_BUCKET_LOWBITS:
            return (d.param < key_bits && d.param < 32 && (1u << d.param) <= B) ? GMS_SUCCESS
                                                                                : GMS_ERR_INVALID_ARG;
        default: return GMS_ERR_INVALID_ARG;
    }
}

static bool aligned8(const void* p) { return (((uintptr_t)p) & 7u) == 0; }

static uint32_t digit_mask(uint32_t B, gms_bucket_desc d) {
    return d.width ? (1u << d.width) - 1u : B - 1u;
}

template <class Fn>
static hipError_t with_functor(uint32_t B, gms_bucket_desc d, Fn&& fn) {
    switch (d.kind) {
        case GMS_BUCKET_RANGE: return fn(RangeF{B});
        case GMS_BUCKET_DELTA: return fn(DeltaF::make(d.param, B));
        case GMS_BUCKET_DIGIT: return fn(DigitF{d.param, digit_mask(B, d)});
        case GMS_BUCKET_LOWBITS: return fn(DigitF{0u, (1u << d.param) - 1u});
        case GMS_BUCKET_IDENTITY: return fn(IdentF{B - 1u});
        default: return hipErrorInvalidValue;
    }
}

static bool vec_aligned(const void* p) { return p == nullptr || (((uintptr_t)p) & 15u) == 0; }

// The plain path scans the matrix in one launch when the scratch has status
// words for it (cleared by the prescan), else with the two-launch scan.
static bool single_scan(const Scratch& s, size_t m) {
    return s.status && scan_chunks(m) <= s.status_cap;
}
static bool scan_fits(const Scratch& s, size_t m) {
    return single_scan(s, m) || scan_partials_needed(m) <= s.part_cap;
}
static hipError_t scan_matrix(const Scratch& s, uint32_t m, hipStream_t stream) {
    if (single_scan(s, m)) return launch_scan_single(s.hist, s.hist, m, s.status, stream);
    return launch_exclusive_scan(s.hist, s.hist, m, s.part, stream);
}

// The launch description of the plain stages for n keys in B buckets, or
// GMS_ERR_CAPACITY when the scratch is too small for it.
static gms_status plan_call(Scratch& s, const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                            uint32_t* vout, uint32_t n, uint32_t B, bool kv, hipStream_t stream,
                            MsCall& c) {
    const int lb = ceil_log2_u32(B);
    const uint32_t tile = tile_keys(lb, kv, n);
    const uint32_t nt = (uint32_t)(((size_t)n + tile - 1) / tile);
    const size_t m = ((size_t)1 << lb) * nt;
    if (m > s.hist_cap || !scan_fits(s, m)) return GMS_ERR_CAPACITY;
    c = MsCall{kin, vin, kout, vout, n, lb, kv, n >= long_tile_min(lb, kv), nt, s.hist,
               vec_aligned(kin) && vec_aligned(vin), stream};
    if (single_scan(s, m)) {
        c.scan_status = s.status;
        c.scan_chunks = scan_chunks(m);
    }
    return GMS_SUCCESS;
}

gms_status run_prescan_only(Scratch& s, const uint32_t* kin, uint32_t n, uint32_t B,
                            gms_bucket_desc d, bool kv, hipStream_t stream, uint32_t* nt_out) {
    const int lb = ceil_log2_u32(B);
    const uint32_t tile = tile_keys(lb, kv, n);
    if (nt_out) *nt_out = (uint32_t)(((size_t)n + tile - 1) / tile);
    if (n == 0 || B <= 1) return GMS_SUCCESS;
    MsCall c;
    gms_status st = plan_call(s, kin, nullptr, nullptr, nullptr, n, B, kv, stream, c);
    if (st != GMS_SUCCESS) return st;
    return hip_status(with_functor(B, d, [&](auto f) { return launch_prescan(c, f); }));
}

// prescan -> scan -> postscan with the identifier that `with_f` hands to its
// callback (the built-in kinds by descriptor, or a given functor).
template <class WithF>
static gms_status run_stages(Scratch& s, const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                             uint32_t* vout, uint32_t n, uint32_t B, uint32_t* d_offsets,
                             hipStream_t stream, Profiler* prof, WithF&& with_f) {
    const bool kv = vin != nullptr;
    if (n == 0) return d_offsets ? hip_status(launch_fill(d_offsets, 0u, B, stream)) : GMS_SUCCESS;
    if (B == 1) {
        hipError_t e = launch_copy(kin, kout, n, stream);
        if (e == hipSuccess && kv) e = launch_copy(vin, vout, n, stream);
        if (e == hipSuccess && d_offsets) e = launch_fill(d_offsets, 0u, 1u, stream);
        return hip_status(e);
    }
    MsCall c;
    gms_status st = plan_call(s, kin, vin, kout, vout, n, B, kv, stream, c);
    if (st != GMS_SUCCESS) return st;
    const uint32_t m = (uint32_t)(((size_t)1 << c.log_b) * c.num_tiles);
    hipEvent_t* ev = (prof && prof->on) ? prof->next(stream) : nullptr;
    hipError_t e = hipSuccess;
    if (ev) e = hipEventRecord(ev[0], stream);
    if (e == hipSuccess) e = with_f([&](auto f) { return launch_prescan(c, f); });
    if (e == hipSuccess && ev) e = hipEventRecord(ev[1], stream);
    if (e == hipSuccess) e = scan_matrix(s, m, stream);
    if (e == hipSuccess && ev) e = hipEventRecord(ev[2], stream);
    if (e == hipSuccess) e = with_f([&](auto f) { return launch_postscan(c, f); });
    if (e == hipSuccess && ev) e = hipEventRecord(ev[3], stream);
    if (e == hipSuccess && d_offsets) e = launch_gather_offsets(s.hist, c.num_tiles, B, d_offsets, stream);
    return hip_status(e);
}

gms_status run_multisplit(Scratch& s, const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                          uint32_t* vout, uint32_t n, uint32_t B, gms_bucket_desc d,
                          uint32_t* d_offsets, hipStream_t stream, Profiler* prof) {
    return run_stages(s, kin, vin, kout, vout, n, B, d_offsets, stream, prof,
                      [&](auto&& launch) { return with_functor(B, d, launch); });
}

gms_status run_multisplit_digit(Scratch& s, const uint32_t* kin, const uint32_t* vin,
                                uint32_t* kout, uint32_t* vout, uint32_t n, uint32_t B,
                                const DigitOfF& f, hipStream_t stream) {
    return run_stages(s, kin, vin, kout, vout, n, B, nullptr, stream, nullptr,
                      [&](auto&& launch) { return launch(f); });
}

int sort_schedule(uint32_t digit_bits, SortPass* p) {
    int k = 0;
    switch (digit_bits) {
        case 8:
            for (uint32_t s = 0; s < 32; s += 8) p[k++] = {s, 8};
            break;
        case 7:
            p[k++] = {0, 4};
            for (uint32_t s = 4; s < 32; s += 7) p[k++] = {s, 7};
            break;
        case 6:
            p[k++] = {0, 4};
            p[k++] = {4, 4};
            for (uint32_t s = 8; s < 32; s += 6) p[k++] = {s, 6};
            break;
        case 5:
            for (uint32_t s = 0; s < 30; s += 5) p[k++] = {s, 5};
            p[k++] = {30, 2};
            break;
        case 4:
            for (uint32_t s = 0; s < 32; s += 4) p[k++] = {s, 4};
            break;
        default:
            return 0;
    }
    return k;
}

// ------------------------------------------------- stages of a caller's functor
// The scratch layout of a multisplit whose identifier is compiled into the
// caller's kernels (gpumultisplit_device.hpp); validation as gms_multisplit_wide.
static gms_status stage_plan(Scratch& s, uint32_t capacity, int allocated, uint32_t B,
                             uint32_t* staging, const void* kin, const void* vin, void* kout,
                             void* vout, uint32_t n, uint32_t key_bytes, uint32_t value_bytes,
                             hipStream_t stream, gms_stage_plan* plan) {
    if (!plan) return GMS_ERR_INVALID_ARG;
    *plan = gms_stage_plan{};
    if ((key_bytes != 4 && key_bytes != 8) || (value_bytes != 0 && value_bytes != 4 && value_bytes != 8))
        return GMS_ERR_INVALID_ARG;
    if (!allocated) return GMS_ERR_NOT_ALLOCATED;
    if (n > capacity) return GMS_ERR_CAPACITY;
    const uint32_t kw = key_bytes / 4, vw = value_bytes / 4;
    const bool wide = kw == 2 || vw == 2;
    if (B > (wide ? 256u : kSinglePassBuckets)) return GMS_ERR_UNSUPPORTED_BUCKETS;
    if (value_bytes && n && (!vin || !vout)) return GMS_ERR_INVALID_ARG;
    if (!value_bytes && (vin || vout)) return GMS_ERR_INVALID_ARG;
    if (n) {
        if (!kin || !kout) return GMS_ERR_INVALID_ARG;
        if ((kw == 2 && (!aligned8(kin) || !aligned8(kout))) ||
            (vw == 2 && (!aligned8(vin) || !aligned8(vout))))
            return GMS_ERR_INVALID_ARG;
        const size_t kb = (size_t)n * key_bytes, vb = (size_t)n * value_bytes;
        if (ranges_overlap(kin, kb, kout, kb)) return GMS_ERR_ALIAS;
        if (vw && (ranges_overlap(vin, vb, vout, vb) || ranges_overlap(kin, kb, vout, vb) ||
                   ranges_overlap(vin, vb, kout, kb) || ranges_overlap(kout, kb, vout, vb)))
            return GMS_ERR_ALIAS;
    }
    plan->num_buckets = B;
    plan->key_words = kw;
    plan->value_words = vw;
    plan->n = n;
    plan->d_partials = s.part;
    plan->d_staging = staging;
    if (n == 0 || B == 1) {  // nothing for the caller's kernels: one bucket is a copy
        plan->done = 1;
        hipError_t e = hipSuccess;
        if (n) e = launch_copy(static_cast<const uint32_t*>(kin), static_cast<uint32_t*>(kout),
                               (size_t)n * kw, stream);
        if (e == hipSuccess && n && vw)
            e = launch_copy(static_cast<const uint32_t*>(vin), static_cast<uint32_t*>(vout),
                            (size_t)n * vw, stream);
        return hip_status(e);
    }
    const int lb = ceil_log2_u32(B);
    const size_t rows = (size_t)1 << lb;
    plan->log_b = (uint32_t)lb;
    plan->vec = vec_aligned(kin) && vec_aligned(vin);  // only the loads are vectors
    plan->d_matrix = s.hist;
    if (wide) {
        const uint32_t tile = wide_tile_elems(kw, vw);
        const uint32_t nt = (uint32_t)(((size_t)n + tile - 1) / tile);
        if (rows * nt > s.hist_cap || scan_partials_needed(rows * nt) > s.part_cap)
            return GMS_ERR_CAPACITY;
        plan->num_tiles = nt;
        plan->scan_words = rows * nt;
        return GMS_SUCCESS;
    }
    MsCall c;
    gms_status st = plan_call(s, static_cast<const uint32_t*>(kin), static_cast<const uint32_t*>(vin),
                              static_cast<uint32_t*>(kout), static_cast<uint32_t*>(vout), n, B,
                              vw != 0, stream, c);
    if (st != GMS_SUCCESS) return st;
    plan->num_tiles = c.num_tiles;
    plan->large = c.large ? 1u : 0u;
    plan->scan_words = rows * c.num_tiles;
    return GMS_SUCCESS;
}

}  // namespace gms

// ---------------------------------------------------------------- no-alloc form
// Scratch of gms_multisplit_noalloc: the single-pass scan's status words
// (cleared by each call's prescan), then the per-tile count matrix.  One copy
// per GPU, held until the process ends.  In host code `scratch` names a
// host-side stand-in, never the device memory: host code uses only the
// address hipGetSymbolAddress returns on each device.
constexpr size_t kNoAllocWords = GMS_NOALLOC_SCRATCH_BYTES / sizeof(uint32_t);
constexpr size_t kNoAllocStatus = (kNoAllocWords + kScanChunk - 1) / kScanChunk;  // 64-bit words
constexpr size_t kNoAllocHead = (2 * kNoAllocStatus + 63) / 64 * 64;             // 32-bit words
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
uint32_t* g_noalloc_ptr[kNoAllocDevices] = {};
}  // namespace

// The current device's scratch address, looked up on the first call there.
static gms_status noalloc_scratch(uint32_t** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices)
        return GMS_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        uint32_t* ptr = nullptr;
        if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
            return GMS_ERR_HIP;
        g_noalloc_ptr[dev] = ptr;
    }
    *out = g_noalloc_ptr[dev];
    return GMS_SUCCESS;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* gms_status_string(gms_status s) {
    switch (s) {
        case GMS_SUCCESS: return "success";
        case GMS_ERR_INVALID_ARG: return "invalid argument";
        case GMS_ERR_UNSUPPORTED_BUCKETS: return "unsupported number of buckets";
        case GMS_ERR_CAPACITY: return "n exceeds the allocated capacity";
        case GMS_ERR_NOT_ALLOCATED: return "context not allocated";
        case GMS_ERR_ALLOC: return "device allocation failed";
        case GMS_ERR_HIP: return "HIP runtime error";
        case GMS_ERR_ALIAS: return "input and output ranges overlap";
        case GMS_ERR_NO_DEVICE: return "no HIP device";
        default: return "unknown status";
    }
}

uint32_t gms_abi_version(void) { return GMS_ABI_VERSION; }

gms_status gms_bucket_ids_host(uint32_t num_buckets, gms_bucket_desc desc, const uint32_t* keys,
                               uint32_t* ids, size_t n) {
    if (num_buckets < 1 || num_buckets > kMaxBuckets) return GMS_ERR_UNSUPPORTED_BUCKETS;
    gms_status st = check_desc(num_buckets, desc);
    if (st != GMS_SUCCESS) return st;
    if (n && (!keys || !ids)) return GMS_ERR_INVALID_ARG;
    (void)with_functor(num_buckets, desc, [&](auto f) {
        for (size_t i = 0; i < n; ++i) ids[i] = num_buckets == 1 ? 0u : f(keys[i]);
        return hipSuccess;
    });
    return GMS_SUCCESS;
}

gms_status gms_bucket_ids64_host(uint32_t num_buckets, gms_bucket_desc64 desc,
                                 const uint64_t* keys, uint32_t* ids, size_t n) {
    if (num_buckets < 1 || num_buckets > 256) return GMS_ERR_UNSUPPORTED_BUCKETS;
    gms_status st = check_desc64(num_buckets, desc, 64);
    if (st != GMS_SUCCESS) return st;
    if (n && (!keys || !ids)) return GMS_ERR_INVALID_ARG;
    const AnyF64 f = AnyF64::make(desc.kind, desc.param, desc.width, num_buckets);
    for (size_t i = 0; i < n; ++i) ids[i] = num_buckets == 1 ? 0u : f(keys[i]);
    return GMS_SUCCESS;
}

uint32_t gms_subproblem_size(uint32_t num_buckets, int key_value) {
    if (num_buckets < 1 || num_buckets > 256) return 0;
    return tile_keys(ceil_log2_u32(num_buckets), key_value != 0, 0);
}

uint32_t gms_tile_keys(uint32_t num_buckets, int key_value, uint32_t n) {
    if (num_buckets < 1 || num_buckets > kSinglePassBuckets) return 0;
    return tile_keys(ceil_log2_u32(num_buckets), key_value != 0, n);
}

// ---------------------------------------------------------------- multisplit
gms_status gms_multisplit_create(uint32_t num_buckets, gms_multisplit_ctx** out) {
    if (!out) return GMS_ERR_INVALID_ARG;
    *out = nullptr;
    if (num_buckets < 1 || num_buckets > kMaxBuckets) return GMS_ERR_UNSUPPORTED_BUCKETS;
    gms_multisplit_ctx* c = new (std::nothrow) gms_multisplit_ctx();
    if (!c) return GMS_ERR_ALLOC;
    c->num_buckets = num_buckets;
    *out = c;
    return GMS_SUCCESS;
}

static void free_wide_temporaries(gms_multisplit_ctx* ctx) {
    if (ctx->wide_k) (void)hipFree(ctx->wide_k);
    if (ctx->wide_v) (void)hipFree(ctx->wide_v);
    ctx->wide_k = ctx->wide_v = nullptr;
}

gms_status gms_multisplit_allocate(gms_multisplit_ctx* ctx, uint32_t n_capacity, int key_value) {
    (void)key_value;  // the matrix is sized for both tile shapes
    if (!ctx) return GMS_ERR_INVALID_ARG;
    // The 64-bit pass temporaries hold capacity x 2 words: a new capacity
    // drops them, and the next call that needs them allocates them anew.
    free_wide_temporaries(ctx);
    const bool large = ctx->num_buckets > 256;
    gms_status st = scratch_alloc(ctx->s, matrix_elems(n_capacity, ctx->num_buckets));
    if (st == GMS_SUCCESS && large)
        st = large_scratch_alloc(ctx->large, n_capacity, ctx->num_buckets,
                                 ctx->num_buckets > kSinglePassBuckets);
    if (st != GMS_SUCCESS) {
        scratch_free(ctx->s);
        return st;
    }
    ctx->capacity = n_capacity;
    ctx->allocated = 1;
    return GMS_SUCCESS;
}

gms_status gms_multisplit_release(gms_multisplit_ctx* ctx) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    free_wide_temporaries(ctx);
    scratch_free(ctx->s);
    large_scratch_free(ctx->large);
    ctx->prof.destroy();
    ctx->allocated = 0;
    ctx->capacity = 0;
    return GMS_SUCCESS;
}

gms_status gms_multisplit_destroy(gms_multisplit_ctx* ctx) {
    if (!ctx) return GMS_SUCCESS;
    gms_multisplit_release(ctx);
    delete ctx;
    return GMS_SUCCESS;
}

// Pointers of one call of n 32-bit elements: present, values both or neither,
// no output overlapping an input or the other output.
static gms_status check_buffers(const uint32_t* kin, const uint32_t* vin, const uint32_t* kout,
                                const uint32_t* vout, uint32_t n) {
    if (n == 0) return GMS_SUCCESS;
    if (!kin || !kout) return GMS_ERR_INVALID_ARG;
    if ((vin == nullptr) != (vout == nullptr)) return GMS_ERR_INVALID_ARG;
    const size_t bytes = (size_t)n * 4;
    if (ranges_overlap(kin, bytes, kout, bytes)) return GMS_ERR_ALIAS;
    if (vin) {
        if (ranges_overlap(vin, bytes, vout, bytes) || ranges_overlap(kin, bytes, vout, bytes) ||
            ranges_overlap(vin, bytes, kout, bytes) || ranges_overlap(kout, bytes, vout, bytes))
            return GMS_ERR_ALIAS;
    }
    return GMS_SUCCESS;
}

static gms_status validate_call(gms_multisplit_ctx* ctx, const uint32_t* kin, const uint32_t* vin,
                                uint32_t* kout, uint32_t* vout, uint32_t n, gms_bucket_desc d) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if (!ctx->allocated) return GMS_ERR_NOT_ALLOCATED;
    if (n > ctx->capacity) return GMS_ERR_CAPACITY;
    gms_status st = check_desc(ctx->num_buckets, d);
    if (st != GMS_SUCCESS) return st;
    return check_buffers(kin, vin, kout, vout, n);
}

// The wave-level kernels only where the context was set to them: AUTO takes
// the block-wide ones, which measured as fast or faster at every bucket count
// and size on MI355X (the wave-level ones at most 5 % faster, at 2 buckets
// and up to 2^20 keys; 1.6-1.9x slower at 64 buckets and 2^24-2^27 keys).
static bool use_wms(const gms_multisplit_ctx* ctx) {
    if (ctx->num_buckets < 2 || ctx->num_buckets > kWmsMaxBuckets) return false;
    return ctx->variant == GMS_VARIANT_WMS || ctx->variant == GMS_VARIANT_DMS;
}

static gms_status run_ctx(gms_multisplit_ctx* ctx, const uint32_t* kin, const uint32_t* vin,
                          uint32_t* kout, uint32_t* vout, uint32_t n, gms_bucket_desc d,
                          uint32_t* d_offsets, hipStream_t s) {
    if (use_wms(ctx))
        return run_wms_multisplit(ctx->s, kin, vin, kout, vout, n, ctx->num_buckets, d, d_offsets,
                                  ctx->variant == GMS_VARIANT_DMS, s);
    if (ctx->num_buckets > kSinglePassBuckets)
        return run_large_multisplit(ctx->large, ctx->s, kin, vin, kout, vout, n, ctx->num_buckets,
                                    d, d_offsets, s);
    return run_multisplit(ctx->s, kin, vin, kout, vout, n, ctx->num_buckets, d, d_offsets, s,
                          &ctx->prof);
}

gms_status gms_multisplit(gms_multisplit_ctx* ctx, const uint32_t* kin, const uint32_t* vin,
                          uint32_t* kout, uint32_t* vout, uint32_t n, gms_bucket_desc d,
                          uint32_t* d_offsets, void* stream) {
    gms_status st = validate_call(ctx, kin, vin, kout, vout, n, d);
    if (st != GMS_SUCCESS) return st;
    return run_ctx(ctx, kin, vin, kout, vout, n, d, d_offsets, (hipStream_t)stream);
}

static gms_status multisplit_host_offsets(gms_multisplit_ctx* ctx, const uint32_t* kin,
                                          const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                                          uint32_t n, gms_bucket_desc d, uint32_t* h_offsets,
                                          void* stream) {
    gms_status st = validate_call(ctx, kin, vin, kout, vout, n, d);
    if (st != GMS_SUCCESS) return st;
    hipStream_t s = (hipStream_t)stream;
    uint32_t* staging = ctx->num_buckets > 256 ? ctx->large.offs : ctx->s.offs;
    st = run_ctx(ctx, kin, vin, kout, vout, n, d, h_offsets ? staging : nullptr, s);
    if (st != GMS_SUCCESS || !h_offsets) return st;
    if (hipMemcpyAsync(h_offsets, staging, (size_t)ctx->num_buckets * sizeof(uint32_t),
                       hipMemcpyDeviceToHost, s) != hipSuccess)
        return GMS_ERR_HIP;
    return hip_status(hipStreamSynchronize(s));
}

gms_status gms_multisplit_noalloc(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                  uint32_t* vout, uint32_t n, uint32_t num_buckets,
                                  gms_bucket_desc d, uint32_t* d_offsets, void* stream) {
    if (num_buckets < 1 || num_buckets > kSinglePassBuckets) return GMS_ERR_UNSUPPORTED_BUCKETS;
    gms_status st = check_desc(num_buckets, d);
    if (st != GMS_SUCCESS) return st;
    // The scan statuses, then the matrix.  Every matrix word is written
    // before it is read, and the prescan clears the statuses the scan reads,
    // so nothing is cleared between calls.
    const bool kv = vin != nullptr;
    const int lb = ceil_log2_u32(num_buckets);
    const uint32_t tile = tile_keys(lb, kv, n);
    const size_t m = ((size_t)1 << lb) * (((size_t)n + tile - 1) / tile);
    if (m > kNoAllocWords - kNoAllocHead) return GMS_ERR_CAPACITY;
    st = check_buffers(kin, vin, kout, vout, n);
    if (st != GMS_SUCCESS) return st;
    uint32_t* base = nullptr;
    st = noalloc_scratch(&base);
    if (st != GMS_SUCCESS) return st;
    Scratch s;
    s.status = reinterpret_cast<uint64_t*>(base);
    s.status_cap = kNoAllocStatus;
    s.hist = base + kNoAllocHead;
    s.hist_cap = kNoAllocWords - kNoAllocHead;
    return run_multisplit(s, kin, vin, kout, vout, n, num_buckets, d, d_offsets,
                          (hipStream_t)stream, nullptr);
}

gms_status gms_multisplit_key_only(gms_multisplit_ctx* ctx, const uint32_t* kin, uint32_t* kout,
                                   uint32_t n, gms_bucket_desc d, uint32_t* h_offsets,
                                   void* stream) {
    return multisplit_host_offsets(ctx, kin, nullptr, kout, nullptr, n, d, h_offsets, stream);
}

gms_status gms_multisplit_key_value(gms_multisplit_ctx* ctx, const uint32_t* kin,
                                    const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                                    uint32_t n, gms_bucket_desc d, uint32_t* h_offsets,
                                    void* stream) {
    if (n && (!vin || !vout)) return GMS_ERR_INVALID_ARG;
    return multisplit_host_offsets(ctx, kin, vin, kout, vout, n, d, h_offsets, stream);
}



gms_status gms_multisplit_wide(gms_multisplit_ctx* ctx, const void* kin, const void* vin,
                               void* kout, void* vout, uint32_t n, uint32_t key_bytes,
                               uint32_t value_bytes, gms_bucket_desc64 d, uint32_t* d_offsets,
                               uint32_t* h_offsets, void* stream) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if ((key_bytes != 4 && key_bytes != 8) || (value_bytes != 0 && value_bytes != 4 && value_bytes != 8))
        return GMS_ERR_INVALID_ARG;
    if (value_bytes && n && (!vin || !vout)) return GMS_ERR_INVALID_ARG;
    if (!value_bytes && (vin || vout)) return GMS_ERR_INVALID_ARG;
    const uint32_t kw = key_bytes / 4, vw = value_bytes / 4;
    hipStream_t s = (hipStream_t)stream;
    gms_status st;
    if (kw == 1 && vw <= 1) {  // the 32-bit path, whose identifiers take 32-bit parameters
        st = check_desc64(ctx->num_buckets, d, 32);
        if (st != GMS_SUCCESS) return st;
        const gms_bucket_desc d32{d.kind, (uint32_t)d.param, d.width};
        const uint32_t* k = static_cast<const uint32_t*>(kin);
        const uint32_t* v = static_cast<const uint32_t*>(vin);
        uint32_t* ko = static_cast<uint32_t*>(kout);
        uint32_t* vo = static_cast<uint32_t*>(vout);
        if (h_offsets) {
            st = multisplit_host_offsets(ctx, k, v, ko, vo, n, d32, h_offsets, stream);
            if (st != GMS_SUCCESS || !d_offsets) return st;
            return hipMemcpyAsync(d_offsets, h_offsets, (size_t)ctx->num_buckets * 4,
                                  hipMemcpyHostToDevice, s) == hipSuccess ? GMS_SUCCESS : GMS_ERR_HIP;
        }
        return gms_multisplit(ctx, k, v, ko, vo, n, d32, d_offsets, stream);
    }
    if (!ctx->allocated) return GMS_ERR_NOT_ALLOCATED;
    if (n > ctx->capacity) return GMS_ERR_CAPACITY;
    st = check_desc64(ctx->num_buckets, d, 32 * kw);
    if (st != GMS_SUCCESS) return st;
    if (n) {
        if (!kin || !kout) return GMS_ERR_INVALID_ARG;
        if ((kw == 2 && (!aligned8(kin) || !aligned8(kout))) ||
            (vw == 2 && (!aligned8(vin) || !aligned8(vout))))
            return GMS_ERR_INVALID_ARG;
        const size_t kb = (size_t)n * key_bytes, vb = (size_t)n * value_bytes;
        if (ranges_overlap(kin, kb, kout, kb)) return GMS_ERR_ALIAS;
        if (vw && (ranges_overlap(vin, vb, vout, vb) || ranges_overlap(kin, kb, vout, vb) ||
                   ranges_overlap(vin, vb, kout, kb) || ranges_overlap(kout, kb, vout, vb)))
            return GMS_ERR_ALIAS;
    }
    const bool large = ctx->num_buckets > 256;
    uint32_t* staging = large ? ctx->large.offs : ctx->s.offs;
    uint32_t* offs = d_offsets ? d_offsets : (h_offsets ? staging : nullptr);
    if (large) {
        if (!ctx->wide_k) {
            const size_t bytes = std::max<size_t>(ctx->capacity, 1) * 2 * sizeof(uint32_t);
            if (hipMalloc((void**)&ctx->wide_k, bytes) != hipSuccess) ctx->wide_k = nullptr;
            if (hipMalloc((void**)&ctx->wide_v, bytes) != hipSuccess) ctx->wide_v = nullptr;
            if (!ctx->wide_k || !ctx->wide_v) {
                if (ctx->wide_k) (void)hipFree(ctx->wide_k);
                if (ctx->wide_v) (void)hipFree(ctx->wide_v);
                ctx->wide_k = ctx->wide_v = nullptr;
                return GMS_ERR_ALLOC;
            }
        }
        st = run_wide_large(ctx->s, ctx->wide_k, ctx->wide_v, static_cast<const uint32_t*>(kin),
                            static_cast<const uint32_t*>(vin), static_cast<uint32_t*>(kout),
                            static_cast<uint32_t*>(vout), n, kw, vw, ctx->num_buckets, d, offs, s);
    } else {
        st = run_wide_multisplit(ctx->s, static_cast<const uint32_t*>(kin),
                                 static_cast<const uint32_t*>(vin), static_cast<uint32_t*>(kout),
                                 static_cast<uint32_t*>(vout), n, kw, vw, ctx->num_buckets, d,
                                 offs, s);
    }
    if (st != GMS_SUCCESS || !h_offsets) return st;
    if (hipMemcpyAsync(h_offsets, offs, (size_t)ctx->num_buckets * sizeof(uint32_t),
                       hipMemcpyDeviceToHost, s) != hipSuccess)
        return GMS_ERR_HIP;
    return hip_status(hipStreamSynchronize(s));
}

gms_status gms_multisplit_set_variant(gms_multisplit_ctx* ctx, int variant) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if (variant != GMS_VARIANT_AUTO && variant != GMS_VARIANT_BMS && variant != GMS_VARIANT_WMS &&
        variant != GMS_VARIANT_DMS)
        return GMS_ERR_INVALID_ARG;
    if ((variant == GMS_VARIANT_WMS || variant == GMS_VARIANT_DMS) &&
        (ctx->num_buckets < 2 || ctx->num_buckets > kWmsMaxBuckets))
        return GMS_ERR_UNSUPPORTED_BUCKETS;
    ctx->variant = variant;
    return GMS_SUCCESS;
}

uint32_t gms_subproblem_size_wms(uint32_t num_buckets, int key_value, uint32_t* size_block) {
    (void)key_value;  // one subproblem shape for key-only and key-value
    if (num_buckets < 2 || num_buckets > kWmsMaxBuckets) return 0;
    if (size_block) *size_block = wms_subproblem_keys() * (kThreads / 64);
    return wms_subproblem_keys();
}

gms_status gms_multisplit_set_profiling(gms_multisplit_ctx* ctx, int enable) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    ctx->prof.on = enable != 0;
    return GMS_SUCCESS;
}

gms_status gms_multisplit_profile(gms_multisplit_ctx* ctx, double* ms_prescan, double* ms_scan,
                                  double* ms_postscan, uint32_t* calls) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    gms_status st = ctx->prof.flush();
    if (st != GMS_SUCCESS) return st;
    if (ms_prescan) *ms_prescan = ctx->prof.ms[0];
    if (ms_scan) *ms_scan = ctx->prof.ms[1];
    if (ms_postscan) *ms_postscan = ctx->prof.ms[2];
    if (calls) *calls = ctx->prof.calls;
    ctx->prof.ms[0] = ctx->prof.ms[1] = ctx->prof.ms[2] = 0;
    ctx->prof.calls = 0;
    return GMS_SUCCESS;
}

gms_status gms_multisplit_prescan_matrix(gms_multisplit_ctx* ctx, const uint32_t* kin, uint32_t n,
                                         gms_bucket_desc d, int key_value, uint32_t* d_matrix,
                                         void* stream) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if (!ctx->allocated) return GMS_ERR_NOT_ALLOCATED;
    // no per-tile matrix where LSD passes over digits of the id run
    if (ctx->num_buckets > kSinglePassBuckets) return GMS_ERR_UNSUPPORTED_BUCKETS;
    if (n > ctx->capacity) return GMS_ERR_CAPACITY;
    gms_status st = check_desc(ctx->num_buckets, d);
    if (st != GMS_SUCCESS) return st;
    if (n == 0 || ctx->num_buckets <= 1) return GMS_SUCCESS;
    if (!kin || !d_matrix) return GMS_ERR_INVALID_ARG;
    uint32_t nt = 0;
    hipStream_t s = (hipStream_t)stream;
    st = run_prescan_only(ctx->s, kin, n, ctx->num_buckets, d, key_value != 0, s, &nt);
    if (st != GMS_SUCCESS) return st;
    // Rows past num_buckets (padding to a power of two) are always empty.
    return hip_status(launch_copy(ctx->s.hist, d_matrix, (size_t)ctx->num_buckets * nt, s));
}

gms_status gms_exclusive_scan_u32(gms_multisplit_ctx* ctx, const uint32_t* d_in, uint32_t* d_out,
                                  uint32_t n, void* stream) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if (!ctx->allocated) return GMS_ERR_NOT_ALLOCATED;
    if (n == 0) return GMS_SUCCESS;
    if (!d_in || !d_out) return GMS_ERR_INVALID_ARG;
    if (scan_partials_needed(n) > ctx->s.part_cap) return GMS_ERR_CAPACITY;
    if (d_in != d_out && ranges_overlap(d_in, (size_t)n * 4, d_out, (size_t)n * 4))
        return GMS_ERR_ALIAS;
    return hip_status(launch_exclusive_scan(d_in, d_out, n, ctx->s.part, (hipStream_t)stream));
}

gms_status gms_multisplit_stage_plan(gms_multisplit_ctx* ctx, const void* kin, const void* vin,
                                     void* kout, void* vout, uint32_t n, uint32_t key_bytes,
                                     uint32_t value_bytes, void* stream, gms_stage_plan* plan) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    uint32_t* staging = ctx->num_buckets > 256 ? ctx->large.offs : ctx->s.offs;
    return stage_plan(ctx->s, ctx->capacity, ctx->allocated, ctx->num_buckets, staging, kin, vin,
                      kout, vout, n, key_bytes, value_bytes, (hipStream_t)stream, plan);
}

gms_status gms_stage_scan(const gms_stage_plan* plan, void* stream) {
    if (!plan) return GMS_ERR_INVALID_ARG;
    if (plan->done) return GMS_SUCCESS;
    if (!plan->d_matrix || !plan->d_partials || plan->scan_words > 0xFFFFFFFFull)
        return GMS_ERR_INVALID_ARG;
    return hip_status(launch_exclusive_scan(plan->d_matrix, plan->d_matrix,
                                            (uint32_t)plan->scan_words, plan->d_partials,
                                            (hipStream_t)stream));
}

gms_status gms_stage_offsets(const gms_stage_plan* plan, uint32_t* d_out, void* stream) {
    if (!plan || !d_out) return GMS_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (plan->done) return hip_status(launch_fill(d_out, 0u, plan->num_buckets, s));
    // bucket b starts at the first word of its row of the scanned matrix
    return hip_status(launch_gather_offsets(plan->d_matrix, plan->num_tiles, plan->num_buckets,
                                            d_out, s));
}

// ---------------------------------------------------------------------- sort

gms_status gms_sort_create(gms_sort_ctx** out) {
    if (!out) return GMS_ERR_INVALID_ARG;
    *out = new (std::nothrow) gms_sort_ctx();
    return *out ? GMS_SUCCESS : GMS_ERR_ALLOC;
}

gms_status gms_sort_release(gms_sort_ctx* ctx) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    scratch_free(ctx->s);
    if (ctx->ktmp) (void)hipFree(ctx->ktmp);
    if (ctx->vtmp) (void)hipFree(ctx->vtmp);
    ctx->ktmp = ctx->vtmp = nullptr;
    ctx->allocated = 0;
    ctx->capacity = 0;
    return GMS_SUCCESS;
}

gms_status gms_sort_allocate(gms_sort_ctx* ctx, uint32_t n_capacity, uint32_t digit_bits,
                             int key_value) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    SortPass p[8];
    const int np = sort_schedule(digit_bits, p);
    if (np == 0) return GMS_ERR_INVALID_ARG;
    gms_sort_release(ctx);
    size_t me = 0;
    for (int i = 0; i < np; ++i) me = std::max(me, matrix_elems(n_capacity, 1u << p[i].bits));
    gms_status st = scratch_alloc(ctx->s, me);
    if (st != GMS_SUCCESS) return st;
    const size_t bytes = std::max<size_t>((size_t)n_capacity, 1) * sizeof(uint32_t);
    if (hipMalloc((void**)&ctx->ktmp, bytes) != hipSuccess) {
        ctx->ktmp = nullptr;
        gms_sort_release(ctx);
        return GMS_ERR_ALLOC;
    }
    if (key_value && hipMalloc((void**)&ctx->vtmp, bytes) != hipSuccess) {
        ctx->vtmp = nullptr;
        gms_sort_release(ctx);
        return GMS_ERR_ALLOC;
    }
    ctx->capacity = n_capacity;
    ctx->digit_bits = digit_bits;
    ctx->kv = key_value != 0;
    ctx->key_words = 1;
    ctx->value_words = key_value ? 1u : 0u;
    ctx->allocated = 1;
    return GMS_SUCCESS;
}

gms_status gms_sort_allocate_wide(gms_sort_ctx* ctx, uint32_t n_capacity, uint32_t digit_bits,
                                  uint32_t key_bytes, uint32_t value_bytes) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if ((key_bytes != 4 && key_bytes != 8) || (value_bytes != 0 && value_bytes != 4 && value_bytes != 8))
        return GMS_ERR_INVALID_ARG;
    const uint32_t kw = key_bytes / 4, vw = value_bytes / 4;
    SortPass p[16];
    const int np = wide_sort_passes(kw, digit_bits, p);
    if (np == 0) return GMS_ERR_INVALID_ARG;
    gms_sort_release(ctx);
    size_t me = 0;
    for (int i = 0; i < np; ++i) me = std::max(me, matrix_elems(n_capacity, 1u << p[i].bits));
    gms_status st = scratch_alloc(ctx->s, me);
    if (st != GMS_SUCCESS) return st;
    const size_t words = std::max<size_t>((size_t)n_capacity, 1);
    if (hipMalloc((void**)&ctx->ktmp, words * kw * sizeof(uint32_t)) != hipSuccess) {
        ctx->ktmp = nullptr;
        gms_sort_release(ctx);
        return GMS_ERR_ALLOC;
    }
    if (vw && hipMalloc((void**)&ctx->vtmp, words * vw * sizeof(uint32_t)) != hipSuccess) {
        ctx->vtmp = nullptr;
        gms_sort_release(ctx);
        return GMS_ERR_ALLOC;
    }
    ctx->capacity = n_capacity;
    ctx->digit_bits = digit_bits;
    ctx->kv = vw != 0;
    ctx->key_words = kw;
    ctx->value_words = vw;
    ctx->allocated = 1;
    return GMS_SUCCESS;
}

gms_status gms_sort_destroy(gms_sort_ctx* ctx) {
    if (!ctx) return GMS_SUCCESS;
    gms_sort_release(ctx);
    delete ctx;
    return GMS_SUCCESS;
}

gms_status gms_sort(gms_sort_ctx* ctx, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                    uint32_t* vout, uint32_t n, uint32_t digit_bits, void* stream) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if (!ctx->allocated) return GMS_ERR_NOT_ALLOCATED;
    if (n > ctx->capacity) return GMS_ERR_CAPACITY;
    SortPass p[8];
    const int np = sort_schedule(digit_bits, p);
    if (np == 0) return GMS_ERR_INVALID_ARG;
    const bool kv = vin != nullptr;
    if ((vin == nullptr) != (vout == nullptr)) return GMS_ERR_INVALID_ARG;
    if (kv && !ctx->vtmp) return GMS_ERR_NOT_ALLOCATED;
    if (n == 0) return GMS_SUCCESS;
    if (!kin || !kout) return GMS_ERR_INVALID_ARG;
    const size_t bytes = (size_t)n * 4;
    if (ranges_overlap(kin, bytes, kout, bytes) ||
        (kv && (ranges_overlap(vin, bytes, vout, bytes) || ranges_overlap(kin, bytes, vout, bytes) ||
                ranges_overlap(vin, bytes, kout, bytes) || ranges_overlap(kout, bytes, vout, bytes))))
        return GMS_ERR_ALIAS;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t* sk = kin;
    const uint32_t* sv = vin;
    for (int i = 0; i < np; ++i) {
        // The last pass writes the output; earlier passes alternate so that
        // the input is never written (multisplit_sort.cuh:320-559).
        const bool to_out = ((np - 1 - i) % 2) == 0;
        uint32_t* dk = to_out ? kout : ctx->ktmp;
        uint32_t* dv = kv ? (to_out ? vout : ctx->vtmp) : nullptr;
        gms_status st = run_multisplit(ctx->s, sk, sv, dk, dv, n, 1u << p[i].bits,
                                       gms_bucket_desc{GMS_BUCKET_DIGIT, p[i].shift, p[i].bits}, nullptr, s,
                                       nullptr);
        if (st != GMS_SUCCESS) return st;
        sk = dk;
        sv = dv;
    }
    return GMS_SUCCESS;
}

gms_status gms_sort_wide(gms_sort_ctx* ctx, const void* kin, void* kout, const void* vin,
                         void* vout, uint32_t n, uint32_t key_bytes, uint32_t value_bytes,
                         void* stream) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if (!ctx->allocated) return GMS_ERR_NOT_ALLOCATED;
    if ((key_bytes != 4 && key_bytes != 8) || (value_bytes != 0 && value_bytes != 4 && value_bytes != 8))
        return GMS_ERR_INVALID_ARG;
    const uint32_t kw = key_bytes / 4, vw = value_bytes / 4;
    if (kw == 1 && vw <= 1)
        return gms_sort(ctx, static_cast<const uint32_t*>(kin), static_cast<uint32_t*>(kout),
                        static_cast<const uint32_t*>(vin), static_cast<uint32_t*>(vout), n,
                        ctx->digit_bits, stream);
    if (n > ctx->capacity) return GMS_ERR_CAPACITY;
    if (kw > ctx->key_words || vw > ctx->value_words) return GMS_ERR_NOT_ALLOCATED;
    if (value_bytes && n && (!vin || !vout)) return GMS_ERR_INVALID_ARG;
    if (!value_bytes && (vin || vout)) return GMS_ERR_INVALID_ARG;
    if (n == 0) return GMS_SUCCESS;
    if (!kin || !kout) return GMS_ERR_INVALID_ARG;
    if ((kw == 2 && (!aligned8(kin) || !aligned8(kout))) ||
        (vw == 2 && (!aligned8(vin) || !aligned8(vout))))
        return GMS_ERR_INVALID_ARG;
    const size_t kb = (size_t)n * key_bytes, vb = (size_t)n * value_bytes;
    if (ranges_overlap(kin, kb, kout, kb) ||
        (vw && (ranges_overlap(vin, vb, vout, vb) || ranges_overlap(kin, kb, vout, vb) ||
                ranges_overlap(vin, vb, kout, kb) || ranges_overlap(kout, kb, vout, vb))))
        return GMS_ERR_ALIAS;
    return run_wide_sort(ctx->s, ctx->ktmp, ctx->vtmp, static_cast<const uint32_t*>(kin),
                         static_cast<const uint32_t*>(vin), static_cast<uint32_t*>(kout),
                         static_cast<uint32_t*>(vout), n, kw, vw, ctx->digit_bits,
                         (hipStream_t)stream);
}

// ---------------------------------------------------------------- compaction
gms_status gms_compaction_create(gms_compaction_ctx** out) {
    if (!out) return GMS_ERR_INVALID_ARG;
    *out = new (std::nothrow) gms_compaction_ctx();
    return *out ? GMS_SUCCESS : GMS_ERR_ALLOC;
}

gms_status gms_compaction_allocate(gms_compaction_ctx* ctx, uint32_t n_capacity, int key_value) {
    (void)key_value;
    if (!ctx) return GMS_ERR_INVALID_ARG;
    gms_status st = scratch_alloc(ctx->s, matrix_elems(n_capacity, 2));
    if (st != GMS_SUCCESS) return st;
    ctx->capacity = n_capacity;
    ctx->allocated = 1;
    return GMS_SUCCESS;
}

gms_status gms_compaction_release(gms_compaction_ctx* ctx) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    scratch_free(ctx->s);
    ctx->allocated = 0;
    ctx->capacity = 0;
    return GMS_SUCCESS;
}

gms_status gms_compaction_destroy(gms_compaction_ctx* ctx) {
    if (!ctx) return GMS_SUCCESS;
    gms_compaction_release(ctx);
    delete ctx;
    return GMS_SUCCESS;
}

gms_status gms_compaction(gms_compaction_ctx* ctx, const uint32_t* kin, const uint32_t* vin,
                          uint32_t* kout, uint32_t* vout, uint32_t n, gms_bucket_desc d,
                          uint32_t* h_count, uint32_t* d_count, void* stream) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if (!ctx->allocated) return GMS_ERR_NOT_ALLOCATED;
    if (n > ctx->capacity) return GMS_ERR_CAPACITY;
    gms_status st = check_desc(2, d);
    if (st != GMS_SUCCESS) return st;
    if ((vin == nullptr) != (vout == nullptr)) return GMS_ERR_INVALID_ARG;
    if (n && (!kin || !kout)) return GMS_ERR_INVALID_ARG;
    const size_t bytes = (size_t)n * 4;
    if (ranges_overlap(kin, bytes, kout, bytes) ||
        (vin && (ranges_overlap(vin, bytes, vout, bytes) || ranges_overlap(kin, bytes, vout, bytes) ||
                 ranges_overlap(vin, bytes, kout, bytes) || ranges_overlap(kout, bytes, vout, bytes))))
        return GMS_ERR_ALIAS;
    hipStream_t s = (hipStream_t)stream;
    // Offsets {0, count0} land in the scratch's offset slots; with n == 0 they are zero.
    st = run_multisplit(ctx->s, kin, vin, kout, vout, n, 2, d, ctx->s.offs, s, nullptr);
    if (st != GMS_SUCCESS) return st;
    if (d_count &&
        hipMemcpyAsync(d_count, ctx->s.offs + 1, sizeof(uint32_t), hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
        return GMS_ERR_HIP;
    if (h_count) {
        if (hipMemcpyAsync(h_count, ctx->s.offs + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s) !=
            hipSuccess)
            return GMS_ERR_HIP;
        return hip_status(hipStreamSynchronize(s));
    }
    return GMS_SUCCESS;
}

gms_status gms_compaction_wide(gms_compaction_ctx* ctx, const void* kin, const void* vin,
                               void* kout, void* vout, uint32_t n, uint32_t key_bytes,
                               uint32_t value_bytes, gms_bucket_desc64 d, uint32_t* h_count,
                               uint32_t* d_count, void* stream) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    if ((key_bytes != 4 && key_bytes != 8) || (value_bytes != 0 && value_bytes != 4 && value_bytes != 8))
        return GMS_ERR_INVALID_ARG;
    if (!value_bytes && (vin || vout)) return GMS_ERR_INVALID_ARG;
    if (value_bytes && n && (!vin || !vout)) return GMS_ERR_INVALID_ARG;
    const uint32_t kw = key_bytes / 4, vw = value_bytes / 4;
    gms_status st = check_desc64(2, d, 32 * kw);
    if (st != GMS_SUCCESS) return st;
    if (kw == 1 && vw <= 1)
        return gms_compaction(ctx, static_cast<const uint32_t*>(kin), static_cast<const uint32_t*>(vin),
                              static_cast<uint32_t*>(kout), static_cast<uint32_t*>(vout), n,
                              gms_bucket_desc{d.kind, (uint32_t)d.param, d.width}, h_count, d_count,
                              stream);
    if (!ctx->allocated) return GMS_ERR_NOT_ALLOCATED;
    if (n > ctx->capacity) return GMS_ERR_CAPACITY;
    if (n) {
        if (!kin || !kout) return GMS_ERR_INVALID_ARG;
        if ((kw == 2 && (!aligned8(kin) || !aligned8(kout))) ||
            (vw == 2 && (!aligned8(vin) || !aligned8(vout))))
            return GMS_ERR_INVALID_ARG;
        const size_t kb = (size_t)n * key_bytes, vb = (size_t)n * value_bytes;
        if (ranges_overlap(kin, kb, kout, kb) ||
            (vw && (ranges_overlap(vin, vb, vout, vb) || ranges_overlap(kin, kb, vout, vb) ||
                    ranges_overlap(vin, vb, kout, kb) || ranges_overlap(kout, kb, vout, vb))))
            return GMS_ERR_ALIAS;
    }
    hipStream_t s = (hipStream_t)stream;
    st = run_wide_multisplit(ctx->s, static_cast<const uint32_t*>(kin),
                             static_cast<const uint32_t*>(vin), static_cast<uint32_t*>(kout),
                             static_cast<uint32_t*>(vout), n, kw, vw, 2, d, ctx->s.offs, s);
    if (st != GMS_SUCCESS) return st;
    if (d_count &&
        hipMemcpyAsync(d_count, ctx->s.offs + 1, sizeof(uint32_t), hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
        return GMS_ERR_HIP;
    if (h_count) {
        if (hipMemcpyAsync(h_count, ctx->s.offs + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s) !=
            hipSuccess)
            return GMS_ERR_HIP;
        return hip_status(hipStreamSynchronize(s));
    }
    return GMS_SUCCESS;
}

gms_status gms_compaction_stage_plan(gms_compaction_ctx* ctx, const void* kin, const void* vin,
                                     void* kout, void* vout, uint32_t n, uint32_t key_bytes,
                                     uint32_t value_bytes, void* stream, gms_stage_plan* plan) {
    if (!ctx) return GMS_ERR_INVALID_ARG;
    return stage_plan(ctx->s, ctx->capacity, ctx->allocated, 2, ctx->s.offs, kin, vin, kout, vout,
                      n, key_bytes, value_bytes, (hipStream_t)stream, plan);
}

// ----------------------------------------------------------------- histogram
gms_status gms_histogram_u32(const uint32_t* kin, uint32_t n, uint32_t* d_hist, uint32_t num_bins,
                             gms_bucket_desc d, void* stream) {
    if (num_bins < 1 || num_bins > 256) return GMS_ERR_UNSUPPORTED_BUCKETS;
    gms_status st = check_desc(num_bins, d);
    if (st != GMS_SUCCESS) return st;
    if (!d_hist) return GMS_ERR_INVALID_ARG;
    if (n == 0) return GMS_SUCCESS;
    if (!kin) return GMS_ERR_INVALID_ARG;
    return hip_status(launch_histogram_u32(kin, n, d_hist, num_bins, d, (hipStream_t)stream));
}

gms_status gms_histogram_even_f32(const float* x, uint32_t n, int32_t* d_hist, uint32_t num_bins,
                                  float lower, float upper, float inv_delta, void* stream) {
    (void)upper;  // the reference passes it but bins only with lower and inv_delta
    if (num_bins < 1 || num_bins > 256) return GMS_ERR_UNSUPPORTED_BUCKETS;
    if (!d_hist) return GMS_ERR_INVALID_ARG;
    if (n == 0) return GMS_SUCCESS;
    if (!x) return GMS_ERR_INVALID_ARG;
    return hip_status(
        launch_histogram_even_f32(x, n, d_hist, num_bins, lower, inv_delta, (hipStream_t)stream));
}

gms_status gms_histogram_wide(const void* d_keys, uint32_t n, uint32_t key_bytes, uint32_t* d_hist,
                              uint32_t num_bins, gms_bucket_desc64 d, void* stream) {
    if (key_bytes != 4 && key_bytes != 8) return GMS_ERR_INVALID_ARG;
    if (num_bins < 1 || num_bins > 256) return GMS_ERR_UNSUPPORTED_BUCKETS;
    gms_status st = check_desc64(num_bins, d, 8 * key_bytes);
    if (st != GMS_SUCCESS) return st;
    if (key_bytes == 4)
        return gms_histogram_u32(static_cast<const uint32_t*>(d_keys), n, d_hist, num_bins,
                                 gms_bucket_desc{d.kind, (uint32_t)d.param, d.width}, stream);
    if (!d_hist) return GMS_ERR_INVALID_ARG;
    if (n == 0) return GMS_SUCCESS;
    if (!d_keys || !aligned8(d_keys)) return GMS_ERR_INVALID_ARG;
    return hip_status(launch_histogram_u64(static_cast<const uint64_t*>(d_keys), n, d_hist,
                                           num_bins, AnyF64::make(d.kind, d.param, d.width, num_bins),
                                           (hipStream_t)stream));
}

gms_status gms_histogram_range_f32(const float* d_samples, uint32_t n, int32_t* d_hist,
                                   uint32_t num_levels, const float* d_levels, void* stream) {
    if (num_levels < 2 || num_levels > 257) return GMS_ERR_UNSUPPORTED_BUCKETS;
    if (!d_hist || !d_levels) return GMS_ERR_INVALID_ARG;
    if (n == 0) return GMS_SUCCESS;
    if (!d_samples) return GMS_ERR_INVALID_ARG;
    return hip_status(launch_histogram_range_f32(d_samples, n, d_hist, num_levels, d_levels,
                                                 (hipStream_t)stream));
}

gms_status gms_histogram_range_u32(const uint32_t* d_keys, uint32_t n, int32_t* d_hist,
                                   uint32_t num_levels, const uint32_t* d_levels, void* stream) {
    if (num_levels < 2 || num_levels > 257) return GMS_ERR_UNSUPPORTED_BUCKETS;
    if (!d_hist || !d_levels) return GMS_ERR_INVALID_ARG;
    if (n == 0) return GMS_SUCCESS;
    if (!d_keys) return GMS_ERR_INVALID_ARG;
    return hip_status(launch_histogram_range_u32(d_keys, n, d_hist, num_levels, d_levels,
                                                 (hipStream_t)stream));
}

// ------------------------------------------------------------------- utility
gms_status gms_copy_u32(const uint32_t* src, uint32_t* dst, size_t n, void* stream) {
    if (n == 0) return GMS_SUCCESS;
    if (!src || !dst) return GMS_ERR_INVALID_ARG;
    if (ranges_overlap(src, n * 4, dst, n * 4)) return GMS_ERR_ALIAS;
    return hip_status(launch_copy(src, dst, n, (hipStream_t)stream));
}

}  // extern "C"
