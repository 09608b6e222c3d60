// gpumultisplit.hpp -- the reference's C++ API rebuilt on the C ABI of
// libgpumultisplit.so (gpumultisplit.h).
//
// Same names, argument order and meaning as the reference's header-only API
// (src/api/multisplit.cuh, multisplit_sort.cuh, multisplit_sort_pairs.cuh,
// src/kernels/compaction/multisplit2_compaction.cuh), so a caller switches by
// changing its include.  Differences, all deliberate:
//  * Errors throw gms::error (the reference returns void and checks nothing).
//  * The call-time n may be anything up to the allocate-time n (the reference
//    reuses the grid frozen at allocate time, so a different n breaks).
//  * Every call takes an optional trailing stream (hipStream_t converts to
//    void*); the default is the null stream, as in the reference.
//  * Contexts free their device memory in the destructor as well (the
//    reference's destructors are no-ops); *_release_memory still works.
//  * A bucket functor crosses the C ABI as a gms_bucket_desc.  The
//    reference's functors are provided (ms_bucket_sort_Nbit at global scope
//    as in multisplit_sort.cuh; delta_func, identity_bucket, test_bucket --
//    which the reference defines in its drivers -- and range_bucket in
//    namespace gms) and run the library's precompiled kernels.  Any other
//    functor type (a hash, a table lookup, ...) runs as device code when the
//    caller includes gpumultisplit_device.hpp instead of this header and
//    compiles with hipcc: multisplits and compactions then instantiate the
//    library's kernel templates with it.  Alternatively a functor equal to a
//    built-in kind can specialise gms::bucket_traits<T> with
//      static gms_bucket_desc desc(const T&, uint32_t num_buckets);
//  * KeyT and ValueT of multisplits, sorts, compactions and functor
//    histograms may be 32- or 64-bit (any 8-byte operand takes the 64-bit
//    kernels).  A 64-bit identifier parameter (a delta above 2^32) needs
//    gms::delta_func64.  Sorts of 64-bit keys or values are
//    allocated through the typed forms, e.g.
//    multisplit_sort_8bit_allocate_pairs<uint64_t, uint64_t>(n, ctx).
#ifndef GPUMULTISPLIT_HPP_
#define GPUMULTISPLIT_HPP_

#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "gpumultisplit.h"

namespace gms {

class error : public std::runtime_error {
  public:
    error(gms_status s, const char* where)
        : std::runtime_error(std::string(where) + ": " + gms_status_string(s)), status(s) {}
    gms_status status;
};

inline void check(gms_status s, const char* where) {
    if (s != GMS_SUCCESS) throw error(s, where);
}

// ---------------------------------------------------------------- functors
// key / delta_ (src/main/main_multisplit.cu:41-47).
struct delta_func {
    delta_func(uint32_t delta) : delta_(delta) {}
    uint32_t delta_;
    uint32_t operator()(uint32_t a) const { return a / delta_; }
};

// key itself (src/main/main_multisplit.cu:34-39).
template <typename key_type>
struct identity_bucket {
    uint32_t operator()(key_type a) const { return uint32_t(a); }
};

// key & ((1 << flag) - 1) (src/main/main_compaction.cu:27-32).
template <uint32_t flag, typename key_type>
struct test_bucket {
    uint32_t operator()(key_type a) const { return a & ((1u << flag) - 1); }
};

// key / delta_ with a 64-bit delta (for 64-bit keys).
struct delta_func64 {
    delta_func64(uint64_t delta) : delta_(delta) {}
    uint64_t delta_;
    uint32_t operator()(uint64_t a) const { return (uint32_t)(a / delta_); }
};

// floor(key * B / 2^32) on 32-bit keys, floor(key * B / 2^64) on 64-bit ones:
// B equal key ranges.
struct range_bucket {
    explicit range_bucket(uint32_t num_buckets) : num_buckets_(num_buckets) {}
    uint32_t num_buckets_;
    uint32_t operator()(uint32_t a) const {
        return (uint32_t)(((uint64_t)a * num_buckets_) >> 32);
    }
};

// No descriptor: the functor runs as device code (gpumultisplit_device.hpp).
template <class T, class Enable = void>
struct bucket_traits {};

template <class T, class = void>
struct has_desc : std::false_type {};
template <class T>
struct has_desc<T, decltype((void)bucket_traits<T>::desc(std::declval<const T&>(), 0u))>
    : std::true_type {};
template <class F>
constexpr bool has_desc_v = has_desc<typename std::decay<F>::type>::value;

template <class T>
struct always_false : std::false_type {};

template <>
struct bucket_traits<delta_func> {
    static gms_bucket_desc desc(const delta_func& f, uint32_t) {
        return gms_bucket_desc{GMS_BUCKET_DELTA, f.delta_, 0};
    }
};
template <typename K>
struct bucket_traits<identity_bucket<K>> {
    static gms_bucket_desc desc(const identity_bucket<K>&, uint32_t) {
        return gms_bucket_desc{GMS_BUCKET_IDENTITY, 0, 0};
    }
};
template <uint32_t flag, typename K>
struct bucket_traits<test_bucket<flag, K>> {
    static gms_bucket_desc desc(const test_bucket<flag, K>&, uint32_t) {
        return gms_bucket_desc{GMS_BUCKET_LOWBITS, flag, 0};
    }
};
template <>
struct bucket_traits<range_bucket> {
    static gms_bucket_desc desc(const range_bucket& f, uint32_t num_buckets) {
        if (f.num_buckets_ != num_buckets)
            throw error(GMS_ERR_INVALID_ARG, "range_bucket: bucket count differs from the context's");
        return gms_bucket_desc{GMS_BUCKET_RANGE, 0, 0};
    }
};

template <class F>
inline gms_bucket_desc desc_of(const F& f, uint32_t num_buckets) {
    static_assert(has_desc_v<F>,
                  "this call takes a built-in identifier (a functor with a gms_bucket_desc)");
    return bucket_traits<typename std::decay<F>::type>::desc(f, num_buckets);
}

// The 64-bit descriptor: bucket_traits<T>::desc64 where a functor has one,
// else its 32-bit descriptor widened.
template <class T, class = void>
struct has_desc64 : std::false_type {};
template <class T>
struct has_desc64<T, decltype((void)bucket_traits<T>::desc64(std::declval<const T&>(), 0u))>
    : std::true_type {};

template <class F>
inline gms_bucket_desc64 desc64_of(const F& f, uint32_t num_buckets) {
    using T = typename std::decay<F>::type;
    static_assert(has_desc_v<F>,
                  "this call takes a built-in identifier (a functor with a gms_bucket_desc)");
    if constexpr (has_desc64<T>::value) {
        return bucket_traits<T>::desc64(f, num_buckets);
    } else {
        const gms_bucket_desc d = bucket_traits<T>::desc(f, num_buckets);
        return gms_bucket_desc64{d.kind, d.width, d.param};
    }
}

template <>
struct bucket_traits<delta_func64> {
    static gms_bucket_desc desc(const delta_func64& f, uint32_t) {
        if (f.delta_ > 0xFFFFFFFFull)
            throw error(GMS_ERR_INVALID_ARG, "delta_func64: delta needs 64-bit keys");
        return gms_bucket_desc{GMS_BUCKET_DELTA, (uint32_t)f.delta_, 0};
    }
    static gms_bucket_desc64 desc64(const delta_func64& f, uint32_t) {
        return gms_bucket_desc64{GMS_BUCKET_DELTA, 0, f.delta_};
    }
};

template <class T>
inline void require_32bit(const char* what) {
    static_assert(sizeof(T) == 4, "keys and values must be 32-bit");
    (void)what;
}

template <class T>
constexpr uint32_t elem_bytes() {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "keys and values must be 32- or 64-bit");
    return (uint32_t)sizeof(T);
}

}  // namespace gms

// (key >> x) & (2^N - 1) (src/api/multisplit_sort.cuh:63-98).
#define GMS_DIGIT_FUNCTOR(N, MASK)                                                      \
    template <uint32_t x, typename key_type>                                            \
    struct ms_bucket_sort_##N##bit {                                                    \
        uint32_t operator()(key_type a) const { return uint32_t((a >> x) & MASK); }     \
    };                                                                                  \
    namespace gms {                                                                     \
    template <uint32_t x, typename K>                                                   \
    struct bucket_traits<::ms_bucket_sort_##N##bit<x, K>> {                             \
        static gms_bucket_desc desc(const ::ms_bucket_sort_##N##bit<x, K>&, uint32_t) { \
            return gms_bucket_desc{GMS_BUCKET_DIGIT, x, N};                             \
        }                                                                               \
    };                                                                                  \
    }
GMS_DIGIT_FUNCTOR(2, 0x03)
GMS_DIGIT_FUNCTOR(4, 0x0F)
GMS_DIGIT_FUNCTOR(5, 0x1F)
GMS_DIGIT_FUNCTOR(6, 0x3F)
GMS_DIGIT_FUNCTOR(7, 0x7F)
GMS_DIGIT_FUNCTOR(8, 0xFF)
#undef GMS_DIGIT_FUNCTOR

// ---------------------------------------------------------------- multisplit
// src/api/multisplit.cuh:29-47.
class multisplit_context {
  public:
    uint32_t size_sub_prob = 0;  // keys per tile of the last allocation
    uint32_t num_sub_prob = 0;   // tiles at the allocated capacity
    uint32_t num_buckets;
    gms_multisplit_ctx* handle = nullptr;

    explicit multisplit_context(uint32_t n_buckets) : num_buckets(n_buckets) {
        gms::check(gms_multisplit_create(n_buckets, &handle), "multisplit_context");
    }
    ~multisplit_context() { gms_multisplit_destroy(handle); }
    multisplit_context(const multisplit_context&) = delete;
    multisplit_context& operator=(const multisplit_context&) = delete;
};

namespace gms {
inline void ms_allocate(uint32_t n, multisplit_context& ctx, int kv, const char* what) {
    check(gms_multisplit_allocate(ctx.handle, n, kv), what);
    ctx.size_sub_prob = gms_subproblem_size(ctx.num_buckets, kv);
    ctx.num_sub_prob = ctx.size_sub_prob ? (n + ctx.size_sub_prob - 1) / ctx.size_sub_prob : 0;
}
}  // namespace gms

// src/api/multisplit.cuh:50-61.
inline void multisplit_allocate_key_only(uint32_t num_elements, multisplit_context& context) {
    gms::ms_allocate(num_elements, context, 0, "multisplit_allocate_key_only");
}
// src/api/multisplit.cuh:63-74.
inline void multisplit_allocate_key_value(uint32_t num_elements, multisplit_context& context) {
    gms::ms_allocate(num_elements, context, 1, "multisplit_allocate_key_value");
}
// src/api/multisplit.cuh:76-79.
inline void multisplit_release_memory(multisplit_context& context) {
    gms::check(gms_multisplit_release(context.handle), "multisplit_release_memory");
}

// src/api/multisplit.cuh:81-197.  h_offsets (host, B entries) makes the call blocking.
template <typename KeyT, typename bucket_t>
inline void multisplit_key_only(KeyT* d_key_in, KeyT* d_key_out, uint32_t num_elements,
                                multisplit_context& context, bucket_t bucket_identifier,
                                uint32_t* h_offsets = nullptr, void* stream = nullptr) {
    if constexpr (!gms::has_desc_v<bucket_t>) {
#ifdef GMS_DEVICE_FUNCTORS
        gms::check(gms::device::multisplit(context.handle, d_key_in, (const KeyT*)nullptr,
                                           d_key_out, (KeyT*)nullptr, num_elements,
                                           bucket_identifier, h_offsets, stream),
                   "multisplit_key_only");
#else
        static_assert(gms::always_false<bucket_t>::value,
                      "this functor has no gms_bucket_desc: include gpumultisplit_device.hpp "
                      "(compiled with hipcc) to run it as device code");
#endif
    } else if constexpr (gms::elem_bytes<KeyT>() == 8) {
        gms::check(gms_multisplit_wide(context.handle, d_key_in, nullptr, d_key_out, nullptr,
                                       num_elements, 8, 0,
                                       gms::desc64_of(bucket_identifier, context.num_buckets),
                                       nullptr, h_offsets, stream),
                   "multisplit_key_only");
    } else {
        gms::check(gms_multisplit_key_only(context.handle, reinterpret_cast<const uint32_t*>(d_key_in),
                                           reinterpret_cast<uint32_t*>(d_key_out), num_elements,
                                           gms::desc_of(bucket_identifier, context.num_buckets),
                                           h_offsets, stream),
                   "multisplit_key_only");
    }
}

// src/api/multisplit.cuh:199-313.
template <typename KeyT, typename ValueT, typename bucket_t>
inline void multisplit_key_value(KeyT* d_key_in, ValueT* d_value_in, KeyT* d_key_out,
                                 ValueT* d_value_out, uint32_t num_elements,
                                 multisplit_context& context, bucket_t bucket_identifier,
                                 uint32_t* h_offsets = nullptr, void* stream = nullptr) {
    constexpr uint32_t kb = gms::elem_bytes<KeyT>(), vb = gms::elem_bytes<ValueT>();
    if constexpr (!gms::has_desc_v<bucket_t>) {
#ifdef GMS_DEVICE_FUNCTORS
        if (num_elements && !d_value_in) throw gms::error(GMS_ERR_INVALID_ARG, "multisplit_key_value");
        gms::check(gms::device::multisplit(context.handle, d_key_in, d_value_in, d_key_out,
                                           d_value_out, num_elements, bucket_identifier, h_offsets,
                                           stream),
                   "multisplit_key_value");
#else
        static_assert(gms::always_false<bucket_t>::value,
                      "this functor has no gms_bucket_desc: include gpumultisplit_device.hpp "
                      "(compiled with hipcc) to run it as device code");
#endif
    } else if constexpr (kb == 8 || vb == 8) {
        gms::check(gms_multisplit_wide(context.handle, d_key_in, d_value_in, d_key_out, d_value_out,
                                       num_elements, kb, vb,
                                       gms::desc64_of(bucket_identifier, context.num_buckets),
                                       nullptr, h_offsets, stream),
                   "multisplit_key_value");
    } else {
        gms::check(gms_multisplit_key_value(
                       context.handle, reinterpret_cast<const uint32_t*>(d_key_in),
                       reinterpret_cast<const uint32_t*>(d_value_in),
                       reinterpret_cast<uint32_t*>(d_key_out), reinterpret_cast<uint32_t*>(d_value_out),
                       num_elements, gms::desc_of(bucket_identifier, context.num_buckets), h_offsets,
                       stream),
                   "multisplit_key_value");
    }
}

// subproblem_size_bms_key_only / _key_value (src/api/bms_api.h:23-89).
inline uint32_t subproblem_size_bms_key_only(uint32_t num_buckets) {
    return gms_subproblem_size(num_buckets, 0);
}
inline uint32_t subproblem_size_bms_key_value(uint32_t num_buckets) {
    return gms_subproblem_size(num_buckets, 1);
}

// subproblem_size_wms_key_only / _key_value (src/api/wms_api.h:25-83): keys per
// wave subproblem of the wave-level variant, keys per workgroup in size_block.
inline uint32_t subproblem_size_wms_key_only(uint32_t num_buckets, uint32_t& size_block) {
    size_block = 0;
    return gms_subproblem_size_wms(num_buckets, 0, &size_block);
}
inline uint32_t subproblem_size_wms_key_value(uint32_t num_buckets, uint32_t& size_block) {
    size_block = 0;
    return gms_subproblem_size_wms(num_buckets, 1, &size_block);
}

namespace gms {
// Block-wide (GMS_VARIANT_BMS), wave-level (GMS_VARIANT_WMS, 2..64 buckets),
// wave-level with direct scatter (GMS_VARIANT_DMS, 2..64 buckets) or automatic
// (GMS_VARIANT_AUTO, which is BMS) kernels for a context's multisplits.
inline void set_variant(multisplit_context& context, gms_variant v) {
    check(gms_multisplit_set_variant(context.handle, (int)v), "set_variant");
}
}  // namespace gms

// ---------------------------------------------------------------- radix sort
// src/api/multisplit_sort.cuh:100-116.
class ms_sort_context {
  public:
    gms_sort_ctx* handle = nullptr;
    ms_sort_context() { gms::check(gms_sort_create(&handle), "ms_sort_context"); }
    ~ms_sort_context() { gms_sort_destroy(handle); }
    ms_sort_context(const ms_sort_context&) = delete;
    ms_sort_context& operator=(const ms_sort_context&) = delete;
};

// src/api/multisplit_sort.cuh:312-318.
inline void multisplit_sort_release_memory(ms_sort_context& context) {
    gms::check(gms_sort_release(context.handle), "multisplit_sort_release_memory");
}

// multisplit_sort_Nbit_allocate[_pairs] (multisplit_sort.cuh:118-309),
// multisplit_sort_Nbit (:320-559), multisplit_sort_Nbit_pairs
// (multisplit_sort_pairs.cuh:21-260): keys in, keys out, values in, values out.
#define GMS_SORT_API(N)                                                                          \
    template <typename KeyT = uint32_t>                                                          \
    inline void multisplit_sort_##N##bit_allocate(uint32_t n_elements, ms_sort_context& ctx) {  \
        constexpr uint32_t kb = gms::elem_bytes<KeyT>();                                         \
        gms::check(kb == 4 ? gms_sort_allocate(ctx.handle, n_elements, N, 0)                     \
                           : gms_sort_allocate_wide(ctx.handle, n_elements, N, kb, 0),           \
                   "multisplit_sort_" #N "bit_allocate");                                        \
    }                                                                                            \
    template <typename KeyT = uint32_t, typename ValueT = uint32_t>                              \
    inline void multisplit_sort_##N##bit_allocate_pairs(uint32_t n_elements,                     \
                                                        ms_sort_context& ctx) {                  \
        constexpr uint32_t kb = gms::elem_bytes<KeyT>(), vb = gms::elem_bytes<ValueT>();         \
        gms::check(kb == 4 && vb == 4 ? gms_sort_allocate(ctx.handle, n_elements, N, 1)          \
                                      : gms_sort_allocate_wide(ctx.handle, n_elements, N, kb, vb), \
                   "multisplit_sort_" #N "bit_allocate_pairs");                                  \
    }                                                                                            \
    template <typename KeyT>                                                                     \
    inline void multisplit_sort_##N##bit(KeyT* d_key_in, KeyT* d_key_out, uint32_t n_elements,   \
                                         ms_sort_context& ctx, void* stream = nullptr) {         \
        constexpr uint32_t kb = gms::elem_bytes<KeyT>();                                         \
        gms::check(kb == 4 ? gms_sort(ctx.handle, reinterpret_cast<const uint32_t*>(d_key_in),   \
                                      reinterpret_cast<uint32_t*>(d_key_out), nullptr, nullptr,  \
                                      n_elements, N, stream)                                     \
                           : gms_sort_wide(ctx.handle, d_key_in, d_key_out, nullptr, nullptr,    \
                                           n_elements, kb, 0, stream),                           \
                   "multisplit_sort_" #N "bit");                                                 \
    }                                                                                            \
    template <typename KeyT, typename ValueT>                                                    \
    inline void multisplit_sort_##N##bit_pairs(KeyT* d_key_in, KeyT* d_key_out,                  \
                                               ValueT* d_value_in, ValueT* d_value_out,          \
                                               uint32_t n_elements, ms_sort_context& ctx,        \
                                               void* stream = nullptr) {                         \
        constexpr uint32_t kb = gms::elem_bytes<KeyT>(), vb = gms::elem_bytes<ValueT>();         \
        gms::check(kb == 4 && vb == 4                                                            \
                       ? gms_sort(ctx.handle, reinterpret_cast<const uint32_t*>(d_key_in),       \
                                  reinterpret_cast<uint32_t*>(d_key_out),                        \
                                  reinterpret_cast<const uint32_t*>(d_value_in),                 \
                                  reinterpret_cast<uint32_t*>(d_value_out), n_elements, N,       \
                                  stream)                                                        \
                       : gms_sort_wide(ctx.handle, d_key_in, d_key_out, d_value_in, d_value_out, \
                                       n_elements, kb, vb, stream),                              \
                   "multisplit_sort_" #N "bit_pairs");                                           \
    }
GMS_SORT_API(4)
GMS_SORT_API(5)
GMS_SORT_API(6)
GMS_SORT_API(7)
GMS_SORT_API(8)
#undef GMS_SORT_API

// ---------------------------------------------------------------- compaction
// src/kernels/compaction/multisplit2_compaction.cuh:477-489.
class compaction_context {
  public:
    gms_compaction_ctx* handle = nullptr;
    compaction_context() { gms::check(gms_compaction_create(&handle), "compaction_context"); }
    ~compaction_context() { gms_compaction_destroy(handle); }
    compaction_context(const compaction_context&) = delete;
    compaction_context& operator=(const compaction_context&) = delete;
};

// multisplit2_compaction.cuh:491-517.
inline void compaction_allocate_key_only(uint32_t num_elements, compaction_context& context) {
    gms::check(gms_compaction_allocate(context.handle, num_elements, 0),
               "compaction_allocate_key_only");
}
inline void compaction_allocate_key_value(uint32_t num_elements, compaction_context& context) {
    gms::check(gms_compaction_allocate(context.handle, num_elements, 1),
               "compaction_allocate_key_value");
}
// multisplit2_compaction.cuh:519-523.
inline void compaction_release_memory(compaction_context& context) {
    gms::check(gms_compaction_release(context.handle), "compaction_release_memory");
}

// multisplit2_compaction.cuh:525-548: stable, bucket 0 first; returns the
// bucket-0 count (blocking, as the reference).
template <typename key_t, typename bucket_t>
inline unsigned compaction_key_only(key_t* d_key_in, key_t* d_key_out, uint32_t num_elements,
                                    compaction_context& context, bucket_t bucket_identifier,
                                    void* stream = nullptr) {
    unsigned count = 0;
    if constexpr (!gms::has_desc_v<bucket_t>) {
#ifdef GMS_DEVICE_FUNCTORS
        gms::check(gms::device::compaction(context.handle, d_key_in, (const key_t*)nullptr,
                                           d_key_out, (key_t*)nullptr, num_elements,
                                           bucket_identifier, &count, stream),
                   "compaction_key_only");
#else
        static_assert(gms::always_false<bucket_t>::value,
                      "this functor has no gms_bucket_desc: include gpumultisplit_device.hpp "
                      "(compiled with hipcc) to run it as device code");
#endif
    } else if constexpr (gms::elem_bytes<key_t>() == 8) {
        gms::check(gms_compaction_wide(context.handle, d_key_in, nullptr, d_key_out, nullptr,
                                       num_elements, 8, 0, gms::desc64_of(bucket_identifier, 2),
                                       &count, nullptr, stream),
                   "compaction_key_only");
    } else {
        gms::check(gms_compaction(context.handle, reinterpret_cast<const uint32_t*>(d_key_in), nullptr,
                                  reinterpret_cast<uint32_t*>(d_key_out), nullptr, num_elements,
                                  gms::desc_of(bucket_identifier, 2), &count, nullptr, stream),
                   "compaction_key_only");
    }
    return count;
}

// multisplit2_compaction.cuh:550-573.
template <typename key_t, typename value_t, typename bucket_t>
inline unsigned compaction_key_value(key_t* d_key_in, value_t* d_value_in, key_t* d_key_out,
                                     value_t* d_value_out, uint32_t num_elements,
                                     compaction_context& context, bucket_t bucket_identifier,
                                     void* stream = nullptr) {
    unsigned count = 0;
    constexpr uint32_t kb = gms::elem_bytes<key_t>(), vb = gms::elem_bytes<value_t>();
    if constexpr (!gms::has_desc_v<bucket_t>) {
#ifdef GMS_DEVICE_FUNCTORS
        if (num_elements && !d_value_in) throw gms::error(GMS_ERR_INVALID_ARG, "compaction_key_value");
        gms::check(gms::device::compaction(context.handle, d_key_in, d_value_in, d_key_out,
                                           d_value_out, num_elements, bucket_identifier, &count,
                                           stream),
                   "compaction_key_value");
#else
        static_assert(gms::always_false<bucket_t>::value,
                      "this functor has no gms_bucket_desc: include gpumultisplit_device.hpp "
                      "(compiled with hipcc) to run it as device code");
#endif
    } else if constexpr (kb == 8 || vb == 8) {
        gms::check(gms_compaction_wide(context.handle, d_key_in, d_value_in, d_key_out, d_value_out,
                                       num_elements, kb, vb, gms::desc64_of(bucket_identifier, 2),
                                       &count, nullptr, stream),
                   "compaction_key_value");
    } else {
        gms::check(gms_compaction(context.handle, reinterpret_cast<const uint32_t*>(d_key_in),
                                  reinterpret_cast<const uint32_t*>(d_value_in),
                                  reinterpret_cast<uint32_t*>(d_key_out),
                                  reinterpret_cast<uint32_t*>(d_value_out), num_elements,
                                  gms::desc_of(bucket_identifier, 2), &count, nullptr, stream),
                   "compaction_key_value");
    }
    return count;
}

// ---------------------------------------------------------------- histogram
// multisplit2_histogram_even_* (src/kernels/histogram/multisplit2_histograms.cuh:36-392):
// float samples, bin = (uint)((x - lower) * inv_delta), added into d_hist[num_bins].
inline void multisplit2_histogram_even(const float* d_key_in, uint32_t num_elements, int* d_hist,
                                       uint32_t num_bins, float lower_level, float upper_level,
                                       float inverse_delta, void* stream = nullptr) {
    gms::check(gms_histogram_even_f32(d_key_in, num_elements, reinterpret_cast<int32_t*>(d_hist),
                                      num_bins, lower_level, upper_level, inverse_delta, stream),
               "multisplit2_histogram_even");
}

// multisplit2_histogram_range_* (src/kernels/histogram/multisplit2_histograms.cuh:397-699):
// num_levels ascending edges, bin i = [levels[i], levels[i+1]), samples outside
// [levels[0], levels[num_levels-1]) and NaN not counted; float or 32-bit integer
// samples (integers compared as uint32).
inline void multisplit2_histogram_range(const float* d_key_in, uint32_t num_elements, int* d_hist,
                                        const float* d_levels, uint32_t num_levels,
                                        void* stream = nullptr) {
    gms::check(gms_histogram_range_f32(d_key_in, num_elements, reinterpret_cast<int32_t*>(d_hist),
                                       num_levels, d_levels, stream),
               "multisplit2_histogram_range");
}
template <typename KeyT>
inline void multisplit2_histogram_range(const KeyT* d_key_in, uint32_t num_elements, int* d_hist,
                                        const KeyT* d_levels, uint32_t num_levels,
                                        void* stream = nullptr) {
    gms::require_32bit<KeyT>("key");
    gms::check(gms_histogram_range_u32(reinterpret_cast<const uint32_t*>(d_key_in), num_elements,
                                       reinterpret_cast<int32_t*>(d_hist), num_levels,
                                       reinterpret_cast<const uint32_t*>(d_levels), stream),
               "multisplit2_histogram_range");
}

// Integer-bin histogram through any bucket functor with a gms_bucket_desc
// (the 64/128/256-bin kernels of multisplit2_histograms.cuh take the bin as a
// bucket id of the key the same way).
template <typename KeyT, typename bucket_t>
inline void multisplit_histogram(const KeyT* d_key_in, uint32_t num_elements, uint32_t* d_hist,
                                 uint32_t num_bins, bucket_t bucket_identifier,
                                 void* stream = nullptr) {
    if constexpr (gms::elem_bytes<KeyT>() == 8) {
        gms::check(gms_histogram_wide(d_key_in, num_elements, 8, d_hist, num_bins,
                                      gms::desc64_of(bucket_identifier, num_bins), stream),
                   "multisplit_histogram");
    } else {
        gms::check(gms_histogram_u32(reinterpret_cast<const uint32_t*>(d_key_in), num_elements,
                                     d_hist, num_bins, gms::desc_of(bucket_identifier, num_bins),
                                     stream),
                   "multisplit_histogram");
    }
}

#endif  // GPUMULTISPLIT_HPP_
