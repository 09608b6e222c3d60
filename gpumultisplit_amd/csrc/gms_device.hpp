// Device building blocks for the gfx950 multisplit kernels: bucket
// identifiers, wave64 scans on DPP, and the XCD-aware tile order.
#pragma once

#ifdef GMS_HOST_EMULATION
// CPU build of the kernels for tests (tests/emu): same sources, emulated runtime.
#include "hip_emu.hpp"
#define GMS_UPDATE_DPP gms_emu::update_dpp
#define GMS_READLANE gms_emu::readlane
#define GMS_READFIRSTLANE(x) (x)
#define GMS_WAVE_SYNC() gms_emu::wave_sync()
#define GMS_BALLOT(p) gms_emu::ballot((p))
#else
#include <hip/hip_runtime.h>
#define GMS_UPDATE_DPP __builtin_amdgcn_update_dpp
#define GMS_READLANE __builtin_amdgcn_readlane
// 64-bit mask of the lanes of the wave whose predicate holds.
#define GMS_BALLOT(p) ((uint64_t)__ballot((p)))
// A value the whole wave holds alike, moved to a scalar register.
#define GMS_READFIRSTLANE(x) __builtin_amdgcn_readfirstlane(x)
// Orders the LDS accesses of one wave before and after this point across its
// lanes.  The lanes of a wave execute in lockstep, so this only stops the
// compiler from moving memory operations across it; it emits no instruction.
#define GMS_WAVE_SYNC() __builtin_amdgcn_wave_barrier()
#endif
#include <stdint.h>

#define GMS_HD __host__ __device__ __forceinline__

namespace gms {

// 16-byte load of streamed input: the line is not kept in the caches for
// reuse (no key or value of a tile is read twice by one kernel).
__device__ __forceinline__ uint4 ld_stream(const uint4* p) {
#ifdef GMS_HOST_EMULATION
    return *p;
#else
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
#endif
}

// ---------------------------------------------------------------------------
// Bucket identifiers.  Host+device so that the host check entry point runs
// exactly the integer code the kernels run.  Every identifier returns an id in
// [0, B): the ones that cannot overflow by construction (range, digit) do no
// clamping, the others clamp into B-1.
// ---------------------------------------------------------------------------

// floor(key * B / 2^32): the range identifier of the plain-multisplit sweep.
struct RangeF {
    uint32_t B;
    GMS_HD uint32_t operator()(uint32_t k) const {
        return (uint32_t)(((uint64_t)k * (uint64_t)B) >> 32);
    }
};

// key / delta (delta_func, src/main/main_multisplit.cu:41-47) by multiplication
// with a precomputed reciprocal: for 1 <= d < 2^32 let l = ceil(log2 d) and
// m = floor(2^32 * (2^l - d) / d) + 1; then for every 32-bit k
//   q = (t + ((k - t) >> min(l,1))) >> max(l-1,0),  t = mulhi(m, k)
// is exactly floor(k / d) (Granlund & Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994, section 4).
struct DeltaF {
    uint32_t m, sh1, sh2, bmax;
    static DeltaF make(uint32_t d, uint32_t num_buckets) {
        uint32_t l = 0;
        while (l < 32 && (1ull << l) < (uint64_t)d) ++l;
        DeltaF f;
        f.m = (uint32_t)(((((uint64_t)1 << l) - d) << 32) / d + 1);
        f.sh1 = l < 1 ? l : 1;
        f.sh2 = l > 0 ? l - 1 : 0;
        f.bmax = num_buckets - 1;
        return f;
    }
    GMS_HD uint32_t operator()(uint32_t k) const {
#ifdef __HIP_DEVICE_COMPILE__
        uint32_t t = __umulhi(m, k);
#else
        uint32_t t = (uint32_t)(((uint64_t)m * k) >> 32);
#endif
        uint32_t q = (t + ((k - t) >> sh1)) >> sh2;
        return q < bmax ? q : bmax;
    }
};

// (key >> shift) & mask with mask = B-1, B a power of two
// (ms_bucket_sort_Nbit<x>, src/api/multisplit_sort.cuh:63-98, and
// test_bucket<flag>, src/main/main_compaction.cu:27-32, with shift 0).
struct DigitF {
    uint32_t shift, mask;
    GMS_HD uint32_t operator()(uint32_t k) const { return (k >> shift) & mask; }
};

// key itself (identity_bucket, src/main/main_multisplit.cu:34-39).
struct IdentF {
    uint32_t bmax;
    GMS_HD uint32_t operator()(uint32_t k) const { return k < bmax ? k : bmax; }
};

// Any of the identifiers above, chosen at run time: the batched kernels give
// every instance of a batch its own.  One branch-free formula covers them all:
//   t = mulhi(k, m);  q = (t + (((k - t) >> sh1) & em)) >> sh2;
//   id = min(q & mask, bmax)
// range:    m = B, em = 0                       -> q = floor(k*B / 2^32)
// delta:    m, sh1, sh2 of DeltaF, em = ~0      -> q = floor(k / d)
// digit:    m = 0, em = ~0, sh1 = 0, sh2 = shift -> q = k >> shift
// identity: m = 0, em = ~0, sh1 = sh2 = 0       -> q = k
struct AnyF {
    uint32_t m, em, sh1, sh2, mask, bmax;
    GMS_HD uint32_t operator()(uint32_t k) const {
#ifdef __HIP_DEVICE_COMPILE__
        const uint32_t t = __umulhi(k, m);
#else
        const uint32_t t = (uint32_t)(((uint64_t)k * m) >> 32);
#endif
        const uint32_t q = ((t + (((k - t) >> sh1) & em)) >> sh2) & mask;
        return q < bmax ? q : bmax;
    }
    // kind: 0 range, 1 delta(param), 2 digit(shift = param, width bits;
    // width 0 = log2 B), 3 identity, 4 low bits(param) -- the gms_bucket_kind
    // numbering.  Arguments are assumed validated.
    static AnyF make(uint32_t kind, uint32_t param, uint32_t width, uint32_t num_buckets) {
        const uint32_t bmax = num_buckets - 1u;
        switch (kind) {
            case 0: return AnyF{num_buckets, 0u, 0u, 0u, ~0u, bmax};
            case 1: {
                const DeltaF d = DeltaF::make(param, num_buckets);
                return AnyF{d.m, ~0u, d.sh1, d.sh2, ~0u, bmax};
            }
            case 2: {
                const uint32_t mask = width ? (1u << width) - 1u : bmax;
                return AnyF{0u, ~0u, 0u, param, mask, bmax};
            }
            case 4: return AnyF{0u, ~0u, 0u, 0u, (1u << param) - 1u, bmax};
            default: return AnyF{0u, ~0u, 0u, 0u, ~0u, bmax};
        }
    }
};

// The identifiers for 64-bit keys, with the same branch-free form as AnyF:
//   t = mulhi64(k, m);  q = (t + (((k - t) >> sh1) & em)) >> sh2;
//   id = min(q & mask, bmax)
// range:    m = B, em = 0                        -> q = floor(k*B / 2^64)
// delta:    64-bit Granlund-Montgomery constants (the 64-bit case of DeltaF's
//           construction, same paper): l = ceil(log2 d),
//           m = floor(2^64 (2^l - d) / d) + 1, sh1 = min(l, 1), sh2 = max(l-1, 0)
//                                                -> q = floor(k / d)
// digit:    m = 0, sh2 = shift (< 64), mask = 2^width - 1 (or B - 1)
// identity: m = 0                                -> q = k
// low bits: m = 0, mask = 2^param - 1
struct AnyF64 {
    uint64_t m, em, mask;
    uint32_t sh1, sh2, bmax;
    GMS_HD uint32_t operator()(uint64_t k) const {
#ifdef __HIP_DEVICE_COMPILE__
        const uint64_t t = __umul64hi(k, m);
#else
        const uint64_t t = (uint64_t)(((unsigned __int128)k * m) >> 64);
#endif
        const uint64_t q = ((t + (((k - t) >> sh1) & em)) >> sh2) & mask;
        return q < bmax ? (uint32_t)q : bmax;
    }
    // kind numbering of gms_bucket_kind; arguments are assumed validated
    // (delta >= 1, shifts < 64).
    static AnyF64 make(uint32_t kind, uint64_t param, uint32_t width, uint32_t num_buckets) {
        const uint32_t bmax = num_buckets - 1u;
        switch (kind) {
            case 0: return AnyF64{num_buckets, 0ull, ~0ull, 0u, 0u, bmax};
            case 1: {
                uint32_t l = 0;
                while (l < 64 && ((unsigned __int128)1 << l) < (unsigned __int128)param) ++l;
                const unsigned __int128 two_l = (unsigned __int128)1 << l;
                const uint64_t m = (uint64_t)((((two_l - param) << 64) / param) + 1);
                return AnyF64{m, ~0ull, ~0ull, l < 1 ? l : 1u, l > 0 ? l - 1u : 0u, bmax};
            }
            case 2: {
                const uint64_t mask = width ? (1ull << width) - 1ull : (uint64_t)bmax;
                return AnyF64{0ull, ~0ull, mask, 0u, (uint32_t)param, bmax};
            }
            case 4: return AnyF64{0ull, ~0ull, (1ull << param) - 1ull, 0u, 0u, bmax};
            default: return AnyF64{0ull, ~0ull, ~0ull, 0u, 0u, bmax};
        }
    }
};

// One digit of a 64-bit key's bucket id (the 64-bit form of DigitOfF below).
struct DigitOf64F {
    AnyF64 base;
    uint32_t shift, mask;
    GMS_HD uint32_t operator()(uint64_t k) const { return (base(k) >> shift) & mask; }
};

// One digit of another identifier's bucket id: (base(k) >> shift) & mask,
// with base's clamp applied first.  The LSD passes of a multisplit into more
// than 256 buckets rank by these digits of the (up to 20-bit) bucket id.
struct DigitOfF {
    AnyF base;
    uint32_t shift, mask;
    GMS_HD uint32_t operator()(uint32_t k) const { return (base(k) >> shift) & mask; }
};

// ---------------------------------------------------------------------------
// Wave64 scans on DPP.  Rows are 16 lanes: four row_shr steps give prefixes
// inside a row, row_bcast:15 / row_bcast:31 carry them across rows.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += GMS_UPDATE_DPP(0u, x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += GMS_UPDATE_DPP(0u, x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += GMS_UPDATE_DPP(0u, x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += GMS_UPDATE_DPP(0u, x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += GMS_UPDATE_DPP(0u, x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += GMS_UPDATE_DPP(0u, x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
}

// Device-scope atomic load / store of a 64-bit status word and a 32-bit
// counter: coherent across the XCDs' L2 caches (the value travels in the
// word itself, so no ordering beyond the word is needed).
__device__ __forceinline__ uint64_t load_device_u64(const uint64_t* p) {
#ifdef GMS_HOST_EMULATION
    return __atomic_load_n(p, __ATOMIC_RELAXED);
#else
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}
__device__ __forceinline__ void store_device_u64(uint64_t* p, uint64_t v) {
#ifdef GMS_HOST_EMULATION
    __atomic_store_n(p, v, __ATOMIC_RELAXED);
#else
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}
__device__ __forceinline__ void store_device_u32(uint32_t* p, uint32_t v) {
#ifdef GMS_HOST_EMULATION
    __atomic_store_n(p, v, __ATOMIC_RELAXED);
#else
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}
// Short pause in a wait loop.
__device__ __forceinline__ void wait_pause() {
#ifndef GMS_HOST_EMULATION
    __builtin_amdgcn_s_sleep(2);
#endif
}

// Value of lane 63 broadcast to the wave (the total after an inclusive scan).
__device__ __forceinline__ uint32_t wave_last(uint32_t x) {
    return (uint32_t)GMS_READLANE((int)x, 63);
}

// Number of lanes below this one whose bit is set in m (v_mbcnt).
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
#ifdef GMS_HOST_EMULATION
    const uint32_t lane = threadIdx.x & 63u;
    return (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
#else
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
#endif
}

// Mask of the lanes of the wave whose LOG_B-bit id equals this lane's, among
// the lanes with `valid` set: one ballot per id bit.
template <int LOG_B>
__device__ __forceinline__ uint64_t wave_peers(uint32_t d, bool valid) {
    uint64_t peers = GMS_BALLOT(valid);
#pragma unroll
    for (int i = 0; i < LOG_B; ++i) {
        const bool bit = (d >> i) & 1u;
        const uint64_t m = GMS_BALLOT(valid && bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// Adds one to cnt[d] for every lane of the wave with `valid` set (cnt in LDS,
// private to the wave).  Keys that share a bucket across the wave (sorted
// runs, one dominant bucket) would make up to 64 lanes add to one address,
// which the LDS serialises: the lanes holding the id of the lowest pending
// lane are found with a ballot and added by that lane in one operation, for
// two such ids; the rest add one each.  On evenly spread ids the match is
// usually that lane alone, which then adds with the rest.
__device__ __forceinline__ void wave_count_add(uint32_t* cnt, uint32_t d, bool valid) {
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t rest = GMS_BALLOT(valid);
    bool pending = valid;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        if (rest == 0) break;  // the same in every lane
        const int first = __builtin_ctzll(rest);
        const uint32_t dl = (uint32_t)GMS_READLANE((int)d, first);
        const uint64_t m = GMS_BALLOT(pending && d == dl);
        if (__builtin_popcountll(m) > 1) {
            if (lane == (uint32_t)first) atomicAdd(&cnt[dl], (uint32_t)__builtin_popcountll(m));
            pending = pending && d != dl;
        }
        rest &= ~m;
    }
    if (pending) atomicAdd(&cnt[d], 1u);
}

// Whether the ids of the wave's lanes coincide: the id of the lowest valid
// lane, or failing that of the next one, is shared by more than 8 lanes.  A
// kernel probes this once per wave and tile and then counts with
// wave_count_add (coinciding ids) or one LDS add per lane (spread ids), which
// keeps the ballots off the path of evenly spread keys.
__device__ __forceinline__ bool wave_ids_coincide(uint32_t d, bool valid) {
    uint64_t rest = GMS_BALLOT(valid);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        if (rest == 0) break;  // the same in every lane
        const int first = __builtin_ctzll(rest);
        const uint32_t dl = (uint32_t)GMS_READLANE((int)d, first);
        const uint64_t m = GMS_BALLOT(valid && d == dl);
        if (__builtin_popcountll(m) > 8) return true;
        rest &= ~m;
    }
    return false;
}

// Exclusive scan across a block of NW waves; `total` receives the block sum.
// s_w must hold NW words; the call contains one __syncthreads and callers
// must separate consecutive uses of the same s_w by a barrier.
template <int NW>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* s_w, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(x);
    if (lane == 63u) s_w[wid] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        uint32_t v = s_w[i];
        pre += ((uint32_t)i < wid) ? v : 0u;
        tot += v;
    }
    total = tot;
    return pre + inc - x;
}

// ---------------------------------------------------------------------------
// Tile order.  Workgroups are handed to the 8 XCDs round-robin (workgroup i
// runs on XCD i mod 8), each XCD with its own L2.  Remap so that every XCD
// walks one contiguous range of tiles: neighbouring tiles then share L2 lines
// of the histogram matrix and of the per-bucket output runs.  A bijection on
// [0, num_tiles) for any num_tiles.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t num_tiles) {
    constexpr uint32_t NX = 8;
    const uint32_t q = num_tiles / NX, r = num_tiles % NX;
    const uint32_t x = bid % NX, i = bid / NX;
    return x * q + (x < r ? x : r) + i;
}

}  // namespace gms
