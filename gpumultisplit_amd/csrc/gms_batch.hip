// Batched multisplit: host engine and the gms_batch_* entry points.
#include <string.h>

#include <algorithm>
#include <memory>
#include <new>

#include "gms_batch.hpp"

namespace gms {

// ---------------------------------------------------------------- buffers
template <class T>
gms_status DevVec<T>::reserve(size_t n, hipStream_t s) {
    if (n <= cap) return GMS_SUCCESS;
    const size_t want = std::max<size_t>(n + n / 2, 64);
    if (p) {
        if (hipStreamSynchronize(s) != hipSuccess) return GMS_ERR_HIP;
        (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    if (hipMalloc((void**)&p, want * sizeof(T)) != hipSuccess) {
        p = nullptr;
        return GMS_ERR_ALLOC;
    }
    cap = want;
    return GMS_SUCCESS;
}

template <class T>
void DevVec<T>::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

template <class T>
gms_status PinnedVec<T>::reserve(size_t n) {
    if (n <= cap) return GMS_SUCCESS;
    const size_t want = std::max<size_t>(n + n / 2, 64);
    T* q = nullptr;
    if (hipHostMalloc((void**)&q, want * sizeof(T), hipHostMallocDefault) != hipSuccess)
        return GMS_ERR_ALLOC;
    if (p) {
        memcpy(q, p, size * sizeof(T));
        (void)hipHostFree(p);
    }
    p = q;
    cap = want;
    return GMS_SUCCESS;
}

template <class T>
void PinnedVec<T>::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = size = 0;
}

template struct DevVec<uint8_t>;
template struct PinnedVec<BatchInst>;
template struct PinnedVec<BatchWideInst>;
template struct PinnedVec<BatchWork>;

// ---------------------------------------------------------------- checks
gms_status check_batch_item(const gms_batch_item& it) {
    if (it.num_buckets < 1 || it.num_buckets > 256) return GMS_ERR_UNSUPPORTED_BUCKETS;
    gms_status st = check_desc(it.num_buckets, it.desc);
    if (st != GMS_SUCCESS) return st;
    if ((it.d_value_in == nullptr) != (it.d_value_out == nullptr)) return GMS_ERR_INVALID_ARG;
    if (it.n && (!it.d_key_in || !it.d_key_out)) return GMS_ERR_INVALID_ARG;
    return GMS_SUCCESS;
}

gms_status ByteRanges::check() {
    // Sorted by start, then a sweep that tracks the furthest end so far and
    // the furthest end of an output.
    std::sort(r.begin(), r.end(), [](const R& x, const R& y) { return x.a < y.a; });
    uintptr_t end_any = 0, end_out = 0;
    for (const R& x : r) {
        if (x.out ? x.a < end_any : x.a < end_out) return GMS_ERR_ALIAS;
        end_any = std::max(end_any, x.b);
        if (x.out) end_out = std::max(end_out, x.b);
    }
    return GMS_SUCCESS;
}

static void add_item_ranges(ByteRanges& R, const gms_batch_item& it) {
    if (it.n == 0) return;
    const size_t bytes = (size_t)it.n * 4;
    R.add(it.d_key_in, bytes, false);
    R.add(it.d_key_out, bytes, true);
    R.add(it.d_value_in, bytes, false);
    R.add(it.d_value_out, bytes, true);
    R.add(it.d_offsets, (size_t)it.num_buckets * 4, true);
}

gms_status check_batch_overlaps(const gms_batch_item* items, uint32_t count) {
    ByteRanges R;
    R.r.reserve((size_t)count * 5);
    for (uint32_t i = 0; i < count; ++i) add_item_ranges(R, items[i]);
    return R.check();
}

// ---------------------------------------------------------------- engine
static int class_of(uint32_t num_buckets) {
    const int l = ceil_log2_u32(num_buckets);
    return l < 1 ? 1 : l;
}

gms_status BatchEngine::reserve_tables(uint32_t instances, uint64_t work_items) {
    for (int k = 0; k < 2; ++k) {
        gms_status st = h_inst_[k].reserve(instances);
        if (st == GMS_SUCCESS) st = h_winst_[k].reserve(instances);
        if (st == GMS_SUCCESS) st = h_work_[k].reserve((size_t)work_items);
        if (st != GMS_SUCCESS) return st;
    }
    return GMS_SUCCESS;
}

gms_status BatchEngine::plan(const gms_batch_item* items, uint32_t count,
                             const BatchWideItem* wide, uint32_t wcount, int slot) {
    const uint32_t K = tuning.items_per_thread;
    if (K != 8 && K != 16 && K != 24 && K != 32) return GMS_ERR_INVALID_ARG;
    const uint32_t tile = (uint32_t)kThreads * K;
    slot_ = slot & 1;
    PinnedVec<BatchInst>& hi = h_inst_[slot_];
    PinnedVec<BatchWideInst>& hwi = h_winst_[slot_];
    PinnedVec<BatchWork>& hw = h_work_[slot_];

    // Group order: (class, kv, alignment); instances keep their order inside
    // a group.
    auto aligned = [&](uint32_t i) {
        return (((uintptr_t)items[i].d_key_in) & 15u) == 0 &&
               (((uintptr_t)items[i].d_value_in) & 15u) == 0;
    };
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; ++i) order[i] = i;
    auto key = [&](uint32_t i) {
        return class_of(items[i].num_buckets) * 4 + (items[i].d_value_in ? 2 : 0) +
               (aligned(i) ? 0 : 1);
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return key(a) < key(b); });

    uint64_t tiles = 0;
    for (uint32_t i = 0; i < count; ++i)
        tiles += items[i].n ? ((uint64_t)items[i].n + tile - 1) / tile : 1;
    for (uint32_t i = 0; i < wcount; ++i) {
        const uint32_t wt = wide_tile_elems(wide[i].kw, wide[i].vw);
        tiles += wide[i].n ? ((uint64_t)wide[i].n + wt - 1) / wt : 1;
    }
    if (tiles >= (1ull << 31)) return GMS_ERR_CAPACITY;
    gms_status st = hi.reserve(count);
    if (st == GMS_SUCCESS) st = hwi.reserve(wcount);
    if (st == GMS_SUCCESS) st = hw.reserve((size_t)tiles);
    if (st != GMS_SUCCESS) return st;

    groups.clear();
    uint64_t mbase = 0, obase = 0;
    uint32_t w = 0;
    for (uint32_t j = 0; j < count; ++j) {
        const gms_batch_item& it = items[order[j]];
        const int lb = class_of(it.num_buckets);
        const bool kv = it.d_value_in != nullptr;
        const bool al = aligned(order[j]);
        if (groups.empty() || groups.back().log_b != lb || groups.back().kv != kv ||
            groups.back().vec != al)
            groups.push_back(BatchGroup{lb, kv, al, w, w});
        const uint32_t nt = it.n ? (uint32_t)(((uint64_t)it.n + tile - 1) / tile) : 1u;
        BatchInst& I = hi.p[j];
        I.kin = it.d_key_in;
        I.vin = it.d_value_in;
        I.kout = it.d_key_out;
        I.vout = it.d_value_out;
        I.offs = it.d_offsets;
        I.n = it.n;
        I.num_tiles = nt;
        I.mbase = (uint32_t)mbase;
        I.obase = (uint32_t)obase;
        I.nb = it.num_buckets;
        I.f = AnyF::make(it.desc.kind, it.desc.param, it.desc.width, it.num_buckets);
        for (uint32_t t = 0; t < nt; ++t) hw.p[w++] = BatchWork{j, t};
        groups.back().w1 = w;
        mbase += ((uint64_t)1 << lb) * nt;
        obase += it.n;
        if (mbase >= 0xFFFFFFFFull || obase >= 0xFFFFFFFFull) return GMS_ERR_CAPACITY;
    }

    // 64-bit elements: groups by (class, key words, value words, alignment),
    // their work items and matrix blocks after the 32-bit ones.
    auto wal = [&](uint32_t i) {
        return (((uintptr_t)wide[i].kin) & 15u) == 0 && (((uintptr_t)wide[i].vin) & 15u) == 0;
    };
    auto wkey = [&](uint32_t i) {
        return ((class_of(wide[i].num_buckets) * 4 + (int)wide[i].kw) * 4 + (int)wide[i].vw) * 2 +
               (wal(i) ? 0 : 1);
    };
    std::vector<uint32_t> worder(wcount);
    for (uint32_t i = 0; i < wcount; ++i) worder[i] = i;
    std::stable_sort(worder.begin(), worder.end(),
                     [&](uint32_t a, uint32_t b) { return wkey(a) < wkey(b); });
    wide_groups.clear();
    for (uint32_t j = 0; j < wcount; ++j) {
        const BatchWideItem& it = wide[worder[j]];
        const int lb = class_of(it.num_buckets);
        const bool al = wal(worder[j]);
        if (wide_groups.empty() || wide_groups.back().log_b != lb || wide_groups.back().kw != it.kw ||
            wide_groups.back().vw != it.vw || wide_groups.back().vec != al)
            wide_groups.push_back(BatchWideGroup{lb, it.kw, it.vw, al, w, w});
        const uint32_t wt = wide_tile_elems(it.kw, it.vw);
        const uint32_t nt = it.n ? (uint32_t)(((uint64_t)it.n + wt - 1) / wt) : 1u;
        BatchWideInst& I = hwi.p[j];
        I.kin = it.kin;
        I.vin = it.vin;
        I.kout = it.kout;
        I.vout = it.vout;
        I.offs = it.offs;
        I.n = it.n;
        I.num_tiles = nt;
        I.mbase = (uint32_t)mbase;
        I.obase = (uint32_t)obase;
        I.nb = it.num_buckets;
        I.f = AnyF::make(it.desc.kind, (uint32_t)it.desc.param, it.desc.width, it.num_buckets);
        I.f64 = AnyF64::make(it.desc.kind, it.desc.param, it.desc.width, it.num_buckets);
        for (uint32_t t = 0; t < nt; ++t) hw.p[w++] = BatchWork{j, t};
        wide_groups.back().w1 = w;
        mbase += ((uint64_t)1 << lb) * nt;
        obase += it.n;
        if (mbase >= 0xFFFFFFFFull || obase >= 0xFFFFFFFFull) return GMS_ERR_CAPACITY;
    }
    hi.size = count;
    hwi.size = wcount;
    hw.size = w;
    num_instances = count;
    num_wide = wcount;
    num_work = w;
    m_total = mbase;
    keys_total = obase;
    return GMS_SUCCESS;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t BatchEngine::device_bytes() const {
    const size_t m = std::max<uint64_t>(m_total, 4);
    return align256(std::max<size_t>(num_instances, 1) * sizeof(BatchInst)) +
           align256(std::max<size_t>(num_wide, 1) * sizeof(BatchWideInst)) +
           align256(std::max<size_t>(num_work, 1) * sizeof(BatchWork)) + align256(m * 4) +
           align256(scan_partials_needed(m) * 4) + align256(16 * std::max<size_t>(groups.size(), 1) * 4);
}

BatchDevice BatchEngine::carve(void* base) const {
    const size_t m = std::max<uint64_t>(m_total, 4);
    char* p = (char*)base;
    BatchDevice d;
    d.inst = (BatchInst*)p;
    p += align256(std::max<size_t>(num_instances, 1) * sizeof(BatchInst));
    d.winst = (BatchWideInst*)p;
    p += align256(std::max<size_t>(num_wide, 1) * sizeof(BatchWideInst));
    d.work = (BatchWork*)p;
    p += align256(std::max<size_t>(num_work, 1) * sizeof(BatchWork));
    d.mat = (uint32_t*)p;
    p += align256(m * 4);
    d.part = (uint32_t*)p;
    p += align256(scan_partials_needed(m) * 4);
    d.queue = (uint32_t*)p;
    return d;
}

gms_status BatchEngine::upload(const BatchDevice& d, hipStream_t s) {
    if (num_cus_ == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess) return GMS_ERR_NO_DEVICE;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return GMS_ERR_HIP;
        num_cus_ = cus > 0 ? cus : 1;
    }
    if (num_instances &&
        hipMemcpyAsync(d.inst, h_inst_[slot_].p, (size_t)num_instances * sizeof(BatchInst),
                       hipMemcpyHostToDevice, s) != hipSuccess)
        return GMS_ERR_HIP;
    if (num_wide &&
        hipMemcpyAsync(d.winst, h_winst_[slot_].p, (size_t)num_wide * sizeof(BatchWideInst),
                       hipMemcpyHostToDevice, s) != hipSuccess)
        return GMS_ERR_HIP;
    if (num_work &&
        hipMemcpyAsync(d.work, h_work_[slot_].p, (size_t)num_work * sizeof(BatchWork),
                       hipMemcpyHostToDevice, s) != hipSuccess)
        return GMS_ERR_HIP;
    return GMS_SUCCESS;
}

uint32_t BatchEngine::grid_for(const BatchGroup& g, bool post) const {
    const uint32_t nw = g.w1 - g.w0;
    if (!tuning.persistent) return nw;
    int occ = 1;
    switch (tuning.items_per_thread) {
        case 8: occ = batch_occupancy_k8(g.log_b, g.kv, post); break;
        case 16: occ = batch_occupancy_k16(g.log_b, g.kv, post); break;
        case 24: occ = batch_occupancy_k24(g.log_b, g.kv, post); break;
        default: occ = batch_occupancy_k32(g.log_b, g.kv, post); break;
    }
    // At least 8 workgroups (one per XCD) whenever there are 8 items, so that
    // every XCD's range of items has a workgroup to drain it.
    const uint64_t resident = std::max<uint64_t>((uint64_t)occ * (uint64_t)num_cus_, 8);
    return (uint32_t)std::min<uint64_t>(nw, resident);
}

uint32_t BatchEngine::launches() const {
    if (num_work == 0) return 0;
    return (tuning.persistent && !groups.empty() ? 1u : 0u) +
           2u * (uint32_t)(groups.size() + wide_groups.size()) + (m_total ? 3u : 0u);
}

gms_status BatchEngine::launch(const BatchDevice& d, hipStream_t s) {
    if (num_work == 0) return GMS_SUCCESS;
    const uint32_t G = (uint32_t)groups.size();
    hipError_t e = tuning.persistent && G ? launch_fill(d.queue, 0u, 16 * G, s) : hipSuccess;
    auto one = [&](const BatchGroup& g, uint32_t gi, bool post) {
        BatchLaunch c{d.inst, d.work + g.w0, g.w1 - g.w0, grid_for(g, post), d.mat,
                      (uint32_t)m_total, (uint32_t)keys_total, d.queue + 16 * gi + (post ? 8 : 0),
                      g.log_b, g.kv, g.vec, tuning.persistent != 0, s,
                      keys_total >= kStreamLoadMin};
        switch (tuning.items_per_thread) {
            case 8: return post ? launch_batch_postscan_k8(c) : launch_batch_prescan_k8(c);
            case 16: return post ? launch_batch_postscan_k16(c) : launch_batch_prescan_k16(c);
            case 24: return post ? launch_batch_postscan_k24(c) : launch_batch_prescan_k24(c);
            default: return post ? launch_batch_postscan_k32(c) : launch_batch_prescan_k32(c);
        }
    };
    auto wide_one = [&](const BatchWideGroup& g, bool post) {
        const BatchWideLaunch c{d.winst, d.work + g.w0, g.w1 - g.w0, d.mat, (uint32_t)m_total,
                                (uint32_t)keys_total, g.log_b, g.kw, g.vw, g.vec, s};
        return launch_batch_wide(c, post);
    };
    for (uint32_t gi = 0; gi < G && e == hipSuccess; ++gi) e = one(groups[gi], gi, false);
    for (size_t gi = 0; gi < wide_groups.size() && e == hipSuccess; ++gi)
        e = wide_one(wide_groups[gi], false);
    if (e == hipSuccess)
        e = launch_exclusive_scan(d.mat, d.mat, (uint32_t)m_total, d.part, s);
    for (uint32_t gi = 0; gi < G && e == hipSuccess; ++gi) e = one(groups[gi], gi, true);
    for (size_t gi = 0; gi < wide_groups.size() && e == hipSuccess; ++gi)
        e = wide_one(wide_groups[gi], true);
    return hip_status(e);
}

void BatchEngine::release() {
    for (int i = 0; i < 2; ++i) {
        h_inst_[i].release();
        h_winst_[i].release();
        h_work_[i].release();
    }
    num_work = num_instances = num_wide = 0;
    m_total = keys_total = 0;
    groups.clear();
    wide_groups.clear();
}

}  // namespace gms

using namespace gms;

// A batch is a sequence of steps, each one batched multisplit (BatchEngine):
// step 0 holds the multisplits and the first pass of every sort, step p the
// p-th pass of the sorts that have one.
struct gms_batch {
    std::vector<std::unique_ptr<BatchEngine>> eng;  // one per step
    std::vector<BatchDevice> dev;
    DevVec<uint8_t> mem;  // every step's tables
    DevVec<uint8_t> tmp;  // the sorts' second pass buffers
    std::vector<gms_batch_item> items;
    std::vector<BatchWideItem> wides;
    std::vector<gms_batch_sort_item> sorts;
    BatchTuning tuning;
    bool planned = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;

    struct Step {
        std::vector<gms_batch_item> items;
        std::vector<BatchWideItem> wides;
    };

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    void release() {
        for (auto& e : eng) e->release();
        eng.clear();
        dev.clear();
        mem.release();
        tmp.release();
    }
    static uint32_t key_words(const gms_batch_sort_item& q) { return q.key_bytes == 8 ? 2u : 1u; }
    static uint32_t value_words(const gms_batch_sort_item& q) {
        return q.d_value_in ? (q.value_bytes == 8 ? 2u : 1u) : 0u;
    }
    // The items of every step.  Pass j of a sort with np passes reads the
    // previous pass's output (the input for j = 0) and writes the output
    // buffers when np-1-j is even, the batch's pass buffer otherwise, so that
    // the last pass writes the output and the input is never written
    // (multisplit_sort.cuh:320-559).
    gms_status step_items(std::vector<Step>& st, hipStream_t s) {
        size_t words = 0;
        for (const auto& q : sorts)
            words += ((size_t)q.n + 63) / 64 * 64 * (key_words(q) + value_words(q));
        gms_status r = tmp.reserve(std::max<size_t>(words, 1) * sizeof(uint32_t), s);
        if (r != GMS_SUCCESS) return r;
        st.clear();
        if (!items.empty() || !wides.empty()) st.push_back(Step{items, wides});
        uint32_t* base = reinterpret_cast<uint32_t*>(tmp.p);
        for (const auto& q : sorts) {
            const uint32_t kw = key_words(q), vw = value_words(q);
            SortPass pass[16];
            const int np = wide_sort_passes(kw, q.digit_bits, pass);
            if (q.n == 0) continue;
            const size_t w = ((size_t)q.n + 63) / 64 * 64;
            uint32_t* tk = base;
            uint32_t* tv = vw ? base + w * kw : nullptr;
            base += w * (kw + vw);
            const uint32_t* sk = static_cast<const uint32_t*>(q.d_key_in);
            const uint32_t* sv = static_cast<const uint32_t*>(q.d_value_in);
            for (int j = 0; j < np; ++j) {
                const bool to_out = ((np - 1 - j) % 2) == 0;
                uint32_t* dk = to_out ? static_cast<uint32_t*>(q.d_key_out) : tk;
                uint32_t* dv = vw ? (to_out ? static_cast<uint32_t*>(q.d_value_out) : tv) : nullptr;
                if ((int)st.size() <= j) st.emplace_back();
                if (kw == 1 && vw <= 1)
                    st[j].items.push_back(gms_batch_item{
                        sk, sv, dk, dv, q.n, 1u << pass[j].bits,
                        gms_bucket_desc{GMS_BUCKET_DIGIT, pass[j].shift, pass[j].bits}, nullptr});
                else
                    st[j].wides.push_back(BatchWideItem{
                        sk, sv, dk, dv, q.n, 1u << pass[j].bits, kw, vw,
                        gms_bucket_desc64{GMS_BUCKET_DIGIT, pass[j].bits, pass[j].shift}, nullptr});
                sk = dk;
                sv = dv;
            }
        }
        return GMS_SUCCESS;
    }
    gms_status replan(hipStream_t s) {
        drop_graph();
        planned = false;
        std::vector<Step> st;
        gms_status r = step_items(st, s);
        if (r != GMS_SUCCESS) return r;
        while (eng.size() < st.size()) eng.push_back(std::make_unique<BatchEngine>());
        while (eng.size() > st.size()) {
            eng.back()->release();
            eng.pop_back();
        }
        size_t bytes = 0;
        for (size_t i = 0; i < st.size() && r == GMS_SUCCESS; ++i) {
            eng[i]->tuning = tuning;
            r = eng[i]->plan(st[i].items.data(), (uint32_t)st[i].items.size(), st[i].wides.data(),
                             (uint32_t)st[i].wides.size(), 0);
            bytes += (eng[i]->device_bytes() + 255) & ~(size_t)255;
        }
        if (r == GMS_SUCCESS) r = mem.reserve(std::max<size_t>(bytes, 256), s);
        dev.assign(st.size(), BatchDevice{});
        char* p = reinterpret_cast<char*>(mem.p);
        for (size_t i = 0; i < st.size() && r == GMS_SUCCESS; ++i) {
            dev[i] = eng[i]->carve(p);
            p += (eng[i]->device_bytes() + 255) & ~(size_t)255;
            r = eng[i]->upload(dev[i], s);
        }
        if (r == GMS_SUCCESS) r = hip_status(hipStreamSynchronize(s));
        planned = r == GMS_SUCCESS;
        return r;
    }
    gms_status launch(hipStream_t s) {
        for (size_t i = 0; i < eng.size(); ++i) {
            const gms_status r = eng[i]->launch(dev[i], s);
            if (r != GMS_SUCCESS) return r;
        }
        return GMS_SUCCESS;
    }
};

static bool aligned_to(const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; }

// A 64-bit item of the C ABI checked and turned into the engine's form: a
// plain item when key and value are 32-bit (*plain set), else a wide one.
static gms_status convert_wide(const gms_batch_wide_item& it, bool* plain, gms_batch_item* pi,
                               BatchWideItem* wi) {
    if (it.num_buckets < 1 || it.num_buckets > 256) return GMS_ERR_UNSUPPORTED_BUCKETS;
    if (it.key_bytes != 4 && it.key_bytes != 8) return GMS_ERR_INVALID_ARG;
    if ((it.d_value_in == nullptr) != (it.d_value_out == nullptr)) return GMS_ERR_INVALID_ARG;
    const bool kv = it.d_value_in != nullptr;
    if (kv && it.value_bytes != 4 && it.value_bytes != 8) return GMS_ERR_INVALID_ARG;
    if (it.n && (!it.d_key_in || !it.d_key_out)) return GMS_ERR_INVALID_ARG;
    const uint32_t kb = it.key_bytes, vb = kv ? it.value_bytes : 0u;
    if (!aligned_to(it.d_key_in, kb) || !aligned_to(it.d_key_out, kb) ||
        (kv && (!aligned_to(it.d_value_in, vb) || !aligned_to(it.d_value_out, vb))))
        return GMS_ERR_INVALID_ARG;
    gms_status st = check_desc64(it.num_buckets, it.desc, kb * 8);
    if (st != GMS_SUCCESS) return st;
    const uint32_t* ki = static_cast<const uint32_t*>(it.d_key_in);
    const uint32_t* vi = static_cast<const uint32_t*>(it.d_value_in);
    uint32_t* ko = static_cast<uint32_t*>(it.d_key_out);
    uint32_t* vo = static_cast<uint32_t*>(it.d_value_out);
    *plain = kb == 4 && vb <= 4;
    if (*plain)
        *pi = gms_batch_item{ki, vi, ko, vo, it.n, it.num_buckets,
                             gms_bucket_desc{it.desc.kind, (uint32_t)it.desc.param, it.desc.width},
                             it.d_offsets};
    else
        *wi = BatchWideItem{ki, vi, ko, vo, it.n, it.num_buckets, kb / 4, vb / 4, it.desc,
                            it.d_offsets};
    return GMS_SUCCESS;
}

extern "C" {

gms_status gms_batch_create(gms_batch** out) {
    if (!out) return GMS_ERR_INVALID_ARG;
    *out = new (std::nothrow) gms_batch();
    return *out ? GMS_SUCCESS : GMS_ERR_ALLOC;
}

gms_status gms_batch_destroy(gms_batch* b) {
    if (!b) return GMS_SUCCESS;
    b->drop_graph();
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    b->release();
    delete b;
    return GMS_SUCCESS;
}

gms_status gms_batch_plan_wide(gms_batch* b, const gms_batch_item* items, uint32_t count,
                               const gms_batch_wide_item* wide, uint32_t num_wide,
                               const gms_batch_sort_item* sorts, uint32_t num_sorts,
                               void* stream) {
    if (!b || (count && !items) || (num_wide && !wide) || (num_sorts && !sorts))
        return GMS_ERR_INVALID_ARG;
    std::vector<gms_batch_item> plain(items, items + count);
    std::vector<BatchWideItem> wides;
    for (uint32_t i = 0; i < count; ++i) {
        gms_status st = check_batch_item(items[i]);
        if (st != GMS_SUCCESS) return st;
    }
    for (uint32_t i = 0; i < num_wide; ++i) {
        bool is_plain = false;
        gms_batch_item pi{};
        BatchWideItem wi{};
        gms_status st = convert_wide(wide[i], &is_plain, &pi, &wi);
        if (st != GMS_SUCCESS) return st;
        if (is_plain) plain.push_back(pi);
        else wides.push_back(wi);
    }
    // Every byte range of the batch: the multisplits' and the sorts'.
    ByteRanges R;
    for (const auto& it : plain) add_item_ranges(R, it);
    for (const auto& it : wides) {
        if (it.n == 0) continue;
        R.add(it.kin, (size_t)it.n * it.kw * 4, false);
        R.add(it.kout, (size_t)it.n * it.kw * 4, true);
        R.add(it.vin, (size_t)it.n * it.vw * 4, false);
        R.add(it.vout, (size_t)it.n * it.vw * 4, true);
        R.add(it.offs, (size_t)it.num_buckets * 4, true);
    }
    for (uint32_t i = 0; i < num_sorts; ++i) {
        const gms_batch_sort_item& q = sorts[i];
        if ((q.key_bytes != 0 && q.key_bytes != 4 && q.key_bytes != 8) ||
            (q.value_bytes != 0 && q.value_bytes != 4 && q.value_bytes != 8))
            return GMS_ERR_INVALID_ARG;
        const uint32_t kb = q.key_bytes == 8 ? 8u : 4u, vb = q.value_bytes == 8 ? 8u : 4u;
        SortPass pass[16];
        if (wide_sort_passes(kb / 4, q.digit_bits, pass) == 0) return GMS_ERR_INVALID_ARG;
        if ((q.d_value_in == nullptr) != (q.d_value_out == nullptr)) return GMS_ERR_INVALID_ARG;
        if (q.n && (!q.d_key_in || !q.d_key_out)) return GMS_ERR_INVALID_ARG;
        if (!aligned_to(q.d_key_in, kb) || !aligned_to(q.d_key_out, kb) ||
            (q.d_value_in && (!aligned_to(q.d_value_in, vb) || !aligned_to(q.d_value_out, vb))))
            return GMS_ERR_INVALID_ARG;
        if (q.n == 0) continue;
        R.add(q.d_key_in, (size_t)q.n * kb, false);
        R.add(q.d_key_out, (size_t)q.n * kb, true);
        R.add(q.d_value_in, (size_t)q.n * vb, false);
        R.add(q.d_value_out, (size_t)q.n * vb, true);
    }
    gms_status st = R.check();
    if (st != GMS_SUCCESS) return st;
    b->items = std::move(plain);
    b->wides = std::move(wides);
    b->sorts.assign(sorts, sorts + num_sorts);
    return b->replan((hipStream_t)stream);
}

gms_status gms_batch_plan_mixed(gms_batch* b, const gms_batch_item* items, uint32_t count,
                                const gms_batch_sort_item* sorts, uint32_t num_sorts,
                                void* stream) {
    return gms_batch_plan_wide(b, items, count, nullptr, 0, sorts, num_sorts, stream);
}

gms_status gms_batch_plan(gms_batch* b, const gms_batch_item* items, uint32_t count,
                          void* stream) {
    return gms_batch_plan_wide(b, items, count, nullptr, 0, nullptr, 0, stream);
}

gms_status gms_batch_run(gms_batch* b, void* stream) {
    if (!b) return GMS_ERR_INVALID_ARG;
    if (!b->planned) return GMS_ERR_NOT_ALLOCATED;
    return b->launch((hipStream_t)stream);
}

gms_status gms_batch_capture(gms_batch* b) {
    if (!b) return GMS_ERR_INVALID_ARG;
    if (!b->planned) return GMS_ERR_NOT_ALLOCATED;
    b->drop_graph();
    if (!b->cap_stream && hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess) {
        b->cap_stream = nullptr;
        return GMS_ERR_HIP;
    }
    if (hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return GMS_ERR_HIP;
    gms_status st = b->launch(b->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (st != GMS_SUCCESS || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return st != GMS_SUCCESS ? st : GMS_ERR_HIP;
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        b->exec = nullptr;
        b->drop_graph();
        return GMS_ERR_HIP;
    }
    return GMS_SUCCESS;
}

gms_status gms_batch_launch_graph(gms_batch* b, void* stream) {
    if (!b) return GMS_ERR_INVALID_ARG;
    if (!b->exec) return GMS_ERR_NOT_ALLOCATED;
    return hip_status(hipGraphLaunch(b->exec, (hipStream_t)stream));
}

gms_status gms_batch_set_tuning(gms_batch* b, gms_batch_tuning t) {
    if (!b) return GMS_ERR_INVALID_ARG;
    if (t.items_per_thread != 8 && t.items_per_thread != 16 && t.items_per_thread != 24 &&
        t.items_per_thread != 32)
        return GMS_ERR_INVALID_ARG;
    if (t.persistent > 1) return GMS_ERR_INVALID_ARG;
    b->tuning.items_per_thread = t.items_per_thread;
    b->tuning.persistent = t.persistent;
    return b->planned ? b->replan(nullptr) : GMS_SUCCESS;
}

gms_status gms_batch_get_tuning(const gms_batch* b, gms_batch_tuning* t) {
    if (!b || !t) return GMS_ERR_INVALID_ARG;
    t->items_per_thread = b->tuning.items_per_thread;
    t->persistent = b->tuning.persistent;
    return GMS_SUCCESS;
}

gms_status gms_batch_autotune(gms_batch* b, uint32_t reps, void* stream, gms_batch_tuning* chosen,
                              float* ms_per_run) {
    if (!b || reps == 0) return GMS_ERR_INVALID_ARG;
    if (!b->planned) return GMS_ERR_NOT_ALLOCATED;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return GMS_ERR_HIP;
    }
    const BatchTuning keep = b->tuning;
    BatchTuning best = keep;
    float best_ms = -1.f;
    gms_status st = GMS_SUCCESS;
    for (uint32_t K : {8u, 16u, 24u, 32u}) {
        for (uint32_t pers : {0u, 1u}) {
            b->tuning.items_per_thread = K;
            b->tuning.persistent = pers;
            st = b->replan(s);
            if (st == GMS_SUCCESS) st = b->launch(s);  // warm-up
            if (st == GMS_SUCCESS) st = hip_status(hipEventRecord(e0, s));
            for (uint32_t r = 0; r < reps && st == GMS_SUCCESS; ++r) st = b->launch(s);
            if (st == GMS_SUCCESS) st = hip_status(hipEventRecord(e1, s));
            if (st == GMS_SUCCESS) st = hip_status(hipEventSynchronize(e1));
            float ms = 0.f;
            if (st == GMS_SUCCESS) st = hip_status(hipEventElapsedTime(&ms, e0, e1));
            if (st != GMS_SUCCESS) break;
            ms /= (float)reps;
            if (best_ms < 0.f || ms < best_ms) {
                best_ms = ms;
                best = b->tuning;
            }
        }
        if (st != GMS_SUCCESS) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tuning = st == GMS_SUCCESS ? best : keep;
    const gms_status st2 = b->replan(s);
    if (st != GMS_SUCCESS) return st;
    if (st2 != GMS_SUCCESS) return st2;
    if (chosen) {
        chosen->items_per_thread = best.items_per_thread;
        chosen->persistent = best.persistent;
    }
    if (ms_per_run) *ms_per_run = best_ms;
    return GMS_SUCCESS;
}

gms_status gms_batch_stats(const gms_batch* b, uint64_t* keys, uint32_t* work_items,
                           uint64_t* matrix_words, uint32_t* launches) {
    if (!b) return GMS_ERR_INVALID_ARG;
    uint64_t k = 0, m = 0;
    uint32_t w = 0, l = 0;
    for (const auto& e : b->eng) {
        k += e->keys_total;
        m += e->m_total;
        w += e->num_work;
        l += e->launches();
    }
    if (keys) *keys = k;
    if (work_items) *work_items = w;
    if (matrix_words) *matrix_words = m;
    if (launches) *launches = l;
    return GMS_SUCCESS;
}

}  // extern "C"
