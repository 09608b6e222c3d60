// Request server with continuous batching (gms_server_* entry points).
//
// A scheduler thread owns one HIP stream and a device scratch arena of the
// configured budget.  Requests (a multisplit = one pass, a radix sort = one
// pass per digit) are admitted in arrival order while their reservation fits
// the budget.  Every step gathers the next pass of every admitted request
// into ONE batched multisplit (gms_batch.hpp) and queues it on the stream;
// up to max_inflight_steps steps are queued at once, so the host plans step
// k+1 while the GPU runs step k, and a request that arrives meanwhile joins
// the next step.  A sort's consecutive passes land in consecutive steps, so
// stream order carries each pass's output to the next.  The arithmetic of a
// request is the same as in the plain path and does not depend on what else
// shares its step.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gms_batch.hpp"

namespace gms {

using Clock = std::chrono::steady_clock;

static size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

// First-fit sub-allocator over one device allocation (256-byte granules,
// neighbouring free ranges coalesced).
class Arena {
  public:
    char* base = nullptr;
    size_t size = 0, used = 0, peak = 0;

    gms_status init(size_t bytes) {
        size = round256(bytes);
        if (hipMalloc((void**)&base, size) != hipSuccess) {
            base = nullptr;
            size = 0;
            return GMS_ERR_ALLOC;
        }
        free_.clear();
        free_[0] = size;
        return GMS_SUCCESS;
    }
    void destroy() {
        if (base) (void)hipFree(base);
        base = nullptr;
        size = used = 0;
        free_.clear();
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = round256(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t o = it->first, rem = it->second - bytes;
            free_.erase(it);
            if (rem) free_[o + bytes] = rem;
            *off = o;
            used += bytes;
            peak = used > peak ? used : peak;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = round256(bytes ? bytes : 1);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<size_t, size_t> free_;
};

struct Req {
    uint64_t ticket = 0;
    gms_request r{};
    uint32_t kw = 1, vw = 0;  // key and value words per element
    SortPass pass[16];
    int np = 1;        // passes
    int next = 0;      // next pass to queue
    int retired = 0;   // passes whose step has completed
    size_t tmp_off = 0, tmp_bytes = 0;
    size_t reserve = 0;  // budget held while admitted
    uint32_t* ktmp = nullptr;
    uint32_t* vtmp = nullptr;
    Clock::time_point t_submit, t_done;
    bool done = false;
    gms_status status = GMS_SUCCESS;

    bool wide() const { return kw == 2 || vw == 2; }
    const uint32_t* kin() const { return static_cast<const uint32_t*>(r.d_key_in); }
    const uint32_t* vin() const { return static_cast<const uint32_t*>(r.d_value_in); }
    uint32_t* kout() const { return static_cast<uint32_t*>(r.d_key_out); }
    uint32_t* vout() const { return static_cast<uint32_t*>(r.d_value_out); }

    // Source and destination words of pass p: the last pass writes the
    // output, the others alternate with the request's pass buffer, and the
    // input is never written.
    void pass_buffers(int p, const uint32_t** sk, const uint32_t** sv, uint32_t** dk,
                      uint32_t** dv) const {
        auto dst_k = [&](int q) { return ((np - 1 - q) % 2 == 0) ? kout() : ktmp; };
        auto dst_v = [&](int q) { return ((np - 1 - q) % 2 == 0) ? vout() : vtmp; };
        const bool kv = vw != 0;
        *sk = p == 0 ? kin() : dst_k(p - 1);
        *sv = kv ? (p == 0 ? vin() : dst_v(p - 1)) : nullptr;
        *dk = dst_k(p);
        *dv = kv ? dst_v(p) : nullptr;
    }
    // The batched multisplit item of pass p (32-bit elements).
    gms_batch_item item(int p) const {
        if (r.op == GMS_OP_MULTISPLIT)
            return gms_batch_item{kin(), vin(), kout(), vout(), r.n, r.num_buckets, r.desc,
                                  r.d_offsets};
        const uint32_t *sk, *sv;
        uint32_t *dk, *dv;
        pass_buffers(p, &sk, &sv, &dk, &dv);
        return gms_batch_item{sk, sv, dk, dv, r.n, 1u << pass[p].bits,
                              gms_bucket_desc{GMS_BUCKET_DIGIT, pass[p].shift, pass[p].bits},
                              nullptr};
    }
    // The same for elements with a 64-bit key or value.
    BatchWideItem wide_item(int p) const {
        if (r.op == GMS_OP_MULTISPLIT)
            return BatchWideItem{kin(), vin(), kout(), vout(), r.n, r.num_buckets, kw, vw, r.desc64,
                                 r.d_offsets};
        const uint32_t *sk, *sv;
        uint32_t *dk, *dv;
        pass_buffers(p, &sk, &sv, &dk, &dv);
        return BatchWideItem{sk, sv, dk, dv, r.n, 1u << pass[p].bits, kw, vw,
                             gms_bucket_desc64{GMS_BUCKET_DIGIT, pass[p].bits, pass[p].shift},
                             nullptr};
    }
};

// Device bytes one pass of n elements with 2^log_b matrix rows can need in a
// step (its share of the step's tables; shares add up to at least the step).
static size_t step_share(uint32_t n, int log_b, uint32_t tile_elems, size_t inst_bytes) {
    const uint64_t tile = tile_elems;
    const uint64_t tiles = n ? (n + tile - 1) / tile : 1;
    const uint64_t m = ((uint64_t)1 << log_b) * tiles;
    return round256(inst_bytes) + round256(tiles * sizeof(BatchWork)) + round256(m * 4) +
           round256(((m + kScanChunk - 1) / kScanChunk + 1) * 4) + 256;
}

}  // namespace gms

using namespace gms;

struct gms_server {
    gms_server_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    Arena arena;
    BatchEngine eng;

    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<std::shared_ptr<Req>> pending;
    std::vector<std::shared_ptr<Req>> active;
    std::unordered_map<uint64_t, std::shared_ptr<Req>> by_ticket;
    struct Step {
        std::vector<std::shared_ptr<Req>> reqs;
        size_t off = 0, bytes = 0;
        hipEvent_t ev = nullptr;
    };
    std::deque<Step> inflight;
    std::vector<hipEvent_t> events;  // free events
    size_t reserved = 0;             // budget held by admitted requests
    uint64_t next_ticket = 1;
    bool stopping = false;
    gms_status error = GMS_SUCCESS;  // sticky scheduler failure
    int slot = 0;
    gms_server_stats st{};
    size_t peak = 0;  // arena high-water mark, published under the lock
    std::thread th;

    size_t budget() const { return arena.size - arena.size / 16; }  // slack for fragmentation

    void finish(const std::shared_ptr<Req>& r, gms_status s) {
        r->done = true;
        r->status = s;
        r->t_done = Clock::now();
        if (r->tmp_bytes) arena.release(r->tmp_off, r->tmp_bytes);
        r->tmp_bytes = 0;
        reserved -= r->reserve;
        r->reserve = 0;
        for (size_t i = 0; i < active.size(); ++i)
            if (active[i] == r) {
                active.erase(active.begin() + (ptrdiff_t)i);
                break;
            }
        ++st.completed;
    }

    void fail_all(gms_status s) {
        error = s;
        for (auto& stp : inflight) {
            arena.release(stp.off, stp.bytes);
            events.push_back(stp.ev);
        }
        inflight.clear();
        while (!active.empty()) finish(active.back(), s);
        while (!pending.empty()) {
            auto r = pending.front();
            pending.pop_front();
            r->done = true;
            r->status = s;
            r->t_done = Clock::now();
            ++st.completed;
        }
        cv_done.notify_all();
    }

    // Oldest step done (or, when `block`, wait for it): retire it.
    bool retire_one(std::unique_lock<std::mutex>& lk, bool block) {
        Step& s = inflight.front();
        hipError_t q = hipEventQuery(s.ev);
        if (q == hipErrorNotReady) {
            if (!block) return false;
            hipEvent_t ev = s.ev;
            lk.unlock();
            q = hipEventSynchronize(ev);
            lk.lock();
        }
        if (q != hipSuccess) {
            fail_all(GMS_ERR_HIP);
            return true;
        }
        Step done = std::move(inflight.front());
        inflight.pop_front();
        arena.release(done.off, done.bytes);
        events.push_back(done.ev);
        for (auto& r : done.reqs)
            if (++r->retired == r->np) finish(r, GMS_SUCCESS);
        cv_done.notify_all();
        return true;
    }

    bool admit(const std::shared_ptr<Req>& r) {
        if (reserved + r->reserve > budget()) return false;
        if (r->tmp_bytes) {
            if (!arena.alloc(r->tmp_bytes, &r->tmp_off)) return false;
            r->ktmp = (uint32_t*)(arena.base + r->tmp_off);
            if (r->vw) r->vtmp = r->ktmp + (round256((size_t)r->r.n * 4 * r->kw) / 4);
        }
        reserved += r->reserve;
        active.push_back(r);
        peak = arena.peak;
        return true;
    }

    // Plans and queues one step; false when there is nothing to queue now.
    // Planning, the table upload and the launches run with the lock released
    // (they touch only scheduler-owned state), so submitters are not held up.
    bool launch_step(std::unique_lock<std::mutex>& lk) {
        std::vector<gms_batch_item> items;
        std::vector<BatchWideItem> wides;
        std::vector<std::shared_ptr<Req>> reqs;
        uint64_t keys = 0;
        const uint64_t cap = cfg.max_step_keys ? cfg.max_step_keys : (1ull << 31);
        for (auto& r : active) {
            if (r->next >= r->np) continue;
            if (!reqs.empty() && keys + r->r.n > cap) break;
            reqs.push_back(r);
            keys += r->r.n;
        }
        if (reqs.empty()) return false;
        // the step's items of the first `count` requests
        auto gather = [&](size_t count) {
            items.clear();
            wides.clear();
            for (size_t i = 0; i < count; ++i) {
                if (reqs[i]->wide()) wides.push_back(reqs[i]->wide_item(reqs[i]->next));
                else items.push_back(reqs[i]->item(reqs[i]->next));
            }
        };
        gather(reqs.size());
        lk.unlock();
        gms_status s = GMS_SUCCESS;
        size_t off = 0, bytes = 0;
        bool placed = false;
        hipEvent_t ev = nullptr;
        while (!reqs.empty()) {
            s = eng.plan(items.data(), (uint32_t)items.size(), wides.data(), (uint32_t)wides.size(),
                         slot);
            if (s != GMS_SUCCESS) break;
            bytes = eng.device_bytes();
            if (arena.alloc(bytes, &off)) {
                placed = true;
                break;
            }
            if (!inflight.empty()) break;  // retry once a step has retired
            // Nothing in flight: the arena holds only admitted sorts'
            // buffers, which can leave the free space in pieces too small for
            // the step's tables even though the reservations fit the budget.
            // Try fewer items; a single item that still does not fit fails
            // its request below instead of waiting for space that no
            // retirement will free.
            if (reqs.size() == 1) break;
            reqs.resize(reqs.size() / 2);
            gather(reqs.size());
        }
        keys = 0;
        for (const auto& r : reqs) keys += r->r.n;
        if (s == GMS_SUCCESS && placed) {
            if (!events.empty()) {
                ev = events.back();
                events.pop_back();
            } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
                ev = nullptr;
                s = GMS_ERR_HIP;
            }
            const BatchDevice d = eng.carve(arena.base + off);
            if (s == GMS_SUCCESS) s = eng.upload(d, stream);
            if (s == GMS_SUCCESS) s = eng.launch(d, stream);
            if (s == GMS_SUCCESS) s = hip_status(hipEventRecord(ev, stream));
        }
        lk.lock();
        if (s != GMS_SUCCESS) {
            if (placed) arena.release(off, bytes);
            if (ev) events.push_back(ev);
            fail_all(s);
            return false;
        }
        if (!placed) {
            if (inflight.empty() && reqs.size() == 1) {
                finish(reqs[0], GMS_ERR_CAPACITY);
                cv_done.notify_all();
                return true;  // progress: its buffers are free again
            }
            return false;
        }
        for (auto& r : reqs) ++r->next;
        st.step_items += reqs.size();
        inflight.push_back(Step{std::move(reqs), off, bytes, ev});
        slot ^= 1;
        ++st.steps;
        st.keys_processed += keys;
        peak = arena.peak;
        return true;
    }

    void run() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            // 1. retire completed steps; block on the oldest when the queue is full
            while (!inflight.empty() &&
                   retire_one(lk, inflight.size() >= cfg.max_inflight_steps)) {
            }
            // 2. admit in arrival order
            while (error == GMS_SUCCESS && !pending.empty() && admit(pending.front()))
                pending.pop_front();
            // 3. queue a step if there is room on the GPU
            if (error == GMS_SUCCESS && inflight.size() < cfg.max_inflight_steps && launch_step(lk))
                continue;
            if (stopping && pending.empty() && active.empty() && inflight.empty()) break;
            if (!inflight.empty())
                cv_sched.wait_for(lk, std::chrono::microseconds(20));
            else if (!active.empty())
                cv_sched.wait_for(lk, std::chrono::milliseconds(1));  // never parked with work left
            else
                cv_sched.wait(lk);
        }
    }
};

extern "C" {

gms_status gms_server_create(const gms_server_config* cfg, gms_server** out) {
    if (!cfg || !out) return GMS_ERR_INVALID_ARG;
    *out = nullptr;
    if (cfg->scratch_budget_bytes < (1u << 20)) return GMS_ERR_INVALID_ARG;
    if (cfg->max_inflight_steps < 1 || cfg->max_inflight_steps > 2) return GMS_ERR_INVALID_ARG;
    const uint32_t K = cfg->items_per_thread ? cfg->items_per_thread : 24;
    if (K != 8 && K != 16 && K != 24 && K != 32) return GMS_ERR_INVALID_ARG;
    gms_server* s = new (std::nothrow) gms_server();
    if (!s) return GMS_ERR_ALLOC;
    s->cfg = *cfg;
    s->cfg.items_per_thread = K;
    s->eng.tuning.items_per_thread = K;
    s->eng.tuning.persistent = 0;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return GMS_ERR_NO_DEVICE;
    }
    gms_status st = s->arena.init(cfg->scratch_budget_bytes);
    // Staging tables for steps of up to min(max_step_keys, 2^30) keys in 4096
    // requests, allocated here rather than on the first large step (pinned
    // allocations take milliseconds); larger steps still grow them.
    if (st == GMS_SUCCESS) {
        const uint64_t keys = std::min<uint64_t>(cfg->max_step_keys ? cfg->max_step_keys : (1ull << 31),
                                                 1ull << 30);
        st = s->eng.reserve_tables(4096, keys / ((uint64_t)kThreads * K) + 4096);
    }
    if (st != GMS_SUCCESS) {
        s->eng.release();
        s->arena.destroy();
        delete s;
        return st;
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        s->arena.destroy();
        delete s;
        return GMS_ERR_HIP;
    }
    s->th = std::thread([s]() { s->run(); });
    *out = s;
    return GMS_SUCCESS;
}

gms_status gms_server_destroy(gms_server* s) {
    if (!s) return GMS_SUCCESS;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_sched.notify_all();
    if (s->th.joinable()) s->th.join();
    (void)hipStreamSynchronize(s->stream);
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(s->stream);
    s->eng.release();
    s->arena.destroy();
    delete s;
    return GMS_SUCCESS;
}

gms_status gms_server_submit(gms_server* s, const gms_request* q, uint64_t* ticket) {
    if (!s || !q || !ticket) return GMS_ERR_INVALID_ARG;
    auto r = std::make_shared<Req>();
    r->r = *q;
    const bool kv = q->d_value_in != nullptr;
    if (kv != (q->d_value_out != nullptr)) return GMS_ERR_INVALID_ARG;
    if (q->n && (!q->d_key_in || !q->d_key_out)) return GMS_ERR_INVALID_ARG;
    if ((q->key_bytes != 0 && q->key_bytes != 4 && q->key_bytes != 8) ||
        (q->value_bytes != 0 && q->value_bytes != 4 && q->value_bytes != 8))
        return GMS_ERR_INVALID_ARG;
    r->kw = q->key_bytes == 8 ? 2u : 1u;
    r->vw = kv ? (q->value_bytes == 8 ? 2u : 1u) : 0u;
    const size_t kb = (size_t)q->n * 4 * r->kw, vb = (size_t)q->n * 4 * r->vw;
    auto al = [](const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; };
    if (!al(q->d_key_in, 4 * r->kw) || !al(q->d_key_out, 4 * r->kw) ||
        (kv && (!al(q->d_value_in, 4 * r->vw) || !al(q->d_value_out, 4 * r->vw))))
        return GMS_ERR_INVALID_ARG;
    if (ranges_overlap(q->d_key_in, kb, q->d_key_out, kb) ||
        (kv && (ranges_overlap(q->d_value_in, vb, q->d_value_out, vb) ||
                ranges_overlap(q->d_key_in, kb, q->d_value_out, vb) ||
                ranges_overlap(q->d_value_in, vb, q->d_key_out, kb) ||
                ranges_overlap(q->d_key_out, kb, q->d_value_out, vb))))
        return GMS_ERR_ALIAS;
    const uint32_t K = s->cfg.items_per_thread;
    // tables of one pass: the 32-bit batched tile, or the 64-bit one
    const uint32_t tile = r->wide() ? wide_tile_elems(r->kw, r->vw) : kThreads * K;
    const size_t inst = r->wide() ? sizeof(BatchWideInst) : sizeof(BatchInst);
    if (q->op == GMS_OP_MULTISPLIT) {
        gms_status st;
        if (r->wide()) {
            if (q->num_buckets < 1 || q->num_buckets > 256) return GMS_ERR_UNSUPPORTED_BUCKETS;
            st = check_desc64(q->num_buckets, q->desc64, 32 * r->kw);
        } else {
            st = check_batch_item(r->item(0));
        }
        if (st != GMS_SUCCESS) return st;
        r->np = 1;
        const int lb = ceil_log2_u32(q->num_buckets);
        r->reserve = 2 * step_share(q->n, lb < 1 ? 1 : lb, tile, inst);
    } else if (q->op == GMS_OP_SORT) {
        r->np = wide_sort_passes(r->kw, q->digit_bits, r->pass);
        if (r->np == 0) return GMS_ERR_INVALID_ARG;
        r->tmp_bytes = round256(kb) + (kv ? round256(vb) : 0);
        r->reserve = round256(r->tmp_bytes) + 2 * step_share(q->n, 8, tile, inst);
    } else {
        return GMS_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->error != GMS_SUCCESS) return s->error;
    if (s->stopping) return GMS_ERR_INVALID_ARG;
    if (r->reserve > s->budget()) {
        ++s->st.rejected;
        return GMS_ERR_CAPACITY;
    }
    r->ticket = s->next_ticket++;
    r->t_submit = Clock::now();
    s->by_ticket[r->ticket] = r;
    s->pending.push_back(r);
    ++s->st.submitted;
    *ticket = r->ticket;
    s->cv_sched.notify_one();
    return GMS_SUCCESS;
}

gms_status gms_server_poll(gms_server* s, uint64_t ticket, int* done, double* latency_us) {
    if (!s || !done) return GMS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->by_ticket.find(ticket);
    if (it == s->by_ticket.end()) return GMS_ERR_INVALID_ARG;
    const Req& r = *it->second;
    *done = r.done ? 1 : 0;
    if (latency_us)
        *latency_us = r.done ? std::chrono::duration<double, std::micro>(r.t_done - r.t_submit).count()
                             : 0.0;
    return r.done ? r.status : GMS_SUCCESS;
}

gms_status gms_server_wait(gms_server* s, uint64_t ticket, double* latency_us) {
    if (!s) return GMS_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->by_ticket.find(ticket);
    if (it == s->by_ticket.end()) return GMS_ERR_INVALID_ARG;
    std::shared_ptr<Req> r = it->second;
    s->cv_done.wait(lk, [&]() { return r->done; });
    if (latency_us)
        *latency_us = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
    s->by_ticket.erase(ticket);
    return r->status;
}

#ifdef GMS_HOST_EMULATION
// Test hooks of the CPU emulation build only (not in the GPU library):
// occupy or free a range of the server's arena, to test how steps are placed
// in a fragmented arena.  The caller makes sure the server is idle.
gms_status gms_emu_server_hold(gms_server* s, size_t bytes, size_t* off) {
    std::lock_guard<std::mutex> lk(s->mu);
    return s->arena.alloc(bytes, off) ? GMS_SUCCESS : GMS_ERR_ALLOC;
}
gms_status gms_emu_server_unhold(gms_server* s, size_t off, size_t bytes) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->arena.release(off, bytes);
    return GMS_SUCCESS;
}
#endif

gms_status gms_server_get_stats(gms_server* s, gms_server_stats* out) {
    if (!s || !out) return GMS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->peak_scratch_bytes = s->peak;
    out->active = (uint32_t)s->active.size();
    out->pending = (uint32_t)s->pending.size();
    return GMS_SUCCESS;
}

}  // extern "C"
