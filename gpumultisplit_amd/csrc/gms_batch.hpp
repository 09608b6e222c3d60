// Host engine of the batched multisplit: plans a list of instances into
// device tables (instances, (instance, tile) work items, the concatenated
// histogram matrix) and launches the batched kernels on a stream.  Used by
// the gms_batch_* entry points (plan once, replay through a hipGraph) and by
// the request server (a new plan every step, tables staged through pinned
// host memory so that planning overlaps the previous step on the GPU).
#pragma once

#include <stdint.h>

#include <vector>

#include "gms_batch_kernels.hpp"
#include "gms_internal.hpp"

namespace gms {

// Device buffer that only grows (the stream is synchronised before a grow).
template <class T>
struct DevVec {
    T* p = nullptr;
    size_t cap = 0;
    gms_status reserve(size_t n, hipStream_t s);
    void release();
};

// Pinned host buffer that only grows.
template <class T>
struct PinnedVec {
    T* p = nullptr;
    size_t cap = 0;
    size_t size = 0;
    gms_status reserve(size_t n);
    void release();
};

struct BatchGroup {
    int log_b;
    bool kv;
    bool vec;         // every instance 16-byte aligned
    uint32_t w0, w1;  // work items [w0, w1)
};

// One multisplit of elements with a 64-bit key and/or value: kw key words and
// vw value words per element (not both at most 1), arrays of 32-bit words.
struct BatchWideItem {
    const uint32_t* kin;
    const uint32_t* vin;
    uint32_t* kout;
    uint32_t* vout;
    uint32_t n;
    uint32_t num_buckets;  // 1..256
    uint32_t kw, vw;
    gms_bucket_desc64 desc;
    uint32_t* offs;
};

struct BatchWideGroup {
    int log_b;
    uint32_t kw, vw;
    bool vec;
    uint32_t w0, w1;  // work items [w0, w1), after every 32-bit group's
};

struct BatchTuning {
    uint32_t items_per_thread = 24;  // 8, 16, 24 or 32 keys per thread (tile = 256 x this)
    uint32_t persistent = 0;         // 0: one workgroup per tile; 1: resident grid + work queue
};

// Device tables of one planned batch, carved from one allocation.
struct BatchDevice {
    BatchInst* inst = nullptr;
    BatchWideInst* winst = nullptr;
    BatchWork* work = nullptr;
    uint32_t* mat = nullptr;    // concatenated histogram matrix
    uint32_t* part = nullptr;   // scan partials
    uint32_t* queue = nullptr;  // 8 per-XCD work-queue counters per launch
};

class BatchEngine {
  public:
    BatchTuning tuning;
    // Plan `count` instances of 32-bit elements and `wcount` of 64-bit ones
    // (validated).  Host tables go to staging slot `slot` (0 or 1) of pinned
    // memory: the caller makes sure no copy from that slot is still pending.
    gms_status plan(const gms_batch_item* items, uint32_t count, const BatchWideItem* wide,
                    uint32_t wcount, int slot);
    gms_status plan(const gms_batch_item* items, uint32_t count, int slot) {
        return plan(items, count, nullptr, 0, slot);
    }
    // Device bytes the current plan needs, and its tables inside `base`
    // (256-byte aligned, device_bytes() long).
    size_t device_bytes() const;
    BatchDevice carve(void* base) const;
    // Copies the staged tables to the device on `stream` (asynchronous).
    gms_status upload(const BatchDevice& d, hipStream_t stream);
    // Queue reset, prescans, one scan, postscans -- all kernels, no memsets.
    gms_status launch(const BatchDevice& d, hipStream_t stream);
    // Grows both staging slots to hold `instances` instances and `work_items`
    // work items, so that plans up to that size allocate no pinned memory.
    gms_status reserve_tables(uint32_t instances, uint64_t work_items);
    void release();

    uint64_t keys_total = 0;
    uint64_t m_total = 0;
    uint32_t num_work = 0;
    uint32_t num_instances = 0;
    uint32_t num_wide = 0;
    std::vector<BatchGroup> groups;
    std::vector<BatchWideGroup> wide_groups;
    uint32_t launches() const;

  private:
    int slot_ = 0;
    PinnedVec<BatchInst> h_inst_[2];
    PinnedVec<BatchWideInst> h_winst_[2];
    PinnedVec<BatchWork> h_work_[2];
    int num_cus_ = 0;
    uint32_t grid_for(const BatchGroup& g, bool post) const;
};

// Checks of one item (bucket count, descriptor, pointers, in/out overlap).
gms_status check_batch_item(const gms_batch_item& it);
// Checks that no output range overlaps any other range of the batch.
gms_status check_batch_overlaps(const gms_batch_item* items, uint32_t count);

// Byte ranges of a batch's buffers; an output may overlap no other range.
struct ByteRanges {
    struct R {
        uintptr_t a, b;
        bool out;
    };
    std::vector<R> r;
    void add(const void* p, size_t bytes, bool out) {
        if (p && bytes) r.push_back({(uintptr_t)p, (uintptr_t)p + bytes, out});
    }
    gms_status check();
};

}  // namespace gms
