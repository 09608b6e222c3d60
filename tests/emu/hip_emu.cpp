// Workgroup execution for the host emulation (see hip_emu.hpp).
#include "hip_emu.hpp"

#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace gms_emu {

thread_local gms_emu_dim3 thread_idx = {0, 0, 0};
gms_emu_dim3 block_idx = {0, 0, 0}, block_dim = {1, 1, 1}, grid_dim = {1, 1, 1};

namespace {
std::unique_ptr<std::barrier<>> g_block_barrier;
std::vector<std::unique_ptr<std::barrier<>>> g_wave_barrier;
uint32_t g_wave_buf[1024];
std::mutex g_launch_mutex;
}  // namespace

void syncthreads() { g_block_barrier->arrive_and_wait(); }

static std::barrier<>& wave_barrier() { return *g_wave_barrier[thread_idx.x >> 6]; }

uint32_t update_dpp(uint32_t old, uint32_t src, int ctrl, int row_mask, int bank_mask,
                    bool bound_ctrl) {
    const uint32_t lane = thread_idx.x & 63u, wave_base = thread_idx.x & ~63u;
    uint32_t* buf = g_wave_buf + wave_base;
    buf[lane] = src;
    wave_barrier().arrive_and_wait();
    const int row = (int)(lane >> 4), bank = (int)((lane >> 2) & 3u);
    uint32_t r;
    if (!((row_mask >> row) & 1) || !((bank_mask >> bank) & 1)) {
        r = old;  // lane not written
    } else {
        int s = -1;
        if (ctrl >= 0x111 && ctrl <= 0x11F) {  // row_shr:n
            const int sh = ctrl - 0x110;
            if ((int)(lane & 15u) >= sh) s = (int)lane - sh;
        } else if (ctrl == 0x142) {  // row_bcast:15
            if (row >= 1) s = row * 16 - 1;
        } else if (ctrl == 0x143) {  // row_bcast:31
            if (row >= 2) s = 31;
        } else {
            fprintf(stderr, "hip_emu: unsupported dpp control 0x%x\n", ctrl);
            abort();
        }
        r = s >= 0 ? buf[s] : (bound_ctrl ? 0u : old);
    }
    wave_barrier().arrive_and_wait();
    return r;
}

void wave_sync() { wave_barrier().arrive_and_wait(); }

int readlane(int x, int lane) {
    const uint32_t wave_base = thread_idx.x & ~63u;
    uint32_t* buf = g_wave_buf + wave_base;
    buf[thread_idx.x & 63u] = (uint32_t)x;
    wave_barrier().arrive_and_wait();
    const int r = (int)buf[lane & 63];
    wave_barrier().arrive_and_wait();
    return r;
}

uint64_t ballot(bool pred) {
    const uint32_t wave_base = thread_idx.x & ~63u;
    uint32_t* buf = g_wave_buf + wave_base;
    buf[thread_idx.x & 63u] = pred ? 1u : 0u;
    wave_barrier().arrive_and_wait();
    uint64_t m = 0;
    const uint32_t lanes = block_dim.x - wave_base < 64u ? block_dim.x - wave_base : 64u;
    for (uint32_t l = 0; l < lanes; ++l) m |= (uint64_t)(buf[l] & 1u) << l;
    wave_barrier().arrive_and_wait();
    return m;
}

void launch(dim3 grid, dim3 block, std::function<void()> body) {
    std::lock_guard<std::mutex> lk(g_launch_mutex);
    const uint32_t nt = block.x;
    if (nt == 0 || nt > 1024 || (nt % 64 != 0 && nt > 64)) {
        fprintf(stderr, "hip_emu: unsupported block size %u\n", nt);
        abort();
    }
    grid_dim = {grid.x, 1, 1};
    block_dim = {nt, 1, 1};
    g_block_barrier = std::make_unique<std::barrier<>>((std::ptrdiff_t)nt);
    g_wave_barrier.clear();
    for (uint32_t w = 0; w < (nt + 63) / 64; ++w)
        g_wave_barrier.push_back(std::make_unique<std::barrier<>>((std::ptrdiff_t)(nt < 64 ? nt : 64)));
    for (uint32_t b = 0; b < grid.x; ++b) {
        block_idx = {b, 0, 0};
        std::vector<std::thread> th;
        th.reserve(nt);
        for (uint32_t t = 0; t < nt; ++t)
            th.emplace_back([t, &body]() {
                thread_idx = {t, 0, 0};
                body();
            });
        for (auto& x : th) x.join();
    }
}

std::atomic<uint64_t> alloc_calls{0};
}  // namespace gms_emu

// Allocation calls (hipMalloc, hipFree, hipHostMalloc, hipHostFree) so far.
extern "C" uint64_t gms_emu_alloc_calls() { return gms_emu::alloc_calls.load(); }
