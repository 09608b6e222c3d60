// Host emulation of the HIP features the library's kernels and host code use,
// so that the unmodified kernel sources build with g++ and run on the CPU
// (test infrastructure: bounds checking with AddressSanitizer and CPU parity
// tests of the real kernel code at small sizes).
//
// Execution model: a kernel launch runs its workgroups one after another;
// every workgroup is blockDim.x host threads that meet at __syncthreads()
// (a std::barrier).  __shared__ arrays become function-static storage, which
// is what LDS is to a sequence of workgroups.  DPP lane moves and readlane
// are wave collectives: the 64 threads of a wave publish their values and
// meet at a wave barrier.  All kernel call sites of these collectives are in
// uniform control flow, as on the GPU.
#pragma once

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <barrier>
#include <chrono>
#include <cstdio>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))

struct gms_emu_dim3 {
    uint32_t x, y, z;
};
typedef struct dim3 {
    uint32_t x, y, z;
    dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
} dim3;

struct alignas(16) uint4 {
    uint32_t x, y, z, w;
};
inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return uint4{a, b, c, d}; }
struct alignas(8) uint2 {
    uint32_t x, y;
};
inline uint2 make_uint2(uint32_t a, uint32_t b) { return uint2{a, b}; }

namespace gms_emu {
extern thread_local gms_emu_dim3 thread_idx;
extern gms_emu_dim3 block_idx, block_dim, grid_dim;
void syncthreads();
uint32_t update_dpp(uint32_t old, uint32_t src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl);
int readlane(int x, int lane);
uint64_t ballot(bool pred);
void wave_sync();
void launch(dim3 grid, dim3 block, std::function<void()> body);
}  // namespace gms_emu

#define threadIdx (gms_emu::thread_idx)
#define blockIdx (gms_emu::block_idx)
#define blockDim (gms_emu::block_dim)
#define gridDim (gms_emu::grid_dim)
#define __syncthreads() gms_emu::syncthreads()

inline uint32_t atomicAdd(uint32_t* p, uint32_t v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline uint32_t atomicMax(uint32_t* p, uint32_t v) {
    uint32_t old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline uint64_t __umul64hi(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}
inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
inline float __uint_as_float(uint32_t u) {
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
}
inline float __fmul_rn(float a, float b) {
    volatile float r = a * b;
    return r;
}
inline float __fsub_rn(float a, float b) {
    volatile float r = a - b;
    return r;
}
inline float __fdiv_rn(float a, float b) {
    volatile float r = a / b;
    return r;
}

// ---- runtime API -----------------------------------------------------------
typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNoDevice = 100,
    hipErrorNotReady = 600
} hipError_t;
typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
} hipMemcpyKind;
typedef void* hipStream_t;
struct gms_emu_event {
    std::chrono::steady_clock::time_point t;
};
typedef gms_emu_event* hipEvent_t;

inline hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
namespace gms_emu {
// Device and pinned allocations and frees made so far (gms_emu_alloc_calls).
extern std::atomic<uint64_t> alloc_calls;
}
// A __device__ variable is a host global here: its address is the device one.
#define HIP_SYMBOL(x) reinterpret_cast<const void*>(&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}
inline hipError_t hipMalloc(void** p, size_t bytes) {
    ++gms_emu::alloc_calls;
    *p = aligned_alloc(256, ((bytes + 255) / 256) * 256 + 256);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes) {
    return hipMalloc((void**)p, bytes);
}
inline hipError_t hipMemset(void* p, int v, size_t bytes) {
    memset(p, v, bytes);
    return hipSuccess;
}
inline hipError_t hipFree(void* p) {
    ++gms_emu::alloc_calls;
    free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t b, hipMemcpyKind, hipStream_t) {
    memmove(d, s, b);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t b, hipMemcpyKind) {
    memmove(d, s, b);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new gms_emu_event();
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) {
    e->t = std::chrono::steady_clock::now();
    return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
// Emulated kernels finish inside their launch, so every recorded event is complete.
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
#define hipEventDisableTiming 2u
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}

// Pinned host memory is ordinary host memory here.
#define hipHostMallocDefault 0u
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
    ++gms_emu::alloc_calls;
    *p = aligned_alloc(256, ((bytes + 255) / 256) * 256);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipHostFree(void* p) {
    ++gms_emu::alloc_calls;
    free(p);
    return hipSuccess;
}

// Device attributes and occupancy of the emulated device.
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
    *v = 4;
    return hipSuccess;
}
template <class Fn>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, Fn, int, size_t) {
    *n = 2;
    return hipSuccess;
}

// Streams exist only as handles; graphs are not emulated.
#define hipStreamNonBlocking 1u
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };
typedef struct gms_emu_graph* hipGraph_t;
typedef struct gms_emu_graph_exec* hipGraphExec_t;
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    static int dummy;
    *s = &dummy;
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorInvalidValue; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    *g = nullptr;
    return hipErrorInvalidValue;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) {
    return hipErrorInvalidValue;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }

#define hipLaunchKernelGGL(kern, grid, block, shmem, stream, ...) \
    gms_emu::launch(dim3(grid), dim3(block), [=]() { kern(__VA_ARGS__); })
