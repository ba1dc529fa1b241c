// Shared host/device helpers for the bmmpy HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bmmpy_hip.h"

namespace bmmhip {

constexpr int kWave = 64;

// ---------------------------------------------------------------- errors --
// Status-carrying exception thrown inside the library and turned into a
// BMM_* status code at the C-ABI boundary.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] inline void fail(int status, const std::string& msg) { throw Error(status, msg); }

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        fail(BMM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define BMM_HIP(call) ::bmmhip::check((call), #call)

void set_last_error(const char* msg);

// Runs `fn`, mapping exceptions to status codes and the thread-local message.
template <typename F> int guarded(F&& fn) {
    try {
        fn();
        set_last_error("");
        return BMM_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return BMM_ERR_DEVICE;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return BMM_ERR_DEVICE;
    }
}

// --------------------------------------------------------- device buffers --
// Grow-only device allocation owned by a per-thread, per-device workspace.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    int device = -1;
    ~DevBuf() { release(); }
    void release() noexcept {
        if (ptr) {
            int cur = 0;
            (void)hipGetDevice(&cur);
            if (device >= 0 && device != cur)
                (void)hipSetDevice(device);
            (void)hipFree(ptr);
            if (device >= 0 && device != cur)
                (void)hipSetDevice(cur);
        }
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T> T* get(size_t count) {
        size_t need = count * sizeof(T);
        if (need == 0)
            need = 256;
        int cur = 0;
        BMM_HIP(hipGetDevice(&cur));
        if (need > bytes || cur != device) {
            release();
            size_t cap = 256;
            while (cap < need)
                cap <<= 1;
            BMM_HIP(hipMalloc(&ptr, cap));
            bytes = cap;
            device = cur;
        }
        return static_cast<T*>(ptr);
    }
};

// Grow-only page-locked host memory (hipHostMalloc): copies between it and
// the device are asynchronous.  Its contents must not be rewritten before
// the copies that read them have completed.
struct HostBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    ~HostBuf() {
        if (ptr)
            (void)hipHostFree(ptr);
    }
    template <typename T> T* get(size_t count) {
        size_t need = count * sizeof(T);
        if (need < 256)
            need = 256;
        if (need > bytes) {
            if (ptr)
                BMM_HIP(hipHostFree(ptr));
            ptr = nullptr;
            bytes = 0;
            size_t cap = 256;
            while (cap < need)
                cap <<= 1;
            BMM_HIP(hipHostMalloc(&ptr, cap, hipHostMallocDefault));
            bytes = cap;
        }
        return static_cast<T*>(ptr);
    }
};

// ------------------------------------------------------------ bit helpers --
__host__ __device__ inline uint64_t gray_of(uint64_t s) { return s ^ (s >> 1); }

__host__ __device__ inline uint64_t gray_inverse(uint64_t g) {
    g ^= g >> 1;
    g ^= g >> 2;
    g ^= g >> 4;
    g ^= g >> 8;
    g ^= g >> 16;
    g ^= g >> 32;
    return g;
}

// popcount(x) + acc as a single VALU op.
__device__ inline uint32_t bcnt_acc(uint32_t x, uint32_t acc) {
    uint32_t r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}

// Wave-wide minimum of a 64-bit value (lanes of one wavefront).
__device__ inline uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint64_t o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}

__device__ inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}

__device__ inline int32_t wave_max_i32(int32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, __shfl_xor(v, off, kWave));
    return v;
}

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kWave);
    return v;
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on XCD
// b % 8), and each XCD has its own L2.  xcd_tile(b, G) renumbers the G
// workgroups of a launch so that each XCD takes one contiguous range of
// tiles: neighbouring tiles share an L2, where partial cache lines at tile
// edges merge before they are written back.
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t G) {
    const uint32_t x = b % kXcds, q = G / kXcds, r = G % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

// Mask of the lanes whose digit (dbits bits) equals this lane's, among the
// lanes where `valid` holds.  Every lane of the wave must call it.
__device__ inline uint64_t match_digit(uint32_t d, int dbits, bool valid) {
    uint64_t match = __ballot(valid);
    for (int b = 0; b < dbits; ++b) {
        const uint64_t bal = __ballot((d >> b) & 1u);
        match &= ((d >> b) & 1u) ? bal : ~bal;
    }
    return match;
}

// Optional HIP-event timing of the dominant kernel of each entry point
// (bmm_hip_timing_begin / bmm_hip_timing_end).  When enabled, a
// ScopedKernelTimer records an event pair on the launch stream around the
// kernel it wraps.
struct KernelTimer {
    bool enabled = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};
KernelTimer& kernel_timer();

struct ScopedKernelTimer {
    hipStream_t stream;
    hipEvent_t stop = nullptr;
    explicit ScopedKernelTimer(hipStream_t s, bool active = true) : stream(s) {
        KernelTimer& t = kernel_timer();
        if (!t.enabled || !active)
            return;
        hipEvent_t a, b;
        BMM_HIP(hipEventCreate(&a));
        BMM_HIP(hipEventCreate(&b));
        BMM_HIP(hipEventRecord(a, stream));
        t.events.emplace_back(a, b);
        stop = b;
    }
    ~ScopedKernelTimer() {
        if (stop)
            (void)hipEventRecord(stop, stream);
    }
};

// Fill kernel used instead of hipMemsetAsync so the same launches can sit in
// a captured graph.
__global__ void fill_u64_kernel(uint64_t* p, size_t n, uint64_t v);
void fill_u64(uint64_t* p, size_t n, uint64_t v, hipStream_t s);
__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v);
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);

} // namespace bmmhip
