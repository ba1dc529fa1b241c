// Walsh-Hadamard kernels on gfx950.
//
// 1. fwht_chunk_kernel / fwht_stride_buf_kernel (fwht_stride_kernel for
//    arrays over 2 GiB): the stand-alone in-place transform of a device
//    array (fwht_inplace, src/math/fwht.cpp:28-57):
//    one LDS pass over the low 14 index bits, then register passes of up to
//    7 bits with coalesced 256-byte wavefront accesses.
//
// 2. column_patterns_kernel: per-column row patterns of a small matrix.
//
// The searches themselves never stage transforms in HBM: they split masks
// m = mh << L | ml and transform 2^L entries per workgroup on chip
// (fwht_block.hpp; batch.hip, search.hip).
#include "common.hpp"
#include "fwht_block.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

namespace bmmhip {

// ------------------------------------------------------ array transform --
// fwht_inplace on a device array of 2^nb elements, wrapping arithmetic: the
// radix-2 stages of the different index bits commute, so the transform is
// applied as
//   * one LDS pass over index bits [0, min(nb, 14)): each workgroup owns a
//     contiguous chunk of 64 KiB (int32), loaded and stored as 16-byte
//     vectors, transformed by block_fwht (radix-32 register rounds between
//     LDS exchanges);
//   * register passes over the remaining bits, at most 7 at a time, split
//     evenly: lane c of a wavefront owns column c of a 2^R x 64 tile at row
//     stride 2^lo, so each of its 2^R loads and stores is one coalesced
//     256-byte wavefront access, and the R stages run in registers with no
//     LDS or barriers.  2^28 int32 takes three passes (14 + 7 + 7 bits).
// Each pass reads and writes the array once (8 bytes per int32 element).
// int16 arrays are computed in 32-bit lanes and truncated on store: the
// transform is a ring homomorphism, so the result equals int16 wrapping.
template <typename T> struct Unsigned;
template <> struct Unsigned<int32_t> { using type = uint32_t; };
template <> struct Unsigned<int16_t> { using type = uint16_t; };

constexpr int kArrBlockBits = 14;
constexpr int kArrMaxR = 7;        // register bits per stride pass (2^7 values per lane)

constexpr int kChunkThreads = 1024;   // 32 waves per CU at two 64 KiB chunks per CU

template <typename T, int L>
__global__ __launch_bounds__(kChunkThreads) void fwht_chunk_kernel(T* __restrict__ data, int vec) {
    __shared__ __align__(16) int32_t s[fwht_lds_words(L) + 4];
    constexpr int N = 1 << L;
    T* chunk = data + (size_t)blockIdx.x * N;
    constexpr int V = 16 / sizeof(T);   // elements per 16-byte vector
    if (N >= V * 256 && vec) {
        for (int i = threadIdx.x; i < N / V; i += blockDim.x) {
            const int4 q = reinterpret_cast<const int4*>(chunk)[i];
            const T* e = reinterpret_cast<const T*>(&q);
#pragma unroll
            for (int k = 0; k < V; ++k)
                s[fwht_pad(i * V + k)] = (int32_t)e[k];
        }
    } else {
        for (int i = threadIdx.x; i < N; i += blockDim.x)
            s[fwht_pad(i)] = (int32_t)chunk[i];
    }
    block_fwht<L>(s);
    if (N >= V * 256 && vec) {
        for (int i = threadIdx.x; i < N / V; i += blockDim.x) {
            int4 q;
            T* e = reinterpret_cast<T*>(&q);
#pragma unroll
            for (int k = 0; k < V; ++k)
                e[k] = (T)s[fwht_pad(i * V + k)];
            reinterpret_cast<int4*>(chunk)[i] = q;
        }
    } else {
        for (int i = threadIdx.x; i < N; i += blockDim.x)
            chunk[i] = (T)s[fwht_pad(i)];
    }
}

// Index bits [lo, lo + R) with lo >= 14: thread (tile, c) transforms the 2^R
// elements h 2^(lo+R) + e 2^lo + c, e < 2^R; 256 consecutive columns per
// workgroup.
template <typename T, int R>
__global__ __launch_bounds__(256) void fwht_stride_kernel(T* __restrict__ data, int lo) {
    const size_t cols = (size_t)1 << lo;
    const size_t col = (size_t)blockIdx.x * 256 + threadIdx.x;   // over (h, c)
    const size_t h = col >> lo, c = col & (cols - 1);
    T* p = data + (h << (lo + R)) + c;
    constexpr int E = 1 << R;
    uint32_t v[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
        v[e] = (uint32_t)(int32_t)p[(size_t)e << lo];
#pragma unroll
    for (int len = 1; len < E; len <<= 1)
#pragma unroll
        for (int i = 0; i < E; ++i)
            if ((i & len) == 0) {
                const uint32_t a = v[i], b = v[i + len];
                v[i] = a + b;
                v[i + len] = a - b;
            }
#pragma unroll
    for (int e = 0; e < E; ++e)
        p[(size_t)e << lo] = (T)(int32_t)v[e];
}

// The same pass through buffer loads and stores (arrays of at most 2 GiB):
// the per-element stride goes in the scalar offset, so a lane keeps one
// 32-bit base offset besides its 2^R values (the global-pointer form keeps a
// 64-bit address per value, which caps the occupancy).
template <typename T> __device__ inline uint32_t buf_load(__amdgpu_buffer_rsrc_t rs, uint32_t v,
                                                          uint32_t s) {
    if constexpr (sizeof(T) == 4)
        return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)v, (int)s, 0);
    else
        return (uint32_t)(int32_t)(int16_t)__builtin_amdgcn_raw_buffer_load_b16(rs, (int)v, (int)s,
                                                                                0);
}

template <typename T> __device__ inline void buf_store(uint32_t x, __amdgpu_buffer_rsrc_t rs,
                                                       uint32_t v, uint32_t s) {
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(x, rs, (int)v, (int)s, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)x, rs, (int)v, (int)s, 0);
}

template <typename T, int R>
__global__ __launch_bounds__(256) void fwht_stride_buf_kernel(T* __restrict__ data,
                                                              uint32_t bytes, int lo) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(data, 0, (int)bytes, 0x00020000);
    const uint32_t col = blockIdx.x * 256u + threadIdx.x;   // over (h, c)
    const uint32_t h = col >> lo, c = col & ((1u << lo) - 1u);
    const uint32_t vbase = ((h << (lo + R)) + c) * (uint32_t)sizeof(T);
    const uint32_t stride = (1u << lo) * (uint32_t)sizeof(T);
    constexpr int E = 1 << R;
    uint32_t v[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
        v[e] = buf_load<T>(rs, vbase, (uint32_t)e * stride);
#pragma unroll
    for (int len = 1; len < E; len <<= 1)
#pragma unroll
        for (int i = 0; i < E; ++i)
            if ((i & len) == 0) {
                const uint32_t a = v[i], b = v[i + len];
                v[i] = a + b;
                v[i + len] = a - b;
            }
#pragma unroll
    for (int e = 0; e < E; ++e)
        buf_store<T>(v[e], rs, vbase, (uint32_t)e * stride);
}

template <typename T> void launch_chunk_pass(T* d, size_t n, int L, hipStream_t stream) {
    const unsigned grid = (unsigned)(n >> L);
    const int vec = ((uintptr_t)d % 16) == 0;   // 16-byte vectors only on aligned arrays
    switch (L) {
#define BMM_L(LL)                                                                              \
    case LL:                                                                                   \
        hipLaunchKernelGGL((fwht_chunk_kernel<T, LL>), dim3(grid), dim3(kChunkThreads), 0,     \
                           stream, d,                                                          \
                           vec);                                                               \
        break;
        BMM_L(1) BMM_L(2) BMM_L(3) BMM_L(4) BMM_L(5) BMM_L(6) BMM_L(7) BMM_L(8) BMM_L(9)
        BMM_L(10) BMM_L(11) BMM_L(12) BMM_L(13) BMM_L(14)
#undef BMM_L
    default: fail(BMM_ERR_INVALID, "fwht_inplace: bad chunk size");
    }
    BMM_HIP(hipGetLastError());
}

template <typename T> void fwht_array_dev(T* d, size_t n, hipStream_t stream) {
    if (n <= 1)
        return;
    if ((n & (n - 1)) != 0)
        fail(BMM_ERR_INVALID, "fwht_inplace: size must be a power of two");
    int nb = 0;
    while (((size_t)1 << nb) < n)
        ++nb;
    const int L = std::min(nb, kArrBlockBits);
    launch_chunk_pass(d, n, L, stream);
    // Remaining bits in the fewest register passes of at most kArrMaxR bits,
    // split evenly (10 bits: 5 + 5; 14 bits: 7 + 7).  Arrays over 2 GiB take
    // the global-pointer passes, at most 5 bits each (a 64-bit address per
    // value otherwise exhausts the registers).
    const bool buf = n * sizeof(T) <= ((size_t)1 << 31);
    const int max_r = buf ? kArrMaxR : 5;
    int lo = L;
    while (lo < nb) {
        const int rem = nb - lo;
        const int passes = (rem + max_r - 1) / max_r;
        const int R = (rem + passes - 1) / passes;
        const unsigned grid = (unsigned)((n >> R) / 256);
        if (buf) {
            const uint32_t bytes = (uint32_t)(n * sizeof(T));
            switch (R) {
#define BMM_R(RR)                                                                              \
    case RR:                                                                                   \
        hipLaunchKernelGGL((fwht_stride_buf_kernel<T, RR>), dim3(grid), dim3(256), 0, stream, d, \
                           bytes, lo);                                                         \
        break;
                BMM_R(1) BMM_R(2) BMM_R(3) BMM_R(4) BMM_R(5) BMM_R(6) BMM_R(7)
#undef BMM_R
            default: fail(BMM_ERR_INVALID, "fwht_inplace: bad stride pass");
            }
        } else {
            switch (R) {
#define BMM_R(RR)                                                                              \
    case RR:                                                                                   \
        hipLaunchKernelGGL((fwht_stride_kernel<T, RR>), dim3(grid), dim3(256), 0, stream, d, lo); \
        break;
                BMM_R(1) BMM_R(2) BMM_R(3) BMM_R(4) BMM_R(5) BMM_R(6) BMM_R(7)
#undef BMM_R
            default: fail(BMM_ERR_INVALID, "fwht_inplace: bad stride pass");
            }
        }
        BMM_HIP(hipGetLastError());
        lo += R;
    }
}

template void fwht_array_dev<int32_t>(int32_t*, size_t, hipStream_t);
template void fwht_array_dev<int16_t>(int16_t*, size_t, hipStream_t);

// ------------------------------------------------------- column patterns --
// pattern[c] for c < cols of an r-row matrix (r <= 32), 0 when inactive.
__global__ void column_patterns_kernel(const uint64_t* __restrict__ rows, int r, int words,
                                       int cols, uint32_t* __restrict__ pat) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols)
        return;
    const int wi = c >> 6, bit = c & 63;
    uint32_t p = 0;
    for (int i = 0; i < r; ++i)
        p |= (uint32_t)((rows[(size_t)i * words + wi] >> bit) & 1ull) << i;
    pat[c] = p;
}

int default_persistent_wgs() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return cus * 8;
}

} // namespace bmmhip
