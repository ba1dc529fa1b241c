// Walsh-Hadamard kernels on gfx950.
//
// 1. fwht_chunk_kernel / fwht_stride_buf_kernel (fwht_stride_kernel for
//    arrays over 2 GiB): the stand-alone in-place transform of a device
//    array (fwht_inplace, src/math/fwht.cpp:28-57):
//    one LDS pass over the low 14 index bits, then register passes of up to
//    7 bits with coalesced 256-byte wavefront accesses.  Aligned 2^28 int32
//    arrays take two passes of 2^16-element register tiles instead
//    (fwht_lo16_kernel, fwht_hi12_kernel).
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
//     LDS or barriers.  2^27 int32 takes three passes (14 + 7 + 6 bits);
//     2^28 int32 takes the two-pass tile kernels below.
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

// ------------------------------------------- two passes for 2^28 int32 --
// A 1 GiB int32 array in two passes of 2^16-element tiles held in registers
// (1024 threads x 64 values = 256 KiB, one workgroup per CU), each tile
// reshuffled between register layouts through LDS in rounds of 128 KiB:
//   * fwht_lo16_kernel: index bits 0-15 of a contiguous 256 KiB tile
//     (16-byte loads, 1 KiB per wavefront access; 4-byte stores, 256 bytes
//     per wavefront access), three register layouts;
//   * fwht_hi12_kernel: index bits 16-27 of a slice of 16 columns (64-byte
//     row segments, 16-byte accesses; the two slices of a 128-byte line are
//     neighbouring tiles on one XCD), three register layouts.
// Each pass reads and writes the array once: 16 B per element in all,
// against 24 B for the LDS pass plus two register passes.  With one
// workgroup per CU the transform work of a tile does not overlap its own
// loads; the CUs drift apart, and those in their memory phase keep HBM busy.
typedef int i32x4 __attribute__((ext_vector_type(4)));
constexpr int kTileThreads = 1024;

// Butterflies over the index bits of v[] that are set in MASK.
template <int MASK, int E> __device__ inline void reg_fwht_bits(uint32_t (&v)[E]) {
#pragma unroll
    for (int len = 1; len < E; len <<= 1)
        if (MASK & len)
#pragma unroll
            for (int i = 0; i < E; ++i)
                if ((i & len) == 0) {
                    const uint32_t a = v[i], b = v[i + len];
                    v[i] = a + b;
                    v[i + len] = a - b;
                }
}

// Pass 1: element e of tile t is data[t 2^16 + e].  Three register
// layouts, 64 values per thread:
//   A: v[q + 4 j] = e = q + 4 l + 256 w + 4096 j   (stages on bits 0, 1, 12-15)
//   B: b[k] = (l & 3) + 4 (k & 31) + 128 (l >> 2) + 2048 w + 32768 (k >> 5)
//                                                 (stages on bits 2-6)
//   C: c[m] = l + 64 m + 4096 w                    (stages on bits 7-11)
// (l = lane, w = wave).  Each exchange runs in two rounds on a bit held in
// registers on both sides (A -> B on bit 15, B -> C on bit 6), so in a
// round every thread writes 32 values and reads 32: at most 64 values are
// live.  A round's 2^15 elements sit at word e' + 4 (e' >> 7) (first
// exchange) or e' + 4 (e' >> 6) (second), e' = e without the round bit:
// the B-layout accesses (lanes over bits 0, 1, 7-10) then fall on 32
// distinct banks modulo 2, the A writes are 16-byte aligned, the C reads
// consecutive.  Every slot is a per-thread base plus a constant.
constexpr int kLo16LdsWords = 32768 + 2048;

__global__ __launch_bounds__(kTileThreads) void fwht_lo16_kernel(int32_t* __restrict__ data,
                                                                 uint32_t bytes) {
    __shared__ __align__(16) uint32_t lds[kLo16LdsWords];
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(data, 0, (int)bytes,
                                                                        0x00020000);
    const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t base = xcd_tile(blockIdx.x, gridDim.x) << 16;
    uint32_t v[64];
    const uint32_t va = (base + (l << 2) + (w << 8)) * 4u;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const i32x4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)va, j << 14, 0);
        v[4 * j] = (uint32_t)a[0];
        v[4 * j + 1] = (uint32_t)a[1];
        v[4 * j + 2] = (uint32_t)a[2];
        v[4 * j + 3] = (uint32_t)a[3];
    }
    reg_fwht_bits<63>(v);
    const uint32_t lb = (l >> 2) & 15u;
    uint32_t b[64];
    {
        uint32_t* const sa = lds + (l << 2) + 264u * w + 4u * (l >> 5);   // + 4224 (j & 7)
        const uint32_t* const sb = lds + (l & 3u) + (lb << 7) + (w << 11) + 4u * (lb | (w << 4));
#pragma unroll
        for (int rho = 0; rho < 2; ++rho) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int jj = j + 8 * rho;
                *reinterpret_cast<uint4*>(sa + 4224 * j) =
                    make_uint4(v[4 * jj], v[4 * jj + 1], v[4 * jj + 2], v[4 * jj + 3]);
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 32; ++k)
                b[k + 32 * rho] = sb[4 * k];
            __syncthreads();
        }
    }
    reg_fwht_bits<31>(b);
    uint32_t c[64];
    {
        uint32_t* const sb2 = lds + (l & 3u) + (lb << 6) + (w << 10) + 4u * (lb | (w << 4));
        const uint32_t* const sc = lds + l + (w << 11) + (w << 7);   // + 68 (m >> 1)
#pragma unroll
        for (int rho = 0; rho < 2; ++rho) {
#pragma unroll
            for (int k = 0; k < 64; ++k)
                if (((k >> 4) & 1) == rho)
                    sb2[4 * (k & 15) + 17408 * (k >> 5)] = b[k];
            __syncthreads();
#pragma unroll
            for (int mh = 0; mh < 32; ++mh)
                c[2 * mh + rho] = sc[68 * mh];
            if (rho == 0)
                __syncthreads();
        }
    }
    reg_fwht_bits<62>(c);
    const uint32_t vc = (base + l + (w << 12)) * 4u;
#pragma unroll
    for (int m = 0; m < 64; ++m)
        __builtin_amdgcn_raw_buffer_store_b32((int)c[m], rs, (int)vc, m << 8, 0);
}

// Pass 2: tile s holds rows r < 2^12 (index bits 16-27) of columns
// 16 s + c (c < 16): element data[r 2^16 + 16 s + c].  Each lane owns four
// consecutive columns (16-byte loads and stores, one wavefront access = 16
// rows of 64 bytes); with cq = l & 3 and the row bits spread over lane bits
// 2-5, waves and registers, three layouts:
//   A: v[4 j + q] = row (l >> 2) + 16 w + 256 j   (stages on row bits 8-11)
//   B: b[4 k + q] = row (l >> 2) + 16 k + 256 w   (stages on row bits 4-7)
//   C: c[4 m + q] = row m + 16 (l >> 2) + 256 w   (stages on row bits 0-3)
// all at column 4 cq + q.  The column bits stay in registers, so each
// exchange runs in two rounds on column bit 1 with 32 values written and 32
// read per thread (8-byte LDS accesses).  A round's element (r, c) sits at
// word e + 8 (e >> 7), e = (c & 1) + 2 (c >> 2) + 8 r: every access of 16
// lanes covers 32 distinct banks.
constexpr int kHi12LdsWords = 32768 + 2048;

// The loads of band j (rows 256 j .. 256 j + 255) and the stores of row
// group m are issued in the order j = i ^ ROT, m = i ^ ROT (i = 0 .. 15),
// with ROT = (s >> 1) mod 16: neighbouring line pairs start on different
// DRAM rows instead of all tiles in flight opening the same ones (a probe of
// the loads and stores alone ran 8% faster so; the kernel 1-2%).  One code
// path per ROT keeps every register index a constant, and a scheduling
// barrier after each access keeps the issue order.
template <int ROT>
__device__ inline void hi12_load_bands(__amdgpu_buffer_rsrc_t rs, uint32_t va, uint32_t (&v)[64]) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int j = i ^ ROT;
        const i32x4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)va, j << 26, 0);
        v[4 * j] = (uint32_t)a[0];
        v[4 * j + 1] = (uint32_t)a[1];
        v[4 * j + 2] = (uint32_t)a[2];
        v[4 * j + 3] = (uint32_t)a[3];
        __builtin_amdgcn_sched_barrier(0);   // keep the issue order
    }
}

template <int ROT>
__device__ inline void hi12_store_groups(__amdgpu_buffer_rsrc_t rs, uint32_t vc,
                                         const uint32_t (&c)[64]) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int m = i ^ ROT;
        i32x4 a;
        a[0] = (int)c[4 * m];
        a[1] = (int)c[4 * m + 1];
        a[2] = (int)c[4 * m + 2];
        a[3] = (int)c[4 * m + 3];
        __builtin_amdgcn_raw_buffer_store_b128(a, rs, (int)vc, m << 18, 0);
        __builtin_amdgcn_sched_barrier(0);
    }
}

#define BMM_ROT16(F, ...)                                                                      \
    switch (rot) {                                                                             \
    case 0: F<0>(__VA_ARGS__); break;                                                          \
    case 1: F<1>(__VA_ARGS__); break;                                                          \
    case 2: F<2>(__VA_ARGS__); break;                                                          \
    case 3: F<3>(__VA_ARGS__); break;                                                          \
    case 4: F<4>(__VA_ARGS__); break;                                                          \
    case 5: F<5>(__VA_ARGS__); break;                                                          \
    case 6: F<6>(__VA_ARGS__); break;                                                          \
    case 7: F<7>(__VA_ARGS__); break;                                                          \
    case 8: F<8>(__VA_ARGS__); break;                                                          \
    case 9: F<9>(__VA_ARGS__); break;                                                          \
    case 10: F<10>(__VA_ARGS__); break;                                                        \
    case 11: F<11>(__VA_ARGS__); break;                                                        \
    case 12: F<12>(__VA_ARGS__); break;                                                        \
    case 13: F<13>(__VA_ARGS__); break;                                                        \
    case 14: F<14>(__VA_ARGS__); break;                                                        \
    default: F<15>(__VA_ARGS__); break;                                                        \
    }

__global__ __launch_bounds__(kTileThreads) void fwht_hi12_kernel(int32_t* __restrict__ data,
                                                                 uint32_t bytes) {
    __shared__ __align__(16) uint32_t lds[kHi12LdsWords];
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(data, 0, (int)bytes,
                                                                        0x00020000);
    const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t cq = l & 3u, lr = l >> 2;
    const uint32_t s = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t rot = (s >> 1) & 15u;
    uint32_t v[64];
    const uint32_t va = (((lr + 16u * w) << 16) + (s << 4) + 4u * cq) * 4u;
    BMM_ROT16(hi12_load_bands, rs, va, v)
    reg_fwht_bits<60>(v);   // row bits 8-11
    uint32_t* const sa = lds + 2u * cq + 8u * lr + 136u * w;    // A: + 2176 j
    uint32_t* const sb = lds + 2u * cq + 8u * lr + 2176u * w;   // B: + 136 k
    const uint32_t* const sc = lds + 2u * cq + 136u * lr + 2176u * w;   // C: + 8 m
    uint32_t b[64];
#pragma unroll
    for (int rho = 0; rho < 2; ++rho) {
#pragma unroll
        for (int j = 0; j < 16; ++j)
            *reinterpret_cast<uint2*>(sa + 2176 * j) =
                make_uint2(v[4 * j + 2 * rho], v[4 * j + 2 * rho + 1]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint2 x = *reinterpret_cast<const uint2*>(sb + 136 * k);
            b[4 * k + 2 * rho] = x.x;
            b[4 * k + 2 * rho + 1] = x.y;
        }
        __syncthreads();
    }
    reg_fwht_bits<60>(b);   // row bits 4-7
    uint32_t c[64];
#pragma unroll
    for (int rho = 0; rho < 2; ++rho) {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            *reinterpret_cast<uint2*>(sb + 136 * k) =
                make_uint2(b[4 * k + 2 * rho], b[4 * k + 2 * rho + 1]);
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const uint2 x = *reinterpret_cast<const uint2*>(sc + 8 * m);
            c[4 * m + 2 * rho] = x.x;
            c[4 * m + 2 * rho + 1] = x.y;
        }
        if (rho == 0)
            __syncthreads();
    }
    reg_fwht_bits<60>(c);   // row bits 0-3
    const uint32_t vc = (((16u * lr + 256u * w) << 16) + (s << 4) + 4u * cq) * 4u;
    BMM_ROT16(hi12_store_groups, rs, vc, c)
}
#undef BMM_ROT16

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
    if constexpr (sizeof(T) == 4) {
        if (nb == 28 && (uintptr_t)d % 16 == 0) {
            const unsigned tiles = 1u << 12;
            const uint32_t bytes = (uint32_t)(n * sizeof(T));
            hipLaunchKernelGGL(fwht_lo16_kernel, dim3(tiles), dim3(kTileThreads), 0, stream,
                               reinterpret_cast<int32_t*>(d), bytes);
            hipLaunchKernelGGL(fwht_hi12_kernel, dim3(tiles), dim3(kTileThreads), 0, stream,
                               reinterpret_cast<int32_t*>(d), bytes);
            BMM_HIP(hipGetLastError());
            return;
        }
    }
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
