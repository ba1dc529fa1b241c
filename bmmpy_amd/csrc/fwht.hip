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
//     row segments; the two slices of a 128-byte line are neighbouring tiles
//     on one XCD); stages on row bits 6-11 in the load layout and 0-5 after
//     the exchange.
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
// 16 s + col (col < 16): element data[r 2^16 + 16 s + col].  A round of the
// exchange holds the rows with one value of row bit 5 at word
// x = e' + 16 (e' >> 9), e' = col + 16 (r & 31) + 512 (r >> 6): the reads of
// the second layout (lanes over col and row bits 6, 7) fall on 32 distinct
// banks modulo 2.
constexpr int kHi12LdsWords = 32768 + 1024;

__device__ inline uint32_t hi12_slot(uint32_t col, uint32_t r0_4, uint32_t r6_11) {
    const uint32_t x = col | (r0_4 << 4) | (r6_11 << 9);
    return x + 16u * (x >> 9);
}

__global__ __launch_bounds__(kTileThreads) void fwht_hi12_kernel(int32_t* __restrict__ data,
                                                                 uint32_t bytes) {
    __shared__ uint32_t lds[kHi12LdsWords];
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(data, 0, (int)bytes,
                                                                        0x00020000);
    const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t col = l & 15u;
    const uint32_t s = xcd_tile(blockIdx.x, gridDim.x);
    // Layout A: v[j] = row rt + 64 j with rt = (l >> 4) + 4 w (row bits 0-5
    // from the thread, 6-11 from j); one wavefront access covers 4 rows of
    // 64 bytes.
    const uint32_t rt = (l >> 4) | (w << 2);
    uint32_t v[64];
    const uint32_t va = ((rt << 16) + (s << 4) + col) * 4u;
#pragma unroll
    for (int j = 0; j < 64; ++j)
        v[j] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)va, j << 24, 0);
    reg_fwht_bits<63>(v);   // row bits 6-11
    // Layout B: b[k] = row k + 64 rt (row bits 0-5 from k, 6-11 from the
    // thread).  Rounds on row bit 5 (a wave bit of A, bit 5 of k).
    uint32_t b[64];
    uint32_t* const sa = lds + hi12_slot(col, rt & 31u, 0);   // + 528 j
    const uint32_t* const sb = lds + hi12_slot(col, 0, rt);   // + 16 k
#pragma unroll
    for (int rho = 0; rho < 2; ++rho) {
        if ((w >> 3) == (uint32_t)rho) {
#pragma unroll
            for (int j = 0; j < 64; ++j)
                sa[528 * j] = v[j];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 32; ++k)
            b[k + 32 * rho] = sb[16 * k];
        if (rho == 0)
            __syncthreads();
    }
    reg_fwht_bits<63>(b);   // row bits 0-5
    const uint32_t vb = ((rt << 22) + (s << 4) + col) * 4u;
#pragma unroll
    for (int k = 0; k < 64; ++k)
        __builtin_amdgcn_raw_buffer_store_b32((int)b[k], rs, (int)vb, k << 18, 0);
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
