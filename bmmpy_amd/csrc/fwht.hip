// Walsh-Hadamard kernels on gfx950.
//
// 1. fwht_pass_kernel: the stand-alone in-place transform of a device array
//    (fwht_inplace, src/math/fwht.cpp:28-57), as passes of up to 12 index
//    bits each: the first pass transforms contiguous 4096-element chunks in
//    LDS, later passes transform 64-column tiles strided by 2^lo, so every
//    global access is a coalesced 256-byte row.
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
template <typename T> struct Unsigned;
template <> struct Unsigned<int32_t> { using type = uint32_t; };
template <> struct Unsigned<int16_t> { using type = uint16_t; };

// Transforms index bits [lo, lo + pbits) of a 2^n array.  lo == 0: each
// workgroup owns one contiguous chunk of 2^pbits elements.  lo > 0: each
// workgroup owns a tile of 2^pbits rows (stride 2^lo) x cw adjacent columns.
template <typename T>
__global__ __launch_bounds__(256) void fwht_pass_kernel(T* __restrict__ data, int n, int lo,
                                                        int pbits) {
    using U = typename Unsigned<T>::type;
    extern __shared__ unsigned char smem_raw[];
    U* s = reinterpret_cast<U*>(smem_raw);
    const int rows = 1 << pbits;
    const int cw = lo == 0 ? 1 : min(64, 1 << lo);
    const int tile = rows * cw;
    // Tile origin: column group within the low part, then the high part.
    const size_t low_groups = lo == 0 ? 1 : ((size_t)1 << lo) / cw;
    const size_t gid = blockIdx.x;
    const size_t col0 = (gid % low_groups) * cw;
    const size_t hi = gid / low_groups;
    const size_t origin = (hi << (lo + pbits)) + col0;
    const size_t rstride = (size_t)1 << lo;
    for (int i = threadIdx.x; i < tile; i += blockDim.x) {
        const int r = i / cw, c = i % cw;
        s[i] = (U)data[origin + (size_t)r * rstride + c];
    }
    __syncthreads();
    for (int len = 1; len < rows; len <<= 1) {
        for (int i = threadIdx.x; i < tile / 2; i += blockDim.x) {
            const int pair = i / cw, c = i % cw;
            const int grp = pair / len, off = pair % len;
            const int r0 = grp * 2 * len + off, r1 = r0 + len;
            const U a = s[r0 * cw + c], b = s[r1 * cw + c];
            s[r0 * cw + c] = (U)(a + b);
            s[r1 * cw + c] = (U)(a - b);
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < tile; i += blockDim.x) {
        const int r = i / cw, c = i % cw;
        data[origin + (size_t)r * rstride + c] = (T)s[i];
    }
}

template <typename T> void fwht_array_dev(T* d, size_t n, hipStream_t stream) {
    if (n <= 1)
        return;
    if ((n & (n - 1)) != 0)
        fail(BMM_ERR_INVALID, "fwht_inplace: size must be a power of two");
    int nb = 0;
    while (((size_t)1 << nb) < n)
        ++nb;
    int lo = 0;
    while (lo < nb) {
        const int pbits = lo == 0 ? std::min(nb, 12) : std::min(nb - lo, 6);
        const int cw = lo == 0 ? 1 : std::min(64, 1 << lo);
        const size_t tile = ((size_t)1 << pbits) * cw;
        const size_t grid = n / tile;
        hipLaunchKernelGGL(fwht_pass_kernel<T>, dim3((unsigned)grid), dim3(256),
                           tile * sizeof(T), stream, d, nb, lo, pbits);
        BMM_HIP(hipGetLastError());
        lo += pbits;
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
