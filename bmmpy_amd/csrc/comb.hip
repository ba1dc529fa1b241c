// fixed_weight_masks_u32 / _u64 (src/math/comb.cpp:13-82): every n-bit mask
// with exactly k set bits, in increasing order (k = 0 gives the one mask 0).
//
// Increasing numeric order of the k-subsets of {0..n-1} is their colex order,
// in which the subset c_k > ... > c_1 has rank sum_j C(c_j, j).  Each thread
// unranks the first of its run of consecutive masks (c_j = the largest c
// with C(c, j) <= the remaining rank, scanning c downwards, at most n steps
// for all j, binomials from an LDS copy of the table) and steps through the
// rest with Gosper's successor,
// m' = (((r ^ m) >> 2) >> ctz(m & -m)) | r with r = m + (m & -m).  The
// masks of a workgroup go through LDS (runs padded by one entry against bank
// conflicts) so that the stores are coalesced, 16 bytes per lane.
#include "common.hpp"

#include <algorithm>

namespace bmmhip {
namespace {

// Consecutive masks per thread: 128 bytes of them (32 u32 / 16 u64), so a
// workgroup stages 32 KiB.
template <typename T> constexpr int comb_run() { return 128 / (int)sizeof(T); }
constexpr int kCombThreads = 256;

// C(c, j) for c, j <= 64 (C(64, 32) < 2^61 fits).
struct BinomTable {
    uint64_t c[65][65];
    constexpr BinomTable() : c() {
        for (int i = 0; i <= 64; ++i) {
            c[i][0] = 1;
            for (int j = 1; j <= i; ++j)
                c[i][j] = c[i - 1][j - 1] + (j < i ? c[i - 1][j] : 0);
        }
    }
};
constexpr BinomTable kBinomHost{};
__constant__ BinomTable kBinom{};

// The rows C(c, j), c <= n, j <= k, of the table go to LDS first: the
// unranking's lookups differ per lane, and from LDS each is a short
// dependent step instead of a memory round trip.
template <typename T>
__global__ __launch_bounds__(kCombThreads) void fixed_weight_kernel(uint32_t n, uint32_t k,
                                                                    uint64_t first, uint64_t count,
                                                                    T* __restrict__ out) {
    constexpr int kCombRun = comb_run<T>();
    constexpr int kRow = kCombRun + 1;   // padded: lanes' runs start on distinct banks
    __shared__ T stage[kCombThreads * kRow];
    extern __shared__ __align__(8) uint64_t tab[];   // (n + 1) x (k + 1)
    const int K1 = (int)k + 1;
    for (int q = threadIdx.x; q < ((int)n + 1) * K1; q += kCombThreads)
        tab[q] = kBinom.c[q / K1][q % K1];
    __syncthreads();
    const uint64_t base = first + (uint64_t)blockIdx.x * kCombThreads * kCombRun;
    const uint64_t i0 = base + (uint64_t)threadIdx.x * kCombRun;
    if (i0 < count) {
        uint64_t rank = i0, mask = 0;
        int c = (int)n;
        for (int j = (int)k; j >= 1; --j) {
            uint64_t b;
            do {
                --c;
                b = tab[c * K1 + j];
            } while (b > rank);
            mask |= 1ull << c;
            rank -= b;
        }
#pragma unroll
        for (int m = 0; m < kCombRun; ++m) {
            stage[threadIdx.x * kRow + m] = (T)mask;
            const uint64_t low = mask & (~mask + 1);
            const uint64_t r = mask + low;
            mask = ((((r ^ mask) >> 2) >> __builtin_ctzll(low)) | r);
        }
    }
    __syncthreads();
    // 16-byte stores of kV consecutive masks (they lie in one thread's run);
    // a 16-byte aligned output takes them whole, the end of the range and
    // an unaligned output go mask by mask.
    constexpr int kV = 16 / (int)sizeof(T);
    static_assert(kCombRun % kV == 0, "a 16-byte group stays in one run");
    const bool vec = (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
    for (int q = threadIdx.x * kV; q < kCombThreads * kCombRun; q += kCombThreads * kV) {
        const uint64_t o = base + (uint64_t)q;
        const T* src = stage + (q / kCombRun) * kRow + q % kCombRun;
        if (vec && o + kV <= count) {
            T v[kV];
#pragma unroll
            for (int e = 0; e < kV; ++e)
                v[e] = src[e];
            *reinterpret_cast<uint4*>(out + o) = __builtin_bit_cast(uint4, v);
        } else {
            for (int e = 0; e < kV; ++e)
                if (o + e < count)
                    out[o + e] = src[e];
        }
    }
}

} // namespace

uint64_t binom_u64(uint32_t n, uint32_t k) { return k > n ? 0 : kBinomHost.c[n][k]; }

// out[0 .. C(n, k)) on the device, 1 <= k <= n <= 8 * sizeof(T).
template <typename T>
void fixed_weight_masks_dev(uint32_t n, uint32_t k, T* d_out, hipStream_t s) {
    const uint64_t count = binom_u64(n, k);
    constexpr uint64_t kPerBlock = (uint64_t)kCombThreads * comb_run<T>();
    constexpr uint64_t kMaxBlocks = 1u << 30;   // per launch
    for (uint64_t first = 0; first < count; first += kPerBlock * kMaxBlocks) {
        const uint64_t blocks = std::min<uint64_t>((count - first + kPerBlock - 1) / kPerBlock,
                                                   kMaxBlocks);
        hipLaunchKernelGGL(fixed_weight_kernel<T>, dim3((unsigned)blocks), dim3(kCombThreads),
                           (size_t)(n + 1) * (k + 1) * 8, s, n, k, first, count, d_out);
        BMM_HIP(hipGetLastError());
    }
}

template <typename T>
void fixed_weight_masks(uint32_t n, uint32_t k, T* out) {
    const uint64_t count = binom_u64(n, k);
    DevBuf buf;
    T* d = buf.get<T>((size_t)count);
    {
        ScopedKernelTimer timer(nullptr);
        fixed_weight_masks_dev<T>(n, k, d, nullptr);
    }
    BMM_HIP(hipMemcpy(out, d, (size_t)count * sizeof(T), hipMemcpyDeviceToHost));
}

template void fixed_weight_masks<uint32_t>(uint32_t, uint32_t, uint32_t*);
template void fixed_weight_masks<uint64_t>(uint32_t, uint32_t, uint64_t*);

} // namespace bmmhip
