// calc_scores_and_order (src/math/fwht.cpp:59-100) on the GPU: from a
// spectrum h of m entries and a column count n, the score of mask u is
// s[u] = clamp((n - h[u]) >> 1, 0, n) (the row-combination weight when h is
// the transform of the column-pattern histogram), s[0] = 0; cnt[v] counts
// the masks u = 1 .. m-1 of score v, and order lists those masks sorted by
// score, stably (ascending u within a score); off[v] ends as the end of
// bucket v in order (the reference's offsets after its scatter).
//
// The stable counting sort is an LSD radix sort over 8-bit digits of the
// score (ceil(bits(n) / 8) passes; one pass for n < 256).  Each pass: a
// digit histogram per tile of 4096 masks, an exclusive scan of the
// digit-major table (digit start + earlier tiles), and a scatter in which a
// tile ranks its masks in index order: for each of its 16 strided slices the
// waves match equal digits with 8 ballots, and per-wave digit counts in LDS
// give the offsets of earlier waves and slices.
#include "common.hpp"

#include <algorithm>
#include <array>
#include <type_traits>

namespace bmmhip {
namespace {

constexpr int kRsTile = 4096;   // masks per tile: 16 slices of 256
constexpr int kScanBlock = 1024;
constexpr int kLdsBuckets = 16384;

template <typename S>
__global__ __launch_bounds__(256) void scores_kernel(const S* __restrict__ h, uint64_t m, int32_t n,
                                                     S* __restrict__ s_out,
                                                     int32_t* __restrict__ keys,
                                                     unsigned int* __restrict__ cnt, int lds_hist,
                                                     unsigned int* __restrict__ overflow) {
    extern __shared__ unsigned int sh_cnt[];
    const int buckets = n + 1;
    if (lds_hist) {
        for (int q = threadIdx.x; q < buckets; q += blockDim.x)
            sh_cnt[q] = 0;
        __syncthreads();
    }
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < m;
         u += (uint64_t)gridDim.x * blockDim.x) {
        if (u == 0) {
            s_out[0] = S(0);
            continue;
        }
        // 64-bit difference: n - h[u] cannot overflow here.
        int64_t su = ((int64_t)n - (int64_t)h[u]) >> 1;
        su = su < 0 ? 0 : (su > n ? n : su);
        if (std::is_same<S, int16_t>::value && su > 32767)
            atomicOr(overflow, 1u);
        s_out[u] = S(su);
        keys[u - 1] = (int32_t)su;
        if (lds_hist)
            atomicAdd(&sh_cnt[su], 1u);
        else
            atomicAdd(&cnt[su], 1u);
    }
    if (lds_hist) {
        __syncthreads();
        for (int q = threadIdx.x; q < buckets; q += blockDim.x)
            if (sh_cnt[q])
                atomicAdd(&cnt[q], sh_cnt[q]);
    }
}

// hist[d * n_tiles + tile] = masks of the tile whose digit is d.
__global__ __launch_bounds__(256) void radix_hist_kernel(const int32_t* __restrict__ keys,
                                                         uint32_t N, int shift, uint32_t n_tiles,
                                                         uint32_t* __restrict__ hist) {
    __shared__ unsigned int c[256];
    c[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kRsTile + threadIdx.x;
    for (int k = 0; k < kRsTile / 256; ++k) {
        const uint64_t e = base + (uint64_t)k * 256;
        if (e < N)
            atomicAdd(&c[((uint32_t)keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * n_tiles + blockIdx.x] = c[threadIdx.x];
}

// Stable scatter of one digit pass; start = scanned hist.  vals_in == nullptr
// stands for the identity values e + 1 (mask u = element e + 1).
__global__ __launch_bounds__(256) void radix_scatter_kernel(
    const int32_t* __restrict__ keys_in, const int32_t* __restrict__ vals_in, uint32_t N,
    int shift, const uint32_t* __restrict__ start, uint32_t n_tiles,
    int32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out) {
    __shared__ uint32_t next[256];      // next output slot of each digit
    __shared__ uint32_t wcnt[4][256];   // this slice's digit counts per wave
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    next[tid] = start[(size_t)tid * n_tiles + blockIdx.x];
    wcnt[0][tid] = wcnt[1][tid] = wcnt[2][tid] = wcnt[3][tid] = 0;
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    for (int k = 0; k < kRsTile / 256; ++k) {
        const uint64_t e = (uint64_t)blockIdx.x * kRsTile + (uint64_t)k * 256 + tid;
        const bool valid = e < N;
        const int32_t key = valid ? keys_in[e] : 0;
        const uint32_t d = ((uint32_t)key >> shift) & 255u;
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            match &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(match & below);
        if (valid && rank == 0)
            wcnt[wave][d] = (uint32_t)__popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t pos = next[d] + rank;
            for (int q = 0; q < wave; ++q)
                pos += wcnt[q][d];
            if (keys_out)
                keys_out[pos] = key;
            vals_out[pos] = vals_in ? vals_in[e] : (int32_t)(e + 1);
        }
        __syncthreads();
        next[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        wcnt[0][tid] = wcnt[1][tid] = wcnt[2][tid] = wcnt[3][tid] = 0;
        __syncthreads();
    }
}

// Scan of blocks of 1024 entries in place (exclusive or inclusive);
// sums[b] = total of block b.
__global__ __launch_bounds__(256) void scan_block_kernel(uint32_t* __restrict__ data, uint64_t L,
                                                         uint32_t* __restrict__ sums,
                                                         int inclusive) {
    __shared__ uint32_t wsum[4];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint64_t b0 = (uint64_t)blockIdx.x * kScanBlock + threadIdx.x * 4u;
    uint32_t v[4], t = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = b0 + q < L ? data[b0 + q] : 0u;
        t += v[q];
    }
    uint32_t x = t;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, kWave);
        if (lane >= o)
            x += y;
    }
    if (lane == kWave - 1)
        wsum[wave] = x;
    __syncthreads();
    uint32_t run = x - t;
    for (int q = 0; q < wave; ++q)
        run += wsum[q];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t before = run;
        run += v[q];
        if (b0 + q < L)
            data[b0 + q] = inclusive ? run : before;
    }
    if (threadIdx.x == blockDim.x - 1)
        sums[blockIdx.x] = run;
}

__global__ __launch_bounds__(256) void scan_add_kernel(uint32_t* __restrict__ data, uint64_t L,
                                                       const uint32_t* __restrict__ sums) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < L && i >= (uint64_t)kScanBlock)
        data[i] += sums[i / kScanBlock];
}

struct ScoresWorkspace {
    DevBuf h, s, cnt, off, overflow, hist, order;
    DevBuf keys[2], vals[2];
    std::array<DevBuf, 6> levels;
};
ScoresWorkspace& scores_ws() {
    thread_local ScoresWorkspace ws;
    return ws;
}

// In-place scan of L entries on the device (totals below 2^32).
void scan_u32(uint32_t* d, uint64_t L, bool inclusive, std::array<DevBuf, 6>& levels, size_t depth,
              hipStream_t s) {
    const uint64_t nb = (L + kScanBlock - 1) / kScanBlock;
    if (depth >= levels.size())
        fail(BMM_ERR_OVERFLOW, "scan: input too long");
    uint32_t* sums = levels[depth].get<uint32_t>((size_t)nb);
    hipLaunchKernelGGL(scan_block_kernel, dim3((unsigned)nb), dim3(256), 0, s, d, L, sums,
                       inclusive ? 1 : 0);
    BMM_HIP(hipGetLastError());
    if (nb > 1) {
        scan_u32(sums, nb, false, levels, depth + 1, s);
        hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, s, d,
                           L, sums);
        BMM_HIP(hipGetLastError());
    }
}

} // namespace

template <typename S>
void scores_order(const S* h, size_t m, int32_t n, S* s_by_mask, int32_t* order, int32_t* cnt,
                  int32_t* off) {
    ScoresWorkspace& ws = scores_ws();
    const size_t buckets = (size_t)n + 1;
    const uint64_t N = m - 1;   // masks 1 .. m-1
    S* d_h = ws.h.get<S>(m);
    S* d_s = ws.s.get<S>(m);
    uint32_t* d_cnt = ws.cnt.get<uint32_t>(buckets);
    uint32_t* d_off = ws.off.get<uint32_t>(buckets);
    uint32_t* d_ovf = ws.overflow.get<uint32_t>(1);
    int32_t* keys[2] = {ws.keys[0].get<int32_t>(std::max<uint64_t>(N, 1)),
                        ws.keys[1].get<int32_t>(std::max<uint64_t>(N, 1))};
    BMM_HIP(hipMemcpy(d_h, h, m * sizeof(S), hipMemcpyHostToDevice));
    fill_u32(d_cnt, buckets, 0, nullptr);
    fill_u32(d_ovf, 1, 0, nullptr);
    const bool lds = buckets <= (size_t)kLdsBuckets;
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((m + 4095) / 4096, 1024));
    hipLaunchKernelGGL(scores_kernel<S>, dim3(grid), dim3(256), lds ? buckets * 4 : 0, nullptr, d_h,
                       (uint64_t)m, n, d_s, keys[0], d_cnt, lds ? 1 : 0, d_ovf);
    BMM_HIP(hipGetLastError());
    uint32_t ovf = 0;
    BMM_HIP(hipMemcpy(&ovf, d_ovf, 4, hipMemcpyDeviceToHost));
    if (ovf)
        fail(BMM_ERR_OVERFLOW, "calc_scores_and_order: a score exceeds the int16 range");
    if (N > 0) {
        ScopedKernelTimer timer(nullptr);
        const int bits = n == 0 ? 1 : 32 - __builtin_clz((uint32_t)n);
        const int passes = (bits + 7) / 8;
        const uint32_t n_tiles = (uint32_t)((N + kRsTile - 1) / kRsTile);
        uint32_t* d_hist = ws.hist.get<uint32_t>((size_t)256 * n_tiles);
        int32_t* d_order = ws.order.get<int32_t>((size_t)N);
        int32_t* vals[2] = {ws.vals[0].get<int32_t>(passes > 1 ? (size_t)N : 1),
                            ws.vals[1].get<int32_t>(passes > 2 ? (size_t)N : 1)};
        const int32_t* kin = keys[0];
        const int32_t* vin = nullptr;
        for (int p = 0; p < passes; ++p) {
            const bool last = p == passes - 1;
            int32_t* kout = last ? nullptr : keys[(p + 1) & 1];
            int32_t* vout = last ? d_order : vals[p & 1];
            hipLaunchKernelGGL(radix_hist_kernel, dim3(n_tiles), dim3(256), 0, nullptr, kin,
                               (uint32_t)N, 8 * p, n_tiles, d_hist);
            BMM_HIP(hipGetLastError());
            scan_u32(d_hist, (uint64_t)256 * n_tiles, false, ws.levels, 0, nullptr);
            hipLaunchKernelGGL(radix_scatter_kernel, dim3(n_tiles), dim3(256), 0, nullptr, kin, vin,
                               (uint32_t)N, 8 * p, d_hist, n_tiles, kout, vout);
            BMM_HIP(hipGetLastError());
            kin = kout;
            vin = vout;
        }
        BMM_HIP(hipMemcpy(order, d_order, N * 4, hipMemcpyDeviceToHost));
    }
    BMM_HIP(hipMemcpy(d_off, d_cnt, buckets * 4, hipMemcpyDeviceToDevice));
    scan_u32(d_off, buckets, true, ws.levels, 0, nullptr);
    BMM_HIP(hipMemcpy(s_by_mask, d_s, m * sizeof(S), hipMemcpyDeviceToHost));
    BMM_HIP(hipMemcpy(cnt, d_cnt, buckets * 4, hipMemcpyDeviceToHost));
    BMM_HIP(hipMemcpy(off, d_off, buckets * 4, hipMemcpyDeviceToHost));
}

template void scores_order<int32_t>(const int32_t*, size_t, int32_t, int32_t*, int32_t*, int32_t*,
                                    int32_t*);
template void scores_order<int16_t>(const int16_t*, size_t, int32_t, int16_t*, int32_t*, int32_t*,
                                    int32_t*);

} // namespace bmmhip
