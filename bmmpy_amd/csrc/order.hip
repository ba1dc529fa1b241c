// Weight order of every nonzero row combination of a window (the candidate
// stream of ExactBasisSolver, src/solver/exact_basis_solver.cpp:29-54 and
// :105-124): the reference visits masks gray(i) for Gray indices
// i = 1 .. 2^t - 1 and counting-sorts them by weight, stably, so the stream is
// ordered by (weight, Gray index).
//
// Here all 2^t weights come from one transform: H = FWHT of the
// column-pattern histogram in HBM (N > 2^15: the two-pass array FWHT of
// fwht.hip), weight(m) = (total - H[m]) / 2.  The first k of the stream are
// then found without a full sort: a per-workgroup weight histogram gives the
// smallest weight w* whose cumulative count reaches k, an order-preserving
// compaction writes every Gray index with weight <= w* (tile counts, host
// prefix, in-tile ranks), and the host sorts that short list stably by weight.
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace bmmhip {

template <typename T> void fwht_array_dev(T* d, size_t n, hipStream_t stream);

namespace {

constexpr int kOrderBuckets = 8192;   // weights >= kOrderBuckets - 1 share the last bucket
constexpr int kTile = 4096;           // Gray indices per compaction tile (16 per thread)

// hist[pattern] += 1 per active column (bit r of the pattern = row r).
__global__ void order_hist_kernel(const uint64_t* __restrict__ rows, int t, int words, int cols,
                                  int32_t* __restrict__ hist, unsigned int* __restrict__ total) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t p = 0;
    if (c < cols) {
        const int wi = c >> 6, bit = c & 63;
        for (int r = 0; r < t; ++r)
            p |= (uint32_t)((rows[(size_t)r * words + wi] >> bit) & 1ull) << r;
        if (p != 0u)
            atomicAdd(&hist[p], 1);
    }
    const uint32_t n = wave_sum_u32(p != 0u ? 1u : 0u);
    if (threadIdx.x % kWave == 0 && n)
        atomicAdd(total, n);
}

__device__ inline uint32_t order_weight(const int32_t* __restrict__ H, uint32_t total,
                                        uint64_t i) {
    return (uint32_t)((int64_t)total - (int64_t)H[i ^ (i >> 1)]) >> 1;
}

// Weight histogram over Gray indices 1 .. n-1 (per-workgroup LDS buckets).
__global__ __launch_bounds__(256) void order_count_kernel(const int32_t* __restrict__ H,
                                                          uint64_t n, uint32_t total,
                                                          unsigned int* __restrict__ counts) {
    __shared__ unsigned int b[kOrderBuckets];
    for (int q = threadIdx.x; q < kOrderBuckets; q += blockDim.x)
        b[q] = 0;
    __syncthreads();
    for (uint64_t i = 1 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&b[min(order_weight(H, total, i), (uint32_t)kOrderBuckets - 1u)], 1u);
    __syncthreads();
    for (int q = threadIdx.x; q < kOrderBuckets; q += blockDim.x)
        if (b[q])
            atomicAdd(&counts[q], b[q]);
}

// Selected (weight <= wmax) Gray indices per tile of kTile indices; thread t
// of a tile owns indices [16 t, 16 t + 16) of it.
__global__ __launch_bounds__(256) void order_tile_count_kernel(const int32_t* __restrict__ H,
                                                               uint64_t n, uint32_t total,
                                                               uint32_t wmax,
                                                               unsigned int* __restrict__ tiles) {
    const uint64_t base = (uint64_t)blockIdx.x * kTile + 16u * threadIdx.x;
    uint32_t c = 0;
    for (int k = 0; k < 16; ++k) {
        const uint64_t i = base + k;
        if (i >= 1 && i < n && order_weight(H, total, i) <= wmax)
            ++c;
    }
    c = wave_sum_u32(c);
    __shared__ unsigned int red[4];
    if (threadIdx.x % kWave == 0)
        red[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0)
        tiles[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Writes the selected Gray indices and weights at tile_off[tile] + rank within
// the tile, in index order.
__global__ __launch_bounds__(256) void order_compact_kernel(
    const int32_t* __restrict__ H, uint64_t n, uint32_t total, uint32_t wmax,
    const uint64_t* __restrict__ tile_off, uint32_t* __restrict__ out_idx,
    uint32_t* __restrict__ out_w) {
    __shared__ unsigned int scan[256];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + 16u * threadIdx.x;
    uint32_t w[16];
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint64_t i = base + k;
        w[k] = (i >= 1 && i < n) ? order_weight(H, total, i) : 0xffffffffu;
        c += w[k] <= wmax;
    }
    scan[threadIdx.x] = c;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {   // inclusive scan (Hillis-Steele)
        const unsigned int v = threadIdx.x >= (unsigned)d ? scan[threadIdx.x - d] : 0u;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t pos = tile_off[blockIdx.x] + scan[threadIdx.x] - c;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        if (w[k] <= wmax) {
            out_idx[pos] = (uint32_t)(base + k);
            out_w[pos] = w[k];
            ++pos;
        }
}

struct OrderWorkspace {
    DevBuf rows, hist, total, counts, tiles, tile_off, idx, w;
};
OrderWorkspace& order_ws() {
    thread_local OrderWorkspace ws;
    return ws;
}

} // namespace

// The first min(k, 2^t - 1) nonzero masks of a t-row window in the order
// (weight, Gray index); masks are gray(i).  Returns how many were written.
size_t weight_order(const uint64_t* rows, size_t t, size_t words, size_t cols, size_t k,
                    uint32_t* out_mask, uint32_t* out_weight) {
    const uint64_t n = uint64_t(1) << t;   // Gray indices 0 .. n-1 (0 excluded)
    const uint64_t states = n - 1;
    k = (size_t)std::min<uint64_t>(k, states);
    if (k == 0)
        return 0;
    OrderWorkspace& ws = order_ws();
    uint64_t* d_rows = ws.rows.get<uint64_t>(std::max<size_t>(t * words, 1));
    int32_t* d_h = ws.hist.get<int32_t>((size_t)n);
    unsigned int* d_total = ws.total.get<unsigned int>(1);
    unsigned int* d_counts = ws.counts.get<unsigned int>(kOrderBuckets);
    BMM_HIP(hipMemcpy(d_rows, rows, t * words * 8, hipMemcpyHostToDevice));
    BMM_HIP(hipMemsetAsync(d_h, 0, (size_t)n * 4, nullptr));
    BMM_HIP(hipMemsetAsync(d_total, 0, 4, nullptr));
    BMM_HIP(hipMemsetAsync(d_counts, 0, kOrderBuckets * 4, nullptr));
    if (cols > 0)
        hipLaunchKernelGGL(order_hist_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0,
                           nullptr, d_rows, (int)t, (int)words, (int)cols, d_h, d_total);
    BMM_HIP(hipGetLastError());
    if (t > 0) {
        ScopedKernelTimer timer(nullptr);
        fwht_array_dev<int32_t>(d_h, (size_t)n, nullptr);
    }
    unsigned int total = 0;
    BMM_HIP(hipMemcpy(&total, d_total, 4, hipMemcpyDeviceToHost));
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(order_count_kernel, dim3(grid), dim3(256), 0, nullptr, d_h, n, total,
                       d_counts);
    BMM_HIP(hipGetLastError());
    std::vector<unsigned int> counts(kOrderBuckets);
    BMM_HIP(hipMemcpy(counts.data(), d_counts, kOrderBuckets * 4, hipMemcpyDeviceToHost));
    // Smallest w* with cumulative count >= k (the last bucket is open-ended).
    uint64_t cum = 0;
    uint32_t wmax = 0xfffffffeu;
    for (int q = 0; q < kOrderBuckets - 1; ++q) {
        cum += counts[q];
        if (cum >= k) {
            wmax = (uint32_t)q;
            break;
        }
    }
    const uint64_t n_tiles = (n + kTile - 1) / kTile;
    unsigned int* d_tiles = ws.tiles.get<unsigned int>((size_t)n_tiles);
    hipLaunchKernelGGL(order_tile_count_kernel, dim3((unsigned)n_tiles), dim3(256), 0, nullptr,
                       d_h, n, total, wmax, d_tiles);
    BMM_HIP(hipGetLastError());
    std::vector<unsigned int> tiles((size_t)n_tiles);
    BMM_HIP(hipMemcpy(tiles.data(), d_tiles, n_tiles * 4, hipMemcpyDeviceToHost));
    std::vector<uint64_t> off((size_t)n_tiles);
    uint64_t sel = 0;
    for (size_t q = 0; q < n_tiles; ++q) {
        off[q] = sel;
        sel += tiles[q];
    }
    uint64_t* d_off = ws.tile_off.get<uint64_t>((size_t)n_tiles);
    uint32_t* d_idx = ws.idx.get<uint32_t>((size_t)std::max<uint64_t>(sel, 1));
    uint32_t* d_w = ws.w.get<uint32_t>((size_t)std::max<uint64_t>(sel, 1));
    BMM_HIP(hipMemcpy(d_off, off.data(), n_tiles * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(order_compact_kernel, dim3((unsigned)n_tiles), dim3(256), 0, nullptr, d_h,
                       n, total, wmax, d_off, d_idx, d_w);
    BMM_HIP(hipGetLastError());
    std::vector<uint32_t> idx((size_t)sel), w((size_t)sel);
    BMM_HIP(hipMemcpy(idx.data(), d_idx, sel * 4, hipMemcpyDeviceToHost));
    BMM_HIP(hipMemcpy(w.data(), d_w, sel * 4, hipMemcpyDeviceToHost));
    // Stable by weight: the selection is already in Gray-index order.
    std::vector<uint32_t> order((size_t)sel);
    for (size_t q = 0; q < sel; ++q)
        order[q] = (uint32_t)q;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return w[a] < w[b]; });
    for (size_t q = 0; q < k; ++q) {
        const uint32_t i = idx[order[q]];
        out_mask[q] = i ^ (i >> 1);
        out_weight[q] = w[order[q]];
    }
    return k;
}

} // namespace bmmhip
