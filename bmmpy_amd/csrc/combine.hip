// GF(2) product D = T * B where row i of T is given as a mask over B's rows:
// D[i] = XOR of B[j] over the set bits j of mask i.  This is BitMatrix::mul
// (src/core/bit_matrix_algebra.cpp:27-50) and the full transform that
// GlobalGreedyApplier applies (src/apply/global_greedy_applier.cpp:93-96).
//
// One thread per (output row block, column word): the thread walks the set
// bits of each mask of its row block (the masks are wave-uniform, so they are
// scalar loads) and XORs the selected words of B, which the rows of the block
// re-read from L2.  Total weight: per-thread popcounts, wave sums, one 64-bit
// atomic per wave.
#include "common.hpp"

#include <algorithm>

namespace bmmhip {
namespace {

constexpr int kCombineRows = 8;   // output rows per workgroup (grid.y)

__global__ __launch_bounds__(256) void row_combine_kernel(
    const uint64_t* __restrict__ B, int words, const uint64_t* __restrict__ masks, int mask_words,
    int row0, int n_out, uint64_t* __restrict__ D, unsigned long long* __restrict__ weight) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int i0 = row0 + (int)blockIdx.y * kCombineRows;
    const int i1 = min(i0 + kCombineRows, n_out);
    uint32_t pop = 0;
    if (w < words) {
        for (int i = i0; i < i1; ++i) {
            uint64_t acc = 0;
            const uint64_t* m = masks + (size_t)i * mask_words;
            for (int q = 0; q < mask_words; ++q)
                for (uint64_t bits = m[q]; bits; bits &= bits - 1)
                    acc ^= B[(size_t)(q * 64 + __builtin_ctzll(bits)) * words + w];
            D[(size_t)i * words + w] = acc;
            pop = bcnt_acc((uint32_t)(acc >> 32), bcnt_acc((uint32_t)acc, pop));
        }
    }
    pop = wave_sum_u32(pop);
    if (threadIdx.x % kWave == 0 && pop)
        atomicAdd(weight, (unsigned long long)pop);
}

struct CombineWorkspace {
    DevBuf rows, masks, out, weight;
};
CombineWorkspace& combine_ws() {
    thread_local CombineWorkspace ws;
    return ws;
}

} // namespace

// Device form: every pointer on the device; *d_weight is overwritten with
// the popcount of D (kernel-side accumulation after a fill on `stream`).
void row_combine_dev(const uint64_t* d_rows, size_t words, const uint64_t* d_masks,
                     size_t mask_words, size_t n_out, uint64_t* d_out, uint64_t* d_weight,
                     hipStream_t stream) {
    fill_u64(d_weight, 1, 0, stream);
    if (n_out == 0 || words == 0)
        return;
    // grid.y stays below 65536: row blocks go in launches of at most
    // 32768 x kCombineRows rows.
    constexpr size_t kRowsPerLaunch = (size_t)32768 * kCombineRows;
    for (size_t row0 = 0; row0 < n_out; row0 += kRowsPerLaunch) {
        const size_t rows = std::min(kRowsPerLaunch, n_out - row0);
        const dim3 grid((unsigned)((words + 255) / 256),
                        (unsigned)((rows + kCombineRows - 1) / kCombineRows));
        hipLaunchKernelGGL(row_combine_kernel, grid, dim3(256), 0, stream, d_rows, (int)words,
                           d_masks, (int)mask_words, (int)row0, (int)n_out, d_out,
                           reinterpret_cast<unsigned long long*>(d_weight));
        BMM_HIP(hipGetLastError());
    }
}

void row_combine(const uint64_t* rows, size_t n_rows, size_t words, const uint64_t* masks,
                 size_t mask_words, size_t n_out, uint64_t* out, uint64_t* out_weight) {
    CombineWorkspace& ws = combine_ws();
    uint64_t* d_rows = ws.rows.get<uint64_t>(std::max<size_t>(n_rows * words, 1));
    uint64_t* d_masks = ws.masks.get<uint64_t>(std::max<size_t>(n_out * mask_words, 1));
    uint64_t* d_out = ws.out.get<uint64_t>(std::max<size_t>(n_out * words, 1));
    uint64_t* d_w = ws.weight.get<uint64_t>(1);
    if (n_rows * words)
        BMM_HIP(hipMemcpy(d_rows, rows, n_rows * words * 8, hipMemcpyHostToDevice));
    if (n_out * mask_words)
        BMM_HIP(hipMemcpy(d_masks, masks, n_out * mask_words * 8, hipMemcpyHostToDevice));
    {
        ScopedKernelTimer timer(nullptr);
        row_combine_dev(d_rows, words, d_masks, mask_words, n_out, d_out, d_w, nullptr);
    }
    if (n_out * words)
        BMM_HIP(hipMemcpy(out, d_out, n_out * words * 8, hipMemcpyDeviceToHost));
    BMM_HIP(hipMemcpy(out_weight, d_w, 8, hipMemcpyDeviceToHost));
}

} // namespace bmmhip
