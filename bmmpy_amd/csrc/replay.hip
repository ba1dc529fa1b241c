// Row-XOR sequence replay on one matrix (src/apply: the XOR replay of
// GreedyApplier::apply, src/apply/greedy_applier.cpp:92-95, through
// BitMatrix::row_xor, src/core/bit_matrix_rows.cpp:12-20).
//
// Every word column is independent under row XORs, so one thread owns one
// 64-bit column slice and applies the whole op list to it in order; adjacent
// threads touch adjacent words of the same rows (coalesced).  The final
// popcount of the owned words gives the matrix weight with one atomic per
// wavefront.
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace bmmhip {

__global__ __launch_bounds__(256) void replay_columns_kernel(uint64_t* __restrict__ rows,
                                                             int n_rows, int words,
                                                             const uint16_t* __restrict__ ops,
                                                             int n_ops,
                                                             unsigned long long* __restrict__ wt) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t pc = 0;
    if (w < words) {
        for (int i = 0; i < n_ops; ++i) {
            const int src = ops[2 * i], dst = ops[2 * i + 1];
            if (src != dst)
                rows[(size_t)dst * words + w] ^= rows[(size_t)src * words + w];
        }
        for (int r = 0; r < n_rows; ++r)
            pc += __builtin_popcountll(rows[(size_t)r * words + w]);
    }
    pc = wave_sum_u32(pc);
    if ((threadIdx.x % kWave) == 0 && pc)
        atomicAdd(wt, (unsigned long long)pc);
}

void replay_dev(uint64_t* d_rows, int n_rows, int words, const uint16_t* d_ops, int n_ops,
                unsigned long long* d_wt, hipStream_t s) {
    fill_u64(reinterpret_cast<uint64_t*>(d_wt), 1, 0, s);
    if (words == 0)
        return;
    hipLaunchKernelGGL(replay_columns_kernel, dim3((words + 255) / 256), dim3(256), 0, s, d_rows,
                       n_rows, words, d_ops, n_ops, d_wt);
    BMM_HIP(hipGetLastError());
}

namespace {
struct ReplayWorkspace {
    DevBuf rows, ops, wt;
};
ReplayWorkspace& replay_ws() {
    thread_local ReplayWorkspace ws;
    return ws;
}
} // namespace

// Host rows in, replayed rows + weight out (ops as (source, target) pairs).
void replay_host(uint64_t* rows, size_t n_rows, size_t words, const uint16_t* ops, size_t n_ops,
                 uint64_t* weight) {
    *weight = 0;
    if (n_rows == 0 || words == 0)
        return;
    ReplayWorkspace& ws = replay_ws();
    const size_t nw = n_rows * words;
    uint64_t* d_rows = ws.rows.get<uint64_t>(nw);
    uint16_t* d_ops = ws.ops.get<uint16_t>(std::max<size_t>(2 * n_ops, 1));
    unsigned long long* d_wt = ws.wt.get<unsigned long long>(1);
    BMM_HIP(hipMemcpy(d_rows, rows, nw * 8, hipMemcpyHostToDevice));
    if (n_ops)
        BMM_HIP(hipMemcpy(d_ops, ops, 2 * n_ops * sizeof(uint16_t), hipMemcpyHostToDevice));
    replay_dev(d_rows, (int)n_rows, (int)words, d_ops, (int)n_ops, d_wt, nullptr);
    BMM_HIP(hipMemcpy(rows, d_rows, nw * 8, hipMemcpyDeviceToHost));
    unsigned long long wt = 0;
    BMM_HIP(hipMemcpy(&wt, d_wt, 8, hipMemcpyDeviceToHost));
    *weight = wt;
}

} // namespace bmmhip
