// Pairwise row overlaps of a matrix: out[i][j] = popcount(r_i & r_j) for
// i != j, 0 on the diagonal (the pairwise synergy table of SASelector,
// src/search/sa_selector_pairwise.cpp:21-52, before its factor 2).
//
// One workgroup per 64 x 64 tile of row pairs above or on the diagonal (the
// table is symmetric).  Word chunks of both 64-row blocks are staged in LDS
// word-major, so thread (ty, tx) reads rows ty + 16 i and tx + 16 j for its
// 4 x 4 micro-tile: the 16 lanes that differ in tx read 16 consecutive words
// (no bank conflicts) and those that share ty read one word (broadcast).  Two
// v_bcnt_u32_b32 per 64-bit word.  The mirrored half of an off-diagonal tile
// is transposed through LDS so both halves are written in rows.
// Work per pair: 4 int32 VALU ops per 64-bit word.
#include "common.hpp"

#include <algorithm>

namespace bmmhip {
namespace {

constexpr int kOvTile = 64;
constexpr int kOvChunk = 32;   // u64 words per LDS stage

__global__ __launch_bounds__(256) void overlap_kernel(const uint64_t* __restrict__ rows, int n,
                                                      int words, int n_blocks,
                                                      uint32_t* __restrict__ out) {
    // Upper-triangle tile (bi <= bj) from the linear block index.
    int bi = 0, rem = (int)blockIdx.x;
    while (rem >= n_blocks - bi) {
        rem -= n_blocks - bi;
        ++bi;
    }
    const int bj = bi + rem;
    __shared__ uint64_t A[kOvChunk][kOvTile + 1];
    __shared__ uint64_t B[kOvChunk][kOvTile + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    uint32_t acc[4][4] = {};
    for (int w0 = 0; w0 < words; w0 += kOvChunk) {
        const int cw = min(kOvChunk, words - w0);
        for (int q = threadIdx.x; q < kOvTile * kOvChunk; q += blockDim.x) {
            const int r = q / kOvChunk, w = q % kOvChunk;
            const int ra = bi * kOvTile + r, rb = bj * kOvTile + r;
            A[w][r] = (w < cw && ra < n) ? rows[(size_t)ra * words + w0 + w] : 0ull;
            B[w][r] = (w < cw && rb < n) ? rows[(size_t)rb * words + w0 + w] : 0ull;
        }
        __syncthreads();
#pragma unroll 4
        for (int w = 0; w < cw; ++w) {
            uint64_t a[4], b[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                a[k] = A[w][ty + 16 * k];
                b[k] = B[w][tx + 16 * k];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint64_t x = a[i] & b[j];
                    acc[i][j] = bcnt_acc((uint32_t)(x >> 32), bcnt_acc((uint32_t)x, acc[i][j]));
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = bi * kOvTile + ty + 16 * i, c = bj * kOvTile + tx + 16 * j;
            if (r < n && c < n)
                out[(size_t)r * n + c] = r == c ? 0u : acc[i][j];
        }
    if (bi == bj)
        return;
    // Mirrored half: T[c][r] in the (now free) staging memory, then rows c.
    uint32_t(*T)[kOvTile + 1] = reinterpret_cast<uint32_t(*)[kOvTile + 1]>(&A[0][0]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            T[tx + 16 * j][ty + 16 * i] = acc[i][j];
    __syncthreads();
    for (int q = threadIdx.x; q < kOvTile * kOvTile; q += blockDim.x) {
        const int cl = q / kOvTile, rl = q % kOvTile;
        const int c = bj * kOvTile + cl, r = bi * kOvTile + rl;
        if (c < n && r < n)
            out[(size_t)c * n + r] = T[cl][rl];
    }
}

struct OverlapWorkspace {
    DevBuf rows, out;
};
OverlapWorkspace& overlap_ws() {
    thread_local OverlapWorkspace ws;
    return ws;
}

} // namespace

void pair_overlaps(const uint64_t* rows, size_t n, size_t words, uint32_t* out) {
    if (n == 0)
        return;
    OverlapWorkspace& ws = overlap_ws();
    uint64_t* d_rows = ws.rows.get<uint64_t>(std::max<size_t>(n * words, 1));
    uint32_t* d_out = ws.out.get<uint32_t>(n * n);
    if (words)
        BMM_HIP(hipMemcpy(d_rows, rows, n * words * 8, hipMemcpyHostToDevice));
    const int nb = (int)((n + kOvTile - 1) / kOvTile);
    const unsigned tiles = (unsigned)((size_t)nb * (nb + 1) / 2);
    {
        ScopedKernelTimer timer(nullptr);
        hipLaunchKernelGGL(overlap_kernel, dim3(tiles), dim3(256), 0, nullptr, d_rows, (int)n,
                           (int)words, nb, d_out);
    }
    BMM_HIP(hipGetLastError());
    BMM_HIP(hipMemcpy(out, d_out, n * n * 4, hipMemcpyDeviceToHost));
}

} // namespace bmmhip
