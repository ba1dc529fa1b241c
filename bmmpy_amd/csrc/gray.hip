// Gray-code coset sweep on gfx950.
//
// Work: for a target row t and a block of b rows (local bit k = block row k),
// visit every block subset in reflected-Gray order and evaluate
// wt(row_t ^ XOR(subset)).  Step s visits subset gray(s) = s ^ (s >> 1); one
// row XOR + one popcount per step (the reference brute-force inner loop,
// src/search/bruteforce_search.cpp:157-169, with the target row folded into the
// starting state).
//
// Mapping (lane per high state): a lane owns one high state h (bits c..b-1 of
// the subset) and sweeps its c low bits itself, holding the whole current row
// in VGPRs, so no cross-lane popcount reduction is needed per step.  A wave is
// 64 consecutive h of one target, a 256-thread workgroup 256 of them.  The
// lowest RES block rows stay resident in VGPRs (they serve 2^RES - 1 of every
// 2^RES steps); the remaining steps read their row from LDS.
//
// Inside lane h, step j (0 <= j < 2^c) visits low subset gray(j); its global
// Gray step is s = gray^-1(h) << c | (j ^ (parity(h) ? 2^c-1 : 0)), and its
// subset is h << c | gray(j).  Each lane keeps two packed u32 minima,
// (w << c | s_low) and (w << c | gray(j)); the workgroup reduces them and
// folds them into per-target u64 keys with one atomicMin each.
#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace bmmhip {

constexpr int kGrayMaxBlock = 40;

struct GrayTask {
    int32_t target;
    int32_t nbits;      // b, the block size (state bits)
    int32_t first_wg;   // first workgroup of this task
    int32_t slot;       // output slot
    int32_t blk[kGrayMaxBlock];
};

template <int NW, int RES>
__global__ __launch_bounds__(256) void gray_coset_kernel(const uint32_t* __restrict__ rows,
                                                         const GrayTask* __restrict__ tasks,
                                                         const int32_t* __restrict__ wg_task,
                                                         int c, uint64_t* __restrict__ first_key,
                                                         uint64_t* __restrict__ mask_key) {
    __shared__ uint32_t srow[kGrayMaxBlock + 1][NW];
    __shared__ uint64_t red1[4], red2[4];

    const int ti = __builtin_amdgcn_readfirstlane(wg_task[blockIdx.x]);
    const GrayTask& task = tasks[ti];
    const int b = task.nbits;
    const int hbits = b - c;

    // Stage the block rows (and the target row, at index b) into LDS.
    for (int i = threadIdx.x; i < (b + 1) * NW; i += blockDim.x) {
        const int r = i / NW, w = i % NW;
        const int src = r < b ? task.blk[r] : task.target;
        srow[r][w] = rows[(size_t)src * NW + w];
    }
    __syncthreads();

    const uint32_t h = (uint32_t)(blockIdx.x - task.first_wg) * 256u + threadIdx.x;

    uint32_t cur[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w)
        cur[w] = srow[b][w];
    // High bits 0..7 of h vary inside the workgroup; the rest are uniform.
    for (int k = 0; k < hbits; ++k) {
        const uint32_t* row = srow[c + k];
        if (k < 8) {
            const uint32_t m = 0u - ((h >> k) & 1u);
#pragma unroll
            for (int w = 0; w < NW; ++w)
                cur[w] ^= row[w] & m;
        } else if ((h >> k) & 1u) {
#pragma unroll
            for (int w = 0; w < NW; ++w)
                cur[w] ^= row[w];
        }
    }

    uint32_t res[RES][NW];
#pragma unroll
    for (int r = 0; r < RES; ++r)
#pragma unroll
        for (int w = 0; w < NW; ++w)
            res[r][w] = srow[r][w];

    const uint32_t lowmask = (1u << c) - 1u;
    const uint32_t flip = (__builtin_popcount(h) & 1) ? lowmask : 0u;
    uint32_t best1 = 0xffffffffu, best2 = 0xffffffffu;

    // Two accumulate chains of v_bcnt_u32_b32 (popcount + add in one op); the
    // compiler would otherwise split each popcount from its add.
    auto weight = [&]() {
        uint32_t a0 = 0, a1 = 0;
#pragma unroll
        for (int w = 0; w < NW; w += 2) {
            a0 = bcnt_acc(cur[w], a0);
            if (w + 1 < NW)
                a1 = bcnt_acc(cur[w + 1], a1);
        }
        return a0 + a1;
    };

    {   // j = 0: low subset 0
        const uint32_t ws = weight() << c;
        best1 = min(best1, ws | flip);
        best2 = min(best2, ws);
    }

    constexpr uint32_t kUnroll = 1u << RES;
    const uint32_t iters = 1u << (c - RES);
    for (uint32_t a = 0; a < iters; ++a) {
        const uint32_t jb = a << RES;
        const uint32_t base1 = jb ^ flip;
        const uint32_t gbase = jb ^ (jb >> 1);
#pragma unroll
        for (uint32_t i = 1; i < kUnroll; ++i) {
            const int bit = __builtin_ctz(i);
#pragma unroll
            for (int w = 0; w < NW; ++w)
                cur[w] ^= res[bit][w];
            const uint32_t ws = weight() << c;
            best1 = min(best1, ws | (base1 ^ i));
            best2 = min(best2, ws | (gbase ^ (i ^ (i >> 1))));
        }
        const uint32_t j = jb + kUnroll;
        if (j <= lowmask) {
            const int bit = __builtin_ctz(j);
            const uint32_t* row = srow[bit];
#pragma unroll
            for (int w = 0; w < NW; ++w)
                cur[w] ^= row[w];
            const uint32_t ws = weight() << c;
            best1 = min(best1, ws | (j ^ flip));
            best2 = min(best2, ws | (j ^ (j >> 1)));
        }
    }

    // Expand lane minima to global keys and reduce over the workgroup.
    const uint64_t hs = (uint64_t)gray_inverse(h) << c;
    const uint64_t k1 = ((uint64_t)(best1 >> c) << 32) | (hs | (best1 & lowmask));
    const uint64_t k2 = ((uint64_t)(best2 >> c) << 32) | (((uint64_t)h << c) | (best2 & lowmask));
    const uint64_t m1 = wave_min_u64(k1);
    const uint64_t m2 = wave_min_u64(k2);
    const int wv = threadIdx.x / kWave;
    if ((threadIdx.x % kWave) == 0) {
        red1[wv] = m1;
        red2[wv] = m2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a1 = red1[0], a2 = red2[0];
        for (int i = 1; i < 4; ++i) {
            a1 = min(a1, red1[i]);
            a2 = min(a2, red2[i]);
        }
        atomicMin((unsigned long long*)&first_key[task.slot], (unsigned long long)a1);
        atomicMin((unsigned long long*)&mask_key[task.slot], (unsigned long long)a2);
    }
}

namespace {

struct GrayWorkspace {
    DevBuf tasks, wg_task, rows32, keys;
    std::vector<GrayTask> plan_tasks;
    std::vector<int32_t> plan_wg;
    const GrayTask* plan_dev = nullptr;
    const int32_t* plan_wg_dev = nullptr;
};

GrayWorkspace& gray_ws() {
    thread_local GrayWorkspace ws;
    return ws;
}

template <int NW, int RES>
void launch_gray(const uint32_t* rows, const GrayTask* d_tasks, const int32_t* d_wg_task,
                 int n_wg, int c, uint64_t* k1, uint64_t* k2, hipStream_t s) {
    ScopedKernelTimer timer(s);
    hipLaunchKernelGGL((gray_coset_kernel<NW, RES>), dim3(n_wg), dim3(256), 0, s, rows, d_tasks,
                       d_wg_task, c, k1, k2);
    BMM_HIP(hipGetLastError());
}

} // namespace

// Chooses the per-lane chunk: enough lanes to fill 256 CUs several times over,
// and at least 8 high bits so a task covers whole 256-lane workgroups.
int gray_auto_chunk(int min_block_bits) {
    int c = 9;
    if (min_block_bits - c < 8)
        c = min_block_bits - 8;
    return c;
}

// Argument checks shared by the host and device entry points (run before any
// device call).
void gray_check_args(size_t n_rows, size_t words_per_row, size_t block_rows,
                     const int32_t* targets, size_t n_targets, size_t chunk_bits) {
    if (words_per_row == 0 || words_per_row > 32)
        fail(BMM_ERR_INVALID, "gray_coset_min: words_per_row must be in [1, 32]");
    if (block_rows > n_rows || block_rows > 64)
        fail(BMM_ERR_INVALID, "gray_coset_min: block_rows must be <= n_rows");
    int min_bits = 64;
    for (size_t i = 0; i < n_targets; ++i) {
        if (targets[i] < 0 || (size_t)targets[i] >= n_rows)
            fail(BMM_ERR_RANGE, "gray_coset_min: target row out of range");
        const int b = (int)block_rows - ((size_t)targets[i] < block_rows ? 1 : 0);
        if (b > kGrayMaxBlock)
            fail(BMM_ERR_INVALID, "gray_coset_min: block larger than 40 rows");
        min_bits = std::min(min_bits, b);
    }
    if (n_targets == 0)
        return;
    const int c = chunk_bits ? (int)chunk_bits : gray_auto_chunk(min_bits);
    if (min_bits - c < 8 || c < 4 || c > 20)
        fail(BMM_ERR_INVALID,
             "gray_coset_min: block must have >= chunk_bits + 8 rows and chunk_bits in [4, 20]");
}

void gray_coset_min_dev(const uint64_t* d_rows, size_t n_rows, size_t words_per_row,
                        size_t block_rows, const int32_t* targets, size_t n_targets,
                        size_t chunk_bits, uint64_t* d_first_key, uint64_t* d_mask_key,
                        hipStream_t stream) {
    gray_check_args(n_rows, words_per_row, block_rows, targets, n_targets, chunk_bits);
    if (n_targets == 0)
        return;
    std::vector<GrayTask> tasks(n_targets);
    int min_bits = 64;
    for (size_t i = 0; i < n_targets; ++i) {
        const int t = targets[i];
        if (t < 0 || (size_t)t >= n_rows)
            fail(BMM_ERR_RANGE, "gray_coset_min: target row out of range");
        GrayTask& g = tasks[i];
        std::memset(&g, 0, sizeof(g));
        g.target = t;
        g.slot = (int)i;
        int b = 0;
        for (size_t r = 0; r < block_rows; ++r)
            if ((int)r != t)
                g.blk[b++] = (int)r;
        if (b > kGrayMaxBlock)
            fail(BMM_ERR_INVALID, "gray_coset_min: block larger than 40 rows");
        g.nbits = b;
        min_bits = std::min(min_bits, b);
    }
    int c = chunk_bits ? (int)chunk_bits : gray_auto_chunk(min_bits);
    if (min_bits - c < 8 || c < 4 || c > 20)
        fail(BMM_ERR_INVALID,
             "gray_coset_min: block must have >= chunk_bits + 8 rows and chunk_bits in [4, 20]");
    std::vector<int32_t> wg_task;
    for (size_t i = 0; i < n_targets; ++i) {
        tasks[i].first_wg = (int32_t)wg_task.size();
        const size_t n_wg = (size_t)1 << (tasks[i].nbits - c - 8);
        if (wg_task.size() + n_wg > (size_t)INT32_MAX)
            fail(BMM_ERR_INVALID, "gray_coset_min: too many workgroups");
        wg_task.insert(wg_task.end(), n_wg, (int32_t)i);
    }
    GrayWorkspace& ws = gray_ws();
    GrayTask* d_tasks = ws.tasks.get<GrayTask>(tasks.size());
    int32_t* d_wg = ws.wg_task.get<int32_t>(wg_task.size());
    // Task tables are re-uploaded only when the plan changes (repeated calls
    // on the same shape enqueue nothing but the fills and the sweep).
    const bool same = ws.plan_dev == d_tasks && ws.plan_wg_dev == d_wg && ws.plan_tasks.size() == tasks.size() &&
                      std::memcmp(ws.plan_tasks.data(), tasks.data(),
                                  sizeof(GrayTask) * tasks.size()) == 0 &&
                      ws.plan_wg == wg_task;
    if (!same) {
        BMM_HIP(hipMemcpyAsync(d_tasks, tasks.data(), sizeof(GrayTask) * tasks.size(),
                               hipMemcpyHostToDevice, stream));
        BMM_HIP(hipMemcpyAsync(d_wg, wg_task.data(), sizeof(int32_t) * wg_task.size(),
                               hipMemcpyHostToDevice, stream));
        BMM_HIP(hipStreamSynchronize(stream));
        ws.plan_tasks = tasks;
        ws.plan_wg = wg_task;
        ws.plan_dev = d_tasks;
        ws.plan_wg_dev = d_wg;
    }
    fill_u64(d_first_key, n_targets, ~0ull, stream);
    fill_u64(d_mask_key, n_targets, ~0ull, stream);
    // Rows are read as u32 words; the u64 layout is the same bytes.
    const uint32_t* rows32 = reinterpret_cast<const uint32_t*>(d_rows);
    const int n_wg = (int)wg_task.size();
    switch (words_per_row) {
    case 1: launch_gray<2, 1>(rows32, d_tasks, d_wg, n_wg, c, d_first_key, d_mask_key, stream); break;
    case 2: launch_gray<4, 2>(rows32, d_tasks, d_wg, n_wg, c, d_first_key, d_mask_key, stream); break;
    case 4: launch_gray<8, 3>(rows32, d_tasks, d_wg, n_wg, c, d_first_key, d_mask_key, stream); break;
    case 8: launch_gray<16, 3>(rows32, d_tasks, d_wg, n_wg, c, d_first_key, d_mask_key, stream); break;
    case 16: launch_gray<32, 3>(rows32, d_tasks, d_wg, n_wg, c, d_first_key, d_mask_key, stream); break;
    case 32: launch_gray<64, 1>(rows32, d_tasks, d_wg, n_wg, c, d_first_key, d_mask_key, stream); break;
    default: {
        // Other widths: zero-pad rows to the next supported width.
        size_t pw = 1;
        while (pw < words_per_row)
            pw <<= 1;
        uint64_t* padded = ws.rows32.get<uint64_t>(n_rows * pw);
        BMM_HIP(hipMemsetAsync(padded, 0, n_rows * pw * sizeof(uint64_t), stream));
        BMM_HIP(hipMemcpy2DAsync(padded, pw * sizeof(uint64_t), d_rows,
                                 words_per_row * sizeof(uint64_t),
                                 words_per_row * sizeof(uint64_t), n_rows,
                                 hipMemcpyDeviceToDevice, stream));
        gray_coset_min_dev(padded, n_rows, pw, block_rows, targets, n_targets, (size_t)c,
                           d_first_key, d_mask_key, stream);
        return;
    }
    }
}

} // namespace bmmhip
