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
#include <mutex>
#include <string>
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
                 int n_wg, int c, uint64_t* k1, uint64_t* k2, hipStream_t s, bool timed = true) {
    ScopedKernelTimer timer(s, timed);
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

// ------------------------------------------------ allocation-free form ----
// The same sweep for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops): the task tables live in `scratch`,
// a file-scope device array each GPU holds until the process ends, and are
// built on the device by gray_setup_kernel from by-value kernel arguments, so
// a captured launch sequence refers to no host memory.  Layout: GrayTask[64]
// at 0, the workgroup -> task table at 16 KiB (61,440 entries), zero-padded
// rows at 256 KiB (98,304 words).  Calls that overlap on one GPU share the
// array and race.
constexpr size_t kGrayScratchBytes = (size_t)1 << 20;
constexpr size_t kGrayScratchWgOff = (size_t)16 << 10;
constexpr size_t kGrayScratchRowsOff = (size_t)256 << 10;
constexpr int kGrayNoAllocTargets = 64;
constexpr size_t kGrayNoAllocWg = (kGrayScratchRowsOff - kGrayScratchWgOff) / 4;
constexpr size_t kGrayNoAllocRowWords = (kGrayScratchBytes - kGrayScratchRowsOff) / 8;
static_assert(kGrayNoAllocTargets * sizeof(GrayTask) <= kGrayScratchWgOff, "task table fits");

static __device__ __align__(256) unsigned char scratch[kGrayScratchBytes];

namespace {

struct GraySetupArgs {
    int32_t targets[kGrayNoAllocTargets];
    int32_t n_targets, block_rows, c, n_wg;
};

// Task i (thread i < n_targets) and workgroup table entry k (thread k < n_wg),
// equal to what gray_coset_min_dev uploads: block = rows [0, block_rows)
// minus the target in ascending order, unused blk entries 0, each task's
// 2^(nbits - c - 8) workgroups after the previous task's.
__global__ void gray_setup_kernel(GraySetupArgs a, GrayTask* __restrict__ tasks,
                                  int32_t* __restrict__ wg_task) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int first = 0;
    for (int k = 0; k < a.n_targets; ++k) {
        const int t = a.targets[k];
        const int nbits = a.block_rows - (t < a.block_rows ? 1 : 0);
        const int n_wg = 1 << (nbits - a.c - 8);
        if (k == i) {
            GrayTask g;
            g.target = t;
            g.nbits = nbits;
            g.first_wg = first;
            g.slot = k;
            int b = 0;
            for (int r = 0; r < a.block_rows; ++r)
                if (r != t)
                    g.blk[b++] = r;
            for (; b < kGrayMaxBlock; ++b)
                g.blk[b] = 0;
            tasks[k] = g;
        }
        if (i >= first && i < first + n_wg)
            wg_task[i] = k;
        first += n_wg;
    }
}

// Rows zero-padded from `words` to `pw` words per row.
__global__ void gray_pad_kernel(const uint64_t* __restrict__ src, size_t words, size_t pw,
                                size_t n, uint64_t* __restrict__ dst) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const size_t r = i / pw, w = i % pw;
    dst[i] = w < words ? src[r * words + w] : 0ull;
}

std::mutex g_scratch_mutex;
void* g_scratch_ptr[256] = {};

} // namespace

// Device address of `scratch` on the current device, looked up once per
// device with that device current (the array's host-side name is not a
// device address).
unsigned char* gray_scratch_ptr() {
    int dev = 0;
    BMM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 256)
        fail(BMM_ERR_DEVICE, "gray_coset_min_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            fail(BMM_ERR_DEVICE, std::string("gray_coset_min_noalloc: hipGetSymbolAddress: ") +
                                     hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    return static_cast<unsigned char*>(g_scratch_ptr[dev]);
}

void gray_coset_min_noalloc(const uint64_t* d_rows, size_t n_rows, size_t words_per_row,
                            size_t block_rows, const int32_t* targets, size_t n_targets,
                            size_t chunk_bits, uint64_t* d_first_key, uint64_t* d_mask_key,
                            hipStream_t stream) {
    if (!d_rows || !d_first_key || !d_mask_key || (n_targets && !targets))
        fail(BMM_ERR_INVALID, "gray_coset_min_noalloc: null pointer");
    gray_check_args(n_rows, words_per_row, block_rows, targets, n_targets, chunk_bits);
    if (n_targets > (size_t)kGrayNoAllocTargets)
        fail(BMM_ERR_INVALID, "gray_coset_min_noalloc: at most 64 targets per call");
    if (n_targets == 0)
        return;
    GraySetupArgs a{};
    int min_bits = 64;
    for (size_t i = 0; i < n_targets; ++i) {
        const int t = targets[i];
        if (t < 0 || (size_t)t >= n_rows)
            fail(BMM_ERR_RANGE, "gray_coset_min: target row out of range");
        const int b = (int)block_rows - ((size_t)t < block_rows ? 1 : 0);
        if (b > kGrayMaxBlock)
            fail(BMM_ERR_INVALID, "gray_coset_min: block larger than 40 rows");
        min_bits = std::min(min_bits, b);
        a.targets[i] = t;
    }
    const int c = chunk_bits ? (int)chunk_bits : gray_auto_chunk(min_bits);
    if (min_bits - c < 8 || c < 4 || c > 20)
        fail(BMM_ERR_INVALID,
             "gray_coset_min: block must have >= chunk_bits + 8 rows and chunk_bits in [4, 20]");
    size_t n_wg = 0;
    for (size_t i = 0; i < n_targets; ++i) {
        const int b = (int)block_rows - ((size_t)a.targets[i] < block_rows ? 1 : 0);
        n_wg += (size_t)1 << (b - c - 8);
    }
    if (n_wg > kGrayNoAllocWg)
        fail(BMM_ERR_INVALID, "gray_coset_min_noalloc: more than 61440 workgroups (use fewer "
                              "targets or a larger chunk_bits)");
    size_t pw = 1;
    while (pw < words_per_row)
        pw <<= 1;
    const bool pad = pw != words_per_row;
    if (pad && n_rows * pw > kGrayNoAllocRowWords)
        fail(BMM_ERR_INVALID, "gray_coset_min_noalloc: padded rows exceed the 768 KiB scratch "
                              "region (use words_per_row 1, 2, 4, 8, 16 or 32)");
    a.n_targets = (int32_t)n_targets;
    a.block_rows = (int32_t)block_rows;
    a.c = c;
    a.n_wg = (int32_t)n_wg;
    unsigned char* base = gray_scratch_ptr();
    GrayTask* d_tasks = reinterpret_cast<GrayTask*>(base);
    int32_t* d_wg = reinterpret_cast<int32_t*>(base + kGrayScratchWgOff);
    const int setup_threads = (int)std::max(n_wg, n_targets);
    hipLaunchKernelGGL(gray_setup_kernel, dim3((setup_threads + 255) / 256), dim3(256), 0, stream,
                       a, d_tasks, d_wg);
    BMM_HIP(hipGetLastError());
    const uint64_t* rows = d_rows;
    if (pad) {
        uint64_t* padded = reinterpret_cast<uint64_t*>(base + kGrayScratchRowsOff);
        const size_t n = n_rows * pw;
        hipLaunchKernelGGL(gray_pad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           stream, d_rows, words_per_row, pw, n, padded);
        BMM_HIP(hipGetLastError());
        rows = padded;
    }
    fill_u64(d_first_key, n_targets, ~0ull, stream);
    fill_u64(d_mask_key, n_targets, ~0ull, stream);
    const uint32_t* rows32 = reinterpret_cast<const uint32_t*>(rows);
    const int nw = (int)n_wg;
    switch (pw) {
    case 1: launch_gray<2, 1>(rows32, d_tasks, d_wg, nw, c, d_first_key, d_mask_key, stream, false); break;
    case 2: launch_gray<4, 2>(rows32, d_tasks, d_wg, nw, c, d_first_key, d_mask_key, stream, false); break;
    case 4: launch_gray<8, 3>(rows32, d_tasks, d_wg, nw, c, d_first_key, d_mask_key, stream, false); break;
    case 8: launch_gray<16, 3>(rows32, d_tasks, d_wg, nw, c, d_first_key, d_mask_key, stream, false); break;
    case 16: launch_gray<32, 3>(rows32, d_tasks, d_wg, nw, c, d_first_key, d_mask_key, stream, false); break;
    default: launch_gray<64, 1>(rows32, d_tasks, d_wg, nw, c, d_first_key, d_mask_key, stream, false); break;
    }
}

} // namespace bmmhip
