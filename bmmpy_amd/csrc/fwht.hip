// Walsh-Hadamard kernels on gfx950.
//
// 1. fwht_pass_kernel: the stand-alone in-place transform of a device array
//    (fwht_inplace, src/math/fwht.cpp:28-57), as passes of up to 12 index
//    bits each: the first pass transforms contiguous 4096-element chunks in
//    LDS, later passes transform 64-column tiles strided by 2^lo, so every
//    global access is a coalesced 256-byte row.
//
// 2. The block transform used by the searches.  For a window of t rows the
//    weight of mask m is (total - H[m]) / 2 with H = FWHT(b), b[p] = number of
//    active columns with row pattern p (src/search/fwht_search.cpp:59-112).
//    Split m = mh << L | ml: H[mh, ml] = FWHT_L(g_mh)[ml] with
//    g_mh[pl] = sum over columns of (-1)^popcount(ph & mh) at pl.  One
//    workgroup builds g_mh in LDS from the column patterns, transforms its
//    2^L entries on chip (radix-32 register rounds between LDS exchanges) and
//    scans the results, so no transform data ever goes to HBM.
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

// ------------------------------------------------- coset minimum (C1/C4) --
struct CosetMatrix {
    int64_t pat_off;   // offset of the matrix's column patterns
    int64_t out_off;   // output slot of target 0
    int32_t rows;      // r
    int32_t cols;
    int32_t total;     // active columns
    int32_t L;         // transform bits per block (min(r, kCosetMaxL))
};

struct CosetTask {
    int32_t matrix;
    int32_t mh;        // high part of the mask
};

constexpr int kCosetMaxL = 13;

template <int L>
__device__ void coset_block_phase1(int32_t* s, const CosetMatrix& M, const uint32_t* pat,
                                   uint32_t mh, unsigned long long* keys) {
    block_build_g<L>(s, pat, M.cols, mh);
    block_fwht<L>(s);
    // Scan: element ml = e * 256 + tid (register bits = ml bits >= 8).
    constexpr int N = 1 << L;
    constexpr int E = N >= 256 ? N / 256 : 1;
    constexpr int HB = L > 8 ? L - 8 : 0;  // register-resident ml bits
    const int tid = threadIdx.x;
    uint32_t overall = 0xffffffffu;
    uint32_t mreg[HB > 0 ? HB : 1];
#pragma unroll
    for (int i = 0; i < (HB > 0 ? HB : 1); ++i)
        mreg[i] = 0xffffffffu;
    if (tid < N) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int ml = e * 256 + tid;
            const int32_t h = s[fwht_pad(ml)];
            const uint32_t w = (uint32_t)(M.total - h) >> 1;
            overall = min(overall, (ml | mh) ? w : 0xffffffffu);
#pragma unroll
            for (int i = 0; i < HB; ++i)
                if ((e >> i) & 1)
                    mreg[i] = min(mreg[i], w);
        }
    }
    // Per-target block minima -> global keys (weight << 32 | mh).
    __shared__ uint32_t red[32][4];
    const int wv = tid / kWave, lane = tid % kWave;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        uint32_t v;
        if (i < 8)
            v = ((tid >> i) & 1) ? overall : 0xffffffffu;
        else
            v = mreg[i >= 8 ? i - 8 : 0];
        v = wave_min_u32(v);
        if (lane == 0)
            red[i][wv] = v;
    }
    const uint32_t wover = wave_min_u32(overall);
    for (int i = L; i < M.rows; ++i)
        if (lane == 0)
            red[i][wv] = ((mh >> (i - L)) & 1u) ? wover : 0xffffffffu;
    __syncthreads();
    if (tid < M.rows) {
        const uint32_t v = min(min(red[tid][0], red[tid][1]), min(red[tid][2], red[tid][3]));
        if (v != 0xffffffffu)
            atomicMin(&keys[M.out_off + tid], ((unsigned long long)v << 32) | mh);
    }
    __syncthreads();
}

template <int L>
__device__ void coset_block_phase2(int32_t* s, const CosetMatrix& M, const uint32_t* pat,
                                   int target, unsigned long long key, uint32_t* low_out) {
    const uint32_t mh = (uint32_t)key;
    const uint32_t wmin = (uint32_t)(key >> 32);
    block_build_g<L>(s, pat, M.cols, mh);
    block_fwht<L>(s);
    constexpr int N = 1 << L;
    __shared__ uint32_t best;
    if (threadIdx.x == 0)
        best = 0xffffffffu;
    __syncthreads();
    uint32_t mine = 0xffffffffu;
    for (int ml = threadIdx.x; ml < N; ml += blockDim.x) {
        const uint32_t w = (uint32_t)(M.total - s[fwht_pad(ml)]) >> 1;
        const bool has = target < L ? ((ml >> target) & 1) : true;
        if (has && w == wmin && (ml | mh))
            mine = min(mine, (uint32_t)ml);
    }
    mine = wave_min_u32(mine);
    if ((threadIdx.x % kWave) == 0 && mine != 0xffffffffu)
        atomicMin(&best, mine);
    __syncthreads();
    if (threadIdx.x == 0)
        *low_out = best;
    __syncthreads();
}

__global__ __launch_bounds__(256) void coset_phase1_kernel(const CosetMatrix* __restrict__ mats,
                                                           const uint32_t* __restrict__ pats,
                                                           const CosetTask* __restrict__ tasks,
                                                           int n_tasks,
                                                           unsigned int* __restrict__ counter,
                                                           unsigned long long* __restrict__ keys) {
    __shared__ int32_t s[fwht_lds_words(kCosetMaxL)];
    __shared__ int task_id;
    for (;;) {
        if (threadIdx.x == 0)
            task_id = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int ti = task_id;
        __syncthreads();
        if (ti >= n_tasks)
            break;
        const CosetTask t = tasks[ti];
        const CosetMatrix M = mats[t.matrix];
        const uint32_t* pat = pats + M.pat_off;
        switch (M.L) {
#define BMM_COSET_CASE(LL)                                                                     \
    case LL: coset_block_phase1<LL>(s, M, pat, (uint32_t)t.mh, keys); break;
            BMM_COSET_CASE(1) BMM_COSET_CASE(2) BMM_COSET_CASE(3) BMM_COSET_CASE(4)
            BMM_COSET_CASE(5) BMM_COSET_CASE(6) BMM_COSET_CASE(7) BMM_COSET_CASE(8)
            BMM_COSET_CASE(9) BMM_COSET_CASE(10) BMM_COSET_CASE(11) BMM_COSET_CASE(12)
            BMM_COSET_CASE(13)
#undef BMM_COSET_CASE
        default: break;
        }
    }
}

// One workgroup per (matrix, target): recompute the first block reaching the
// target's minimum and pick its lowest low part; then write weight + index.
__global__ __launch_bounds__(256) void coset_phase2_kernel(
    const CosetMatrix* __restrict__ mats, const uint32_t* __restrict__ pats,
    const int32_t* __restrict__ slot_matrix, const unsigned long long* __restrict__ keys,
    uint32_t* __restrict__ out_w, uint64_t* __restrict__ out_idx) {
    __shared__ int32_t s[fwht_lds_words(kCosetMaxL)];
    __shared__ uint32_t low;
    const int slot = blockIdx.x;
    const int mi = slot_matrix[slot];
    const CosetMatrix M = mats[mi];
    const int target = (int)(slot - M.out_off);
    const unsigned long long key = keys[slot];
    const uint32_t* pat = pats + M.pat_off;
    switch (M.L) {
#define BMM_COSET_CASE(LL)                                                                     \
    case LL: coset_block_phase2<LL>(s, M, pat, target, key, &low); break;
        BMM_COSET_CASE(1) BMM_COSET_CASE(2) BMM_COSET_CASE(3) BMM_COSET_CASE(4)
        BMM_COSET_CASE(5) BMM_COSET_CASE(6) BMM_COSET_CASE(7) BMM_COSET_CASE(8)
        BMM_COSET_CASE(9) BMM_COSET_CASE(10) BMM_COSET_CASE(11) BMM_COSET_CASE(12)
        BMM_COSET_CASE(13)
#undef BMM_COSET_CASE
    default: break;
    }
    if (threadIdx.x == 0) {
        const uint64_t mask = ((uint64_t)(uint32_t)key << M.L) | low;
        // Coset index: drop the target bit and close the gap.
        const uint64_t below = mask & ((1ull << target) - 1ull);
        const uint64_t above = mask >> (target + 1);
        out_w[slot] = (uint32_t)(key >> 32);
        out_idx[slot] = below | (above << target);
    }
}

namespace {
struct CosetWorkspace {
    DevBuf mats, pats, tasks, keys, counter, slot_matrix, rows, outw, outi;
};
CosetWorkspace& coset_ws() {
    thread_local CosetWorkspace ws;
    return ws;
}
} // namespace

struct CosetPlan {
    std::vector<CosetMatrix> mats;
    std::vector<CosetTask> tasks;
    std::vector<int32_t> slot_matrix;
    int64_t n_slots = 0;
    int64_t n_pats = 0;
};

// Builds matrix descriptors and the LPT-ordered task list (largest blocks
// first; within a matrix, ascending mh).
CosetPlan make_coset_plan(const int32_t* rows, const int32_t* cols, size_t n_mat,
                          const int64_t* out_offset) {
    CosetPlan p;
    p.mats.resize(n_mat);
    int64_t pat_off = 0;
    for (size_t m = 0; m < n_mat; ++m) {
        if (rows[m] < 2 || rows[m] > 31)
            fail(BMM_ERR_INVALID, "coset_min: rows must be in [2, 31]");
        if (cols[m] < 0)
            fail(BMM_ERR_INVALID, "coset_min: cols must be non-negative");
        CosetMatrix& M = p.mats[m];
        M.rows = rows[m];
        M.cols = cols[m];
        M.pat_off = pat_off;
        M.out_off = out_offset ? out_offset[m] : p.n_slots;
        M.L = std::min(rows[m], kCosetMaxL);
        M.total = 0;
        pat_off += std::max(cols[m], 1);
        p.n_slots = std::max<int64_t>(p.n_slots, M.out_off + M.rows);
    }
    p.n_pats = pat_off;
    std::vector<int> order(n_mat);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return p.mats[a].L > p.mats[b].L; });
    for (int m : order) {
        const int nb = 1 << (p.mats[m].rows - p.mats[m].L);
        for (int mh = 0; mh < nb; ++mh)
            p.tasks.push_back(CosetTask{m, mh});
    }
    p.slot_matrix.assign((size_t)p.n_slots, -1);
    for (size_t m = 0; m < n_mat; ++m)
        for (int i = 0; i < p.mats[m].rows; ++i) {
            int32_t& sm = p.slot_matrix[(size_t)(p.mats[m].out_off + i)];
            if (sm != -1)
                fail(BMM_ERR_INVALID, "coset_min: overlapping output offsets");
            sm = (int32_t)m;
        }
    for (int32_t sm : p.slot_matrix)
        if (sm < 0)
            fail(BMM_ERR_INVALID, "coset_min: output offsets leave gaps");
    return p;
}

__global__ void count_active_kernel(const uint32_t* __restrict__ pats,
                                    CosetMatrix* __restrict__ mats, int n_mat) {
    const int m = blockIdx.x;
    if (m >= n_mat)
        return;
    CosetMatrix& M = mats[m];
    uint32_t cnt = 0;
    for (int c = threadIdx.x; c < M.cols; c += blockDim.x)
        cnt += pats[M.pat_off + c] != 0u;
    cnt = wave_sum_u32(cnt);
    __shared__ uint32_t red[4];
    if (threadIdx.x % kWave == 0)
        red[threadIdx.x / kWave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0)
        M.total = (int32_t)(red[0] + red[1] + red[2] + red[3]);
}

// Device-resident batched coset minimum.  d_rows holds each matrix's rows at
// word_offset[m] (host array); outputs are device arrays of n_slots entries.
void coset_min_batch_dev(const uint64_t* d_rows, const int64_t* word_offset, const int32_t* rows,
                         const int32_t* cols, size_t n_mat, const int64_t* out_offset,
                         uint32_t* d_out_w, uint64_t* d_out_idx, int persistent_wgs,
                         hipStream_t stream) {
    CosetPlan plan = make_coset_plan(rows, cols, n_mat, out_offset);
    CosetWorkspace& ws = coset_ws();
    uint32_t* d_pats = ws.pats.get<uint32_t>((size_t)plan.n_pats);
    CosetMatrix* d_mats = ws.mats.get<CosetMatrix>(n_mat);
    CosetTask* d_tasks = ws.tasks.get<CosetTask>(plan.tasks.size());
    unsigned long long* d_keys = ws.keys.get<unsigned long long>((size_t)plan.n_slots);
    unsigned int* d_counter = ws.counter.get<unsigned int>(1);
    int32_t* d_slot = ws.slot_matrix.get<int32_t>((size_t)plan.n_slots);
    for (size_t m = 0; m < n_mat; ++m) {
        const int words = (cols[m] + 63) / 64;
        if (cols[m] == 0)
            continue;
        hipLaunchKernelGGL(column_patterns_kernel, dim3((cols[m] + 255) / 256), dim3(256), 0,
                           stream, d_rows + word_offset[m], rows[m], words, cols[m],
                           d_pats + plan.mats[m].pat_off);
    }
    BMM_HIP(hipGetLastError());
    BMM_HIP(hipMemcpyAsync(d_mats, plan.mats.data(), sizeof(CosetMatrix) * n_mat,
                           hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(count_active_kernel, dim3((unsigned)n_mat), dim3(256), 0, stream, d_pats,
                       d_mats, (int)n_mat);
    BMM_HIP(hipMemcpyAsync(d_tasks, plan.tasks.data(), sizeof(CosetTask) * plan.tasks.size(),
                           hipMemcpyHostToDevice, stream));
    BMM_HIP(hipMemcpyAsync(d_slot, plan.slot_matrix.data(),
                           sizeof(int32_t) * plan.slot_matrix.size(), hipMemcpyHostToDevice,
                           stream));
    fill_u64(reinterpret_cast<uint64_t*>(d_keys), (size_t)plan.n_slots, ~0ull, stream);
    fill_u32(d_counter, 1, 0u, stream);
    const int grid = std::max(1, std::min<int>(persistent_wgs, (int)plan.tasks.size()));
    hipLaunchKernelGGL(coset_phase1_kernel, dim3(grid), dim3(256), 0, stream, d_mats, d_pats,
                       d_tasks, (int)plan.tasks.size(), d_counter, d_keys);
    BMM_HIP(hipGetLastError());
    hipLaunchKernelGGL(coset_phase2_kernel, dim3((unsigned)plan.n_slots), dim3(256), 0, stream,
                       d_mats, d_pats, d_slot, d_keys, d_out_w, d_out_idx);
    BMM_HIP(hipGetLastError());
}

int default_persistent_wgs() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return cus * 8;
}

} // namespace bmmhip
