// Batched per-target coset minimum (many matrices of uneven size, one
// persistent launch fed by a device work queue).
//
// For a matrix of r rows every row is a target: min over subsets S of the other
// rows of wt(row_i ^ XOR_S).  Equivalently, over the full-window ranking of
// the reference FwhtSearch (src/search/fwht_search.cpp:59-126), the lowest
// (weight, mask) among masks that contain bit i.  One full 2^r transform per
// matrix serves all r targets: each block mh of 2^L masks is transformed on
// chip (fwht_block.hpp) and scanned once, keeping per-target block minima:
//   * target bits below L: per-thread minima over the elements with that bit
//     (register bits) or the thread's own minimum (thread bits);
//   * target bits at or above L: the block minimum when mh has the bit.
// Each block folds (weight << 32 | mh) into the target's key with atomicMin,
// so the key ends at the minimum weight and the lowest block reaching it;
// phase 2 recomputes that one block per target and picks its lowest mask.
//
// Work queue: matrices are ordered by block size (largest first, LPT) and
// their blocks numbered through a prefix sum; persistent workgroups take
// `chunk` consecutive block numbers per atomicAdd and locate the matrix by
// binary search over the prefix.  A plan (device descriptors, launch shape)
// is built once; its whole launch sequence is captured in a hipGraph, and the
// launch shape can be picked by timing candidates on the plan's own data.
#include "common.hpp"
#include "fwht_block.hpp"
#include "fwht_pk16.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <tuple>
#include <numeric>
#include <vector>

namespace bmmhip {

struct CosetMatrix {
    int64_t word_off;  // first word of the matrix in the caller's row buffer
    int64_t pat_off;   // first column pattern
    int64_t out_off;   // output slot of target 0
    int32_t rows;
    int32_t cols;
    int32_t words;
    int32_t total;     // active columns (set on the device)
    int32_t L;         // transform bits per block
    int32_t pad;
};

constexpr int kBatchMaxL = 13;

// Column patterns (bit i = row i) and the active-column count; one workgroup
// per matrix.
__global__ __launch_bounds__(256) void batch_patterns_kernel(const uint64_t* __restrict__ words,
                                                             CosetMatrix* __restrict__ mats,
                                                             uint32_t* __restrict__ pats) {
    CosetMatrix& M = mats[blockIdx.x];
    const uint64_t* rows = words + M.word_off;
    uint32_t cnt = 0;
    for (int c = threadIdx.x; c < M.cols; c += blockDim.x) {
        const int wi = c >> 6, bit = c & 63;
        uint32_t p = 0;
        for (int i = 0; i < M.rows; ++i)
            p |= (uint32_t)((rows[(size_t)i * M.words + wi] >> bit) & 1ull) << i;
        pats[M.pat_off + c] = p;
        cnt += p != 0u;
    }
    cnt = wave_sum_u32(cnt);
    __shared__ uint32_t red[4];
    if (threadIdx.x % kWave == 0)
        red[threadIdx.x / kWave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0)
        M.total = (int32_t)(red[0] + red[1] + red[2] + red[3]);
}

// Per-target block maxima of H from per-thread values, folded into the keys
// as (weight << 32 | mh).  Thread t's elements have ml bits 0..7 = t;
// `overall` is the thread's maximum, mreg[j] its maximum over the elements
// with ml bit 8 + j set (j < L - 8).  Target bits below 8 reduce row 0 over
// the threads with that bit, bits 8..L-1 reduce their rows, bits >= L take
// the block maximum when mh has the bit.  With SHIFTED (the packed block's
// paired last round) thread t's elements have ml bits 1..8 = t instead, and
// mreg[0..1] stand for ml bits 0 and 9, mreg[2..] for bits 10 and up.
template <int L, bool SHIFTED = false>
__device__ void coset_block_finish(int32_t* s, const CosetMatrix& M, uint32_t mh,
                                   unsigned long long* keys, int32_t overall,
                                   const int32_t* mreg) {
    constexpr int HB = L > 8 ? L - 8 : 0;
    const int tid = threadIdx.x;
    int32_t* red = s;
    __shared__ int32_t res[kPkL + 1];
    __syncthreads();
    red[tid] = overall;
#pragma unroll
    for (int i = 0; i < HB; ++i)
        red[(1 + i) * 256 + tid] = mreg[i];
    __syncthreads();
    const int wv = tid / kWave, lane = tid % kWave;
    for (int q = wv; q <= L; q += 4) {
        int32_t v = INT32_MIN;
        for (int x = lane; x < 256; x += kWave) {
            int32_t val;
            if (q < L && q < 8)
                val = ((x >> q) & 1) ? red[x] : INT32_MIN;
            else if (q < L)
                val = red[(1 + q - 8) * 256 + x];
            else
                val = red[x];
            v = max(v, val);
        }
        v = wave_max_i32(v);
        if (lane == 0)
            res[q] = v;
    }
    __syncthreads();
    if (tid < M.rows) {
        const bool has = tid < L || ((mh >> (tid - L)) & 1u);
        // Row of target bit b: thread bit b (or b - 1), else the register
        // bit's row 8 + its index.
        int row = tid < L ? tid : L;
        if (SHIFTED && tid < L)
            row = tid == 0 ? 8 : tid < 9 ? tid - 1 : tid;
        const int32_t h = res[row];
        if (has && h != INT32_MIN) {
            // Keys only decrease, so a stale read never skips an improvement.
            const uint32_t w = (uint32_t)(M.total - h) >> 1;
            const unsigned long long k = ((unsigned long long)w << 32) | mh;
            unsigned long long* dst = &keys[M.out_off + tid];
            if (k < __atomic_load_n(dst, __ATOMIC_RELAXED))
                atomicMin(dst, k);
        }
    }
    __syncthreads();
}

// Transform of block mh (built into LDS by `build`), int32 entries: thread t
// holds elements ml = i * 256 + t, so ml bits 8..L-1 index its register
// values; maximising H minimises the weight (weight = (total - H) / 2).
template <int L, typename Build>
__device__ void coset_block_phase1(int32_t* s, const CosetMatrix& M, uint32_t mh,
                                   unsigned long long* keys, Build&& build) {
    build(s);
    constexpr int N = 1 << L;
    constexpr int HB = L > 8 ? L - 8 : 0;
    const int tid = threadIdx.x;
    int32_t overall = INT32_MIN;
    int32_t mreg[HB > 0 ? HB : 1];
#pragma unroll
    for (int i = 0; i < (HB > 0 ? HB : 1); ++i)
        mreg[i] = INT32_MIN;
    // v[e] is register index q + PER * e: subset maxima over e, then the bits
    // of q (fixed per call) take the group maximum.
    auto visit = [&](int q, const auto& v) {
        constexpr int E = sizeof(v) / sizeof(v[0]);
        constexpr int LE = ilog2c(E);
        constexpr int LP = HB - LE;
        int32_t x[E], bm[LE > 0 ? LE : 1];
#pragma unroll
        for (int e = 0; e < E; ++e)
            x[e] = (int32_t)v[e];
        const int32_t all = subset_max<E>(x, bm);
#pragma unroll
        for (int j = 0; j < LE; ++j)
            mreg[LP + j] = max(mreg[LP + j], bm[j]);
#pragma unroll
        for (int i = 0; i < LP; ++i)
            if ((q >> i) & 1)
                mreg[i] = max(mreg[i], all);
        overall = max(overall, all);
    };
    if constexpr (L >= 11) {
        block_fwht_visit<L, 0>(s, visit);
    } else {
        constexpr int E = N >= 256 ? N / 256 : 1;
        block_fwht<L>(s);
        if (tid < N) {
            uint32_t v[E];
#pragma unroll
            for (int e = 0; e < E; ++e)
                v[e] = (uint32_t)s[fwht_pad(e * 256 + tid)];
            visit(0, v);
        }
    }
    coset_block_finish<L>(s, M, mh, keys, overall, mreg);
}

// Packed-int16 block of 2^14 entries (matrices of >= 14 rows and at most
// 32767 columns).  The last round hands thread t the words
// 2 t + j + 512 h + 1024 e, each holding entries w and w + 2^13: ml bit 0 is
// j, bits 1-8 the thread, bit 9 h, bits 10-12 e and bit 13 the half; the
// visit's q = j | h << 1.
template <typename Build>
__device__ void coset_block_pk16(uint32_t* s, const CosetMatrix& M, uint32_t mh,
                                 unsigned long long* keys, Build&& build) {
    build(s);
    const s16x2 lowest = {INT16_MIN, INT16_MIN};
    s16x2 overall = lowest;
    s16x2 mreg[5];
#pragma unroll
    for (int i = 0; i < 5; ++i)
        mreg[i] = lowest;
    pk_block_fwht_visit_paired(s, [&](int q, const uint32_t (&v)[8]) {
        s16x2 x[8], bm[3];
#pragma unroll
        for (int e = 0; e < 8; ++e)
            x[e] = pk_as(v[e]);
        const s16x2 all = pk_subset_max<8>(x, bm);
#pragma unroll
        for (int j = 0; j < 3; ++j)
            mreg[2 + j] = pk_max(mreg[2 + j], bm[j]);
        if (q & 1)
            mreg[0] = pk_max(mreg[0], all);
        if (q & 2)
            mreg[1] = pk_max(mreg[1], all);
        overall = pk_max(overall, all);
    });
    int32_t m[6];
#pragma unroll
    for (int i = 0; i < 5; ++i)
        m[i] = max(pk_lo(mreg[i]), pk_hi(mreg[i]));
    m[5] = pk_hi(overall);
    coset_block_finish<kPkL, true>(reinterpret_cast<int32_t*>(s), M, mh, keys,
                             max(pk_lo(overall), pk_hi(overall)), m);
}

// Column patterns of the current matrix kept in registers across its blocks
// (up to kPatRegs * 256 columns; wider matrices read them from memory).
constexpr int kPatRegs = 8;

template <int L>
__device__ inline void build_from_regs(int32_t* s, const uint32_t (&preg)[kPatRegs], uint32_t mh) {
    constexpr int N = 1 << L;
    int4* s4 = reinterpret_cast<int4*>(s);
    for (int i = threadIdx.x; i < (fwht_lds_words(L) + 3) / 4; i += blockDim.x)
        s4[i] = make_int4(0, 0, 0, 0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPatRegs; ++k) {
        const uint32_t p = preg[k];
        if (p != 0u) {
            const int neg = __builtin_popcount((p >> L) & mh) & 1;
            atomicAdd(&s[fwht_pad((int)(p & (N - 1)))], neg ? -1 : 1);
        }
    }
}

// Blocks of the matrices with at least `L` rows (all of them 2^L wide):
// persistent workgroups take `chunk` consecutive block numbers per atomicAdd.
// The matrix of a block is found by galloping forward from the previous
// one (block numbers handed to a workgroup only grow), then bisecting.
template <int L>
__global__ __launch_bounds__(256) void coset_full_kernel(
    const CosetMatrix* __restrict__ mats, const int64_t* __restrict__ prefix, int n_full,
    int64_t total_blocks, int chunk, const uint32_t* __restrict__ pats,
    unsigned long long* __restrict__ counter, unsigned long long* __restrict__ keys) {
    constexpr int kWords = L == kPkL ? pk_lds_words() : fwht_lds_words(L);
    __shared__ __align__(16) int32_t s[(kWords > 7 * 256 ? kWords : 7 * 256) + 4];
    __shared__ long long first;
    int idx = 0, cached = -1;
    uint32_t preg[kPatRegs];
    CosetMatrix M{};
    for (;;) {
        if (threadIdx.x == 0)
            first = (long long)atomicAdd(counter, (unsigned long long)chunk);
        __syncthreads();
        const int64_t g0 = first;
        __syncthreads();
        if (g0 >= total_blocks)
            break;
        const int64_t g1 = min(g0 + (int64_t)chunk, total_blocks);
        if (prefix[idx + 1] <= g0) {
            int step = 1, lo = idx + 1, hi = idx + 1;
            while (hi < n_full - 1 && prefix[hi + 1] <= g0) {
                lo = hi + 1;
                hi = min(hi + step, n_full - 1);
                step *= 2;
            }
            // prefix[lo] <= g0 < prefix[hi + 1]
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (prefix[mid] <= g0)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            idx = lo;
        }
        for (int64_t g = g0; g < g1; ++g) {
            while (prefix[idx + 1] <= g)
                ++idx;
            if (idx != cached) {
                M = mats[idx];
                const uint32_t* pat = pats + M.pat_off;
#pragma unroll
                for (int k = 0; k < kPatRegs; ++k) {
                    const int c = (int)threadIdx.x + 256 * k;
                    preg[k] = c < M.cols ? pat[c] : 0u;
                }
                cached = idx;
            }
            const uint32_t mh = (uint32_t)(g - prefix[idx]);
            if constexpr (L == kPkL) {
                uint32_t* su = reinterpret_cast<uint32_t*>(s);
                if (M.cols <= kPatRegs * 256)
                    coset_block_pk16(su, M, mh, keys,
                                     [&](uint32_t* sp) { pk_build(sp, preg, mh); });
                else
                    coset_block_pk16(su, M, mh, keys, [&](uint32_t* sp) {
                        pk_build_mem(sp, pats + M.pat_off, M.cols, mh);
                    });
            } else if (M.cols <= kPatRegs * 256) {
                coset_block_phase1<L>(s, M, mh, keys,
                                      [&](int32_t* sp) { build_from_regs<L>(sp, preg, mh); });
            } else {
                coset_block_phase1<L>(s, M, mh, keys, [&](int32_t* sp) {
                    block_build_g<L>(sp, pats + M.pat_off, M.cols, mh);
                });
            }
        }
    }
}

// Matrices with fewer rows than the block size: one workgroup each, one
// block of 2^rows masks.
__global__ __launch_bounds__(256) void coset_small_kernel(const CosetMatrix* __restrict__ mats,
                                                         const uint32_t* __restrict__ pats,
                                                         unsigned long long* __restrict__ keys) {
    __shared__ __align__(16) int32_t s[fwht_lds_words(kBatchMaxL) + 4];
    const CosetMatrix M = mats[blockIdx.x];
    const uint32_t* pat = pats + M.pat_off;
    switch (M.L) {
#define BMM_CASE(LL)                                                                           \
    case LL:                                                                                   \
        coset_block_phase1<LL>(s, M, 0u, keys,                                                 \
                               [&](int32_t* sp) { block_build_g<LL>(sp, pat, M.cols, 0u); });  \
        break;
        BMM_CASE(1) BMM_CASE(2) BMM_CASE(3) BMM_CASE(4) BMM_CASE(5) BMM_CASE(6) BMM_CASE(7)
        BMM_CASE(8) BMM_CASE(9) BMM_CASE(10) BMM_CASE(11) BMM_CASE(12) BMM_CASE(13)
#undef BMM_CASE
    default: break;
    }
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
        if (has && w == wmin)
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

// Phase 2 for a packed-int16 block of 2^14 entries.
__device__ void coset_block_phase2_pk16(uint32_t* s, const CosetMatrix& M, const uint32_t* pat,
                                        int target, unsigned long long key,
                                        uint32_t* low_out) {
    const uint32_t mh = (uint32_t)key;
    const int32_t want = M.total - 2 * (int32_t)(uint32_t)(key >> 32);  // H of the minimum
    __shared__ uint32_t best;
    if (threadIdx.x == 0)
        best = 0xffffffffu;
    pk_build_mem(s, pat, M.cols, mh);
    uint32_t mine = 0xffffffffu;
    pk_block_fwht_visit(s, [&](int q, const uint32_t (&v)[8]) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t w = (uint32_t)((q + 4 * e) * 256 + (int)threadIdx.x);
            const s16x2 x = pk_as(v[e]);
            const uint32_t ml_lo = w, ml_hi = w + (1u << kPkW);
            const bool has_lo = target < kPkL ? ((ml_lo >> target) & 1u) : true;
            const bool has_hi = target < kPkL ? ((ml_hi >> target) & 1u) : true;
            if (has_lo && pk_lo(x) == want)
                mine = min(mine, ml_lo);
            if (has_hi && pk_hi(x) == want)
                mine = min(mine, ml_hi);
        }
    });
    mine = wave_min_u32(mine);
    if ((threadIdx.x % kWave) == 0 && mine != 0xffffffffu)
        atomicMin(&best, mine);
    __syncthreads();
    if (threadIdx.x == 0)
        *low_out = best;
    __syncthreads();
}

// One workgroup per output slot (matrix, target).  Each workgroup also
// resets its key to ~0 after reading it, and workgroup 0 the work-queue
// counter, so the next pass starts from the cleared state without fill
// launches (the plan clears both once when it is built).
__global__ __launch_bounds__(256) void coset_phase2_kernel(
    const CosetMatrix* __restrict__ mats, const int32_t* __restrict__ slot_code,
    const uint32_t* __restrict__ pats, unsigned long long* __restrict__ keys,
    unsigned long long* __restrict__ counter, uint32_t* __restrict__ out_w,
    uint64_t* __restrict__ out_idx) {
    constexpr int kWords = fwht_lds_words(kBatchMaxL) > pk_lds_words() ? fwht_lds_words(kBatchMaxL)
                                                                         : pk_lds_words();
    __shared__ __align__(16) int32_t s[kWords + 4];
    __shared__ uint32_t low;
    const int slot = blockIdx.x;
    const int32_t code = slot_code[slot];
    const CosetMatrix M = mats[code >> 5];
    const int target = code & 31;
    const unsigned long long key = keys[slot];
    __syncthreads();   // every thread has read the key before it is cleared
    if (threadIdx.x == 0) {
        keys[slot] = ~0ull;
        if (slot == 0)
            *counter = 0ull;
    }
    const uint32_t* pat = pats + M.pat_off;
    switch (M.L) {
#define BMM_CASE(LL)                                                                           \
    case LL: coset_block_phase2<LL>(s, M, pat, target, key, &low); break;
        BMM_CASE(1) BMM_CASE(2) BMM_CASE(3) BMM_CASE(4) BMM_CASE(5) BMM_CASE(6) BMM_CASE(7)
        BMM_CASE(8) BMM_CASE(9) BMM_CASE(10) BMM_CASE(11) BMM_CASE(12) BMM_CASE(13)
#undef BMM_CASE
    case kPkL:
        coset_block_phase2_pk16(reinterpret_cast<uint32_t*>(s), M, pat, target, key, &low);
        break;
    default: break;
    }
    if (threadIdx.x == 0) {
        const uint64_t mask = ((uint64_t)(uint32_t)key << M.L) | low;
        // Coset index over the other rows: drop the target bit, close the gap.
        const uint64_t below = mask & ((1ull << target) - 1ull);
        const uint64_t above = mask >> (target + 1);
        out_w[slot] = (uint32_t)(key >> 32);
        out_idx[slot] = below | (above << target);
    }
}

// ----------------------------------------------- single-launch form -----
// Small batches (every matrix at most kPatRegs * 256 columns, few blocks):
// one workgroup per block of 2^L masks (L = min(rows, 13)), one launch in
// all.  The workgroup reads its matrix's column patterns from the rows, builds
// and transforms its block, and takes for every target the block's best
// element among the masks that hold the target bit as the packed key
// K = H * 2^L + (2^L - 1 - ml) (largest H, then lowest ml); the exact
// (weight << 32 | mask) then goes to the target's key with one atomicMin, so
// the key ends at the minimum weight and its lowest mask and no second pass
// is needed.  The last workgroup to finish (completion counter) turns the
// keys into the outputs and clears keys and counter for the next launch.
constexpr int kFusedMaxRows = 31;
constexpr int64_t kFusedMaxBlocks = 1024;
constexpr int kFusedL = 13;   // 2^13-mask blocks (2^12: 256 workgroups, measured no faster)

template <int L>
__device__ void coset_block_exact(int32_t* s, const CosetMatrix& M, int32_t total, uint32_t mh,
                                  const uint32_t (&preg)[kPatRegs],
                                  unsigned long long* __restrict__ keys) {
    constexpr int N = 1 << L;
    build_from_regs<L>(s, preg, mh);
    const int tid = threadIdx.x;
    int32_t best[kFusedMaxRows];
#pragma unroll
    for (int b = 0; b < kFusedMaxRows; ++b)
        best[b] = INT32_MIN;
    int32_t overall = INT32_MIN;
    if constexpr (L >= 11) {
        // The last transform round stays in registers: entry ml = (q + 4 e) * 256
        // + tid, so bits 0-7 of ml are the thread's own and bits 8.. follow q and e.
        block_fwht_visit<L, 0>(s, [&](int q, const auto& v) {
#pragma unroll
            for (int e = 0; e < (1 << (L - 10)); ++e) {
                const int ml = (q + 4 * e) * 256 + tid;
                const int32_t k = (int32_t)v[e] * N + (N - 1 - ml);
                overall = max(overall, k);
#pragma unroll
                for (int b = 8; b < L; ++b)
                    if ((ml >> b) & 1)
                        best[b] = max(best[b], k);
            }
        });
#pragma unroll
        for (int b = 0; b < 8; ++b)
            best[b] = ((tid >> b) & 1) ? overall : INT32_MIN;
    } else {
        constexpr int E = N >= 256 ? N / 256 : 1;
        block_fwht<L>(s);
        if (tid < N) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int ml = e * 256 + tid;
                const int32_t k = s[fwht_pad(ml)] * N + (N - 1 - ml);
                overall = max(overall, k);
#pragma unroll
                for (int b = 0; b < L; ++b)
                    if ((ml >> b) & 1)
                        best[b] = max(best[b], k);
            }
        }
    }
    __shared__ int32_t red[4][kFusedMaxRows + 1];
    const int wv = tid / kWave, lane = tid % kWave;
#pragma unroll
    for (int b = 0; b < L; ++b) {
        const int32_t v = wave_max_i32(best[b]);
        if (lane == 0)
            red[wv][b] = v;
    }
    {
        const int32_t v = wave_max_i32(overall);
        if (lane == 0)
            red[wv][kFusedMaxRows] = v;
    }
    __syncthreads();
    if (tid < M.rows) {
        const int b = tid;
        const int row = b < L ? b : kFusedMaxRows;
        const bool has = b < L || ((mh >> (b - L)) & 1u);
        const int32_t k = max(max(red[0][row], red[1][row]), max(red[2][row], red[3][row]));
        if (has && k != INT32_MIN) {
            const int32_t h = k >> L;   // k = h * 2^L + (N - 1 - ml), 0 <= N - 1 - ml < 2^L
            const uint32_t ml = (uint32_t)(N - 1) - ((uint32_t)k & (uint32_t)(N - 1));
            const uint32_t w = (uint32_t)(total - h) >> 1;
            const unsigned long long key =
                ((unsigned long long)w << 32) | (((unsigned long long)mh << L) | ml);
            atomicMin(&keys[M.out_off + b], key);
        }
    }
}

// A one-matrix plan (n_mat == 1) takes its matrix from the argument `first`
// and reads no descriptor from memory; otherwise a workgroup finds its
// matrix in `mats` by bisecting the block prefix.
__global__ __launch_bounds__(256) void coset_fused_kernel(
    const uint64_t* __restrict__ words, CosetMatrix first, const CosetMatrix* __restrict__ mats,
    const int64_t* __restrict__ prefix, int n_mat, const int32_t* __restrict__ slot_code,
    int n_slots, unsigned long long* __restrict__ keys, unsigned int* __restrict__ done,
    uint32_t* __restrict__ out_w, uint64_t* __restrict__ out_idx) {
    __shared__ __align__(16) int32_t s[fwht_lds_words(kBatchMaxL) + 4];
    __shared__ uint32_t cnt[4];
    __shared__ int last;
    const int64_t g = blockIdx.x;
    CosetMatrix M = first;
    uint32_t mh = (uint32_t)g;
    if (n_mat > 1) {
        int lo = 0, hi = n_mat - 1;     // prefix[lo] <= g < prefix[lo + 1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= g)
                lo = mid;
            else
                hi = mid - 1;
        }
        M = mats[lo];
        mh = (uint32_t)(g - prefix[lo]);
    }
    // The matrix (at most 31 x 32 words) goes to LDS in one coalesced pass,
    // then each thread forms the patterns of its columns from there.
    const uint64_t* rows = words + M.word_off;
    uint64_t* sw = reinterpret_cast<uint64_t*>(s);
    const int nw = M.rows * M.words;
    for (int i = (int)threadIdx.x; i < nw; i += blockDim.x)
        sw[i] = rows[i];
    __syncthreads();
    uint32_t preg[kPatRegs];
    uint32_t active = 0;
#pragma unroll
    for (int k = 0; k < kPatRegs; ++k) {
        const int c = (int)threadIdx.x + 256 * k;
        uint32_t p = 0;
        if (c < M.cols) {
            const uint64_t* col = sw + (c >> 6);
            for (int i = 0; i < M.rows; ++i)
                p |= (uint32_t)((col[i * M.words] >> (c & 63)) & 1ull) << i;
        }
        preg[k] = p;
        active += p != 0u;
    }
    active = wave_sum_u32(active);
    if (threadIdx.x % kWave == 0)
        cnt[threadIdx.x / kWave] = active;
    __syncthreads();   // patterns read before the block build clears LDS
    const int32_t total = (int32_t)(cnt[0] + cnt[1] + cnt[2] + cnt[3]);
    switch (M.L) {
#define BMM_CASE(LL)                                                                           \
    case LL: coset_block_exact<LL>(s, M, total, mh, preg, keys); break;
        BMM_CASE(1) BMM_CASE(2) BMM_CASE(3) BMM_CASE(4) BMM_CASE(5) BMM_CASE(6) BMM_CASE(7)
        BMM_CASE(8) BMM_CASE(9) BMM_CASE(10) BMM_CASE(11) BMM_CASE(12) BMM_CASE(13)
#undef BMM_CASE
    default: break;
    }
    // Completion: each workgroup takes a ticket with release semantics after
    // its key updates; the last one (acquire) sees all of them.
    __syncthreads();
    if (threadIdx.x == 0)
        last = __hip_atomic_fetch_add(done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
               gridDim.x - 1;
    __syncthreads();
    if (!last)
        return;
    for (int slot = threadIdx.x; slot < n_slots; slot += blockDim.x) {
        const unsigned long long key =
            __hip_atomic_load(&keys[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int target = n_mat == 1 ? (int)(slot - first.out_off) : (slot_code[slot] & 31);
        const uint64_t mask = (uint32_t)key;
        const uint64_t below = mask & ((1ull << target) - 1ull);
        const uint64_t above = mask >> (target + 1);
        out_w[slot] = (uint32_t)(key >> 32);
        out_idx[slot] = below | (above << target);
        keys[slot] = ~0ull;
    }
    if (threadIdx.x == 0)
        *done = 0u;
}

int default_persistent_wgs();

// ------------------------------------------------------------- plan ------
struct BatchPlan {
    int device = -1;
    size_t n_mat = 0;
    int64_t n_slots = 0;
    int64_t total_blocks = 0;   // blocks of all matrices
    int64_t full_blocks = 0;    // blocks of the matrices with >= lmax rows
    int n_full = 0;             // those matrices (first in the sorted order)
    double subset_evals = 0.0;   // sum of r * 2^(r-1)
    double transform_elems = 0.0; // sum of 2^r
    int grid = 0;
    int chunk = 1;
    int lmax = kBatchMaxL;
    bool fused = false;         // single-launch form (coset_fused_kernel)
    const uint64_t* d_words = nullptr;
    uint32_t* d_out_w = nullptr;
    uint64_t* d_out_idx = nullptr;
    // slot_code[slot] = (index of the slot's matrix in the sorted order) << 5 | target row
    DevBuf mats, prefix, pats, keys, counter, slot_code, done;
    CosetMatrix first{};        // sorted matrix 0 (a one-matrix plan passes it by value)
    std::vector<int32_t> rows, cols;
    std::vector<int64_t> word_off, out_off;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~BatchPlan() {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        if (cap_stream)
            (void)hipStreamDestroy(cap_stream);
    }
};

namespace {

// Uploads descriptors for a given maximum block size.
void plan_upload(BatchPlan& P, int lmax) {
    P.lmax = lmax;
    const size_t n = P.n_mat;
    std::vector<CosetMatrix> mats(n);
    int64_t pat_off = 0;
    for (size_t m = 0; m < n; ++m) {
        CosetMatrix& M = mats[m];
        M.word_off = P.word_off[m];
        M.pat_off = pat_off;
        M.out_off = P.out_off[m];
        M.rows = P.rows[m];
        M.cols = P.cols[m];
        M.words = (P.cols[m] + 63) / 64;
        M.total = 0;
        M.L = std::min(P.rows[m], lmax);
        M.pad = 0;
        pat_off += std::max(P.cols[m], 1);
    }
    // LPT order: biggest blocks first, then most blocks.
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (mats[a].L != mats[b].L)
            return mats[a].L > mats[b].L;
        return mats[a].rows > mats[b].rows;
    });
    std::vector<CosetMatrix> sorted(n);
    std::vector<int64_t> prefix(n + 1, 0);
    int n_full = 0;
    for (size_t i = 0; i < n; ++i) {
        sorted[i] = mats[order[i]];
        prefix[i + 1] = prefix[i] + (int64_t(1) << (sorted[i].rows - sorted[i].L));
        if (sorted[i].L == lmax)
            n_full = (int)i + 1;
    }
    P.n_full = n_full;
    P.full_blocks = prefix[n_full];
    P.total_blocks = prefix[n];
    std::vector<int32_t> slot_code((size_t)P.n_slots, -1);
    for (size_t i = 0; i < n; ++i)
        for (int t = 0; t < sorted[i].rows; ++t) {
            int32_t& sm = slot_code[(size_t)(sorted[i].out_off + t)];
            if (sm != -1)
                fail(BMM_ERR_INVALID, "coset_min: overlapping output offsets");
            sm = (int32_t)(i << 5) | t;
        }
    P.first = sorted[0];
    for (int32_t sm : slot_code)
        if (sm < 0)
            fail(BMM_ERR_INVALID, "coset_min: output offsets leave gaps");
    CosetMatrix* d_m = P.mats.get<CosetMatrix>(n);
    int64_t* d_p = P.prefix.get<int64_t>(n + 1);
    int32_t* d_s = P.slot_code.get<int32_t>((size_t)P.n_slots);
    P.pats.get<uint32_t>((size_t)std::max<int64_t>(pat_off, 1));
    // Cleared once here; every pass's phase 2 leaves them cleared again.
    fill_u64(reinterpret_cast<uint64_t*>(P.keys.get<unsigned long long>((size_t)P.n_slots)),
             (size_t)P.n_slots, ~0ull, nullptr);
    fill_u64(reinterpret_cast<uint64_t*>(P.counter.get<unsigned long long>(1)), 1, 0ull, nullptr);
    BMM_HIP(hipMemcpy(d_m, sorted.data(), sizeof(CosetMatrix) * n, hipMemcpyHostToDevice));
    BMM_HIP(hipMemcpy(d_p, prefix.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    BMM_HIP(hipMemcpy(d_s, slot_code.data(), sizeof(int32_t) * slot_code.size(),
                      hipMemcpyHostToDevice));
}

void plan_enqueue(BatchPlan& P, hipStream_t s, bool timed) {
    CosetMatrix* d_m = static_cast<CosetMatrix*>(P.mats.ptr);
    int64_t* d_p = static_cast<int64_t*>(P.prefix.ptr);
    uint32_t* d_pat = static_cast<uint32_t*>(P.pats.ptr);
    unsigned long long* d_k = static_cast<unsigned long long*>(P.keys.ptr);
    unsigned long long* d_c = static_cast<unsigned long long*>(P.counter.ptr);
    int32_t* d_s = static_cast<int32_t*>(P.slot_code.ptr);
    if (P.fused) {
        ScopedKernelTimer timer(s, timed);
        hipLaunchKernelGGL(coset_fused_kernel, dim3((unsigned)P.total_blocks), dim3(256), 0, s,
                           P.d_words, P.first, d_m, d_p, (int)P.n_mat, d_s, (int)P.n_slots, d_k,
                           static_cast<unsigned int*>(P.done.ptr), P.d_out_w, P.d_out_idx);
        BMM_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(batch_patterns_kernel, dim3((unsigned)P.n_mat), dim3(256), 0, s, P.d_words,
                       d_m, d_pat);
    BMM_HIP(hipGetLastError());
    if (P.full_blocks > 0) {
        ScopedKernelTimer timer(s, timed);
        auto kern = P.lmax == kPkL ? coset_full_kernel<kPkL>
                    : P.lmax == 13 ? coset_full_kernel<13> : coset_full_kernel<12>;
        hipLaunchKernelGGL(kern, dim3(P.grid), dim3(256), 0, s, d_m, d_p, P.n_full,
                           P.full_blocks, P.chunk, d_pat, d_c, d_k);
        BMM_HIP(hipGetLastError());
    }
    if ((int64_t)P.n_mat > P.n_full) {
        hipLaunchKernelGGL(coset_small_kernel, dim3((unsigned)(P.n_mat - P.n_full)), dim3(256), 0,
                           s, d_m + P.n_full, d_pat, d_k);
        BMM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(coset_phase2_kernel, dim3((unsigned)P.n_slots), dim3(256), 0, s, d_m, d_s,
                       d_pat, d_k, d_c, P.d_out_w, P.d_out_idx);
    BMM_HIP(hipGetLastError());
}

// Times the sequence for each candidate launch shape and keeps the fastest.
void plan_autotune(BatchPlan& P) {
    int cus = default_persistent_wgs() / 8;
    struct Cand { int lmax, grid, chunk; };
    std::vector<Cand> cands;
    const bool pk = *std::max_element(P.cols.begin(), P.cols.end()) <= kPkMaxTotal;
    std::vector<int> lmaxes = {12, 13};
    if (pk)
        lmaxes.push_back(kPkL);
    for (int lmax : lmaxes)
        for (int per_cu : {4, 8})
            for (int chunk : {1, 8})
                cands.push_back({lmax, cus * per_cu, chunk});
    hipStream_t s;
    BMM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t a, b;
    BMM_HIP(hipEventCreate(&a));
    BMM_HIP(hipEventCreate(&b));
    float best = 1e30f;
    Cand bestc = {kBatchMaxL, cus * 8, 1};
    int cur_lmax = -1;
    for (const Cand& c : cands) {
        if (c.lmax != cur_lmax) {
            plan_upload(P, c.lmax);
            cur_lmax = c.lmax;
        }
        P.grid = (int)std::max<int64_t>(1, std::min<int64_t>(c.grid, P.full_blocks));
        P.chunk = c.chunk;
        plan_enqueue(P, s, false);   // warm
        // Best of several timed passes (one pass for passes of >= 5 ms, up to
        // 16 for short ones, where launch jitter dominates a single sample).
        float ms = 1e30f;
        for (int rep = 0, reps = 1; rep < reps; ++rep) {
            BMM_HIP(hipEventRecord(a, s));
            plan_enqueue(P, s, false);
            BMM_HIP(hipEventRecord(b, s));
            BMM_HIP(hipEventSynchronize(b));
            float t = 0.f;
            BMM_HIP(hipEventElapsedTime(&t, a, b));
            ms = std::min(ms, t);
            if (rep == 0 && t < 5.f)
                reps = 16;
        }
        if (ms < best) {
            best = ms;
            bestc = c;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    BMM_HIP(hipStreamSynchronize(s));
    (void)hipStreamDestroy(s);
    if (bestc.lmax != cur_lmax)
        plan_upload(P, bestc.lmax);
    P.grid = (int)std::max<int64_t>(1, std::min<int64_t>(bestc.grid, P.full_blocks));
    P.chunk = bestc.chunk;
}

// Untuned block size: the largest of 14 (packed int16, when every matrix has
// at most 32767 columns), 13, 12 that still gives at least two blocks per
// CU; else 12.
int default_block_bits(const BatchPlan& P) {
    const int64_t want = 2 * (default_persistent_wgs() / 8);
    const bool pk = *std::max_element(P.cols.begin(), P.cols.end()) <= kPkMaxTotal;
    for (int bits : {kPkL, 13, 12}) {
        if (bits == kPkL && !pk)
            continue;
        int64_t blocks = 0;
        for (size_t m = 0; m < P.n_mat; ++m)
            if (P.rows[m] >= bits)
                blocks += int64_t(1) << (P.rows[m] - bits);
        if (blocks >= want)
            return bits;
    }
    return 12;
}

// Launch shapes chosen by autotuning, keyed by the batch's shape signature.
std::map<std::vector<int64_t>, std::tuple<int, int, int>>& tune_cache() {
    static std::map<std::vector<int64_t>, std::tuple<int, int, int>> c;
    return c;
}

} // namespace

BatchPlan* batch_plan_create(const uint64_t* d_words, const int64_t* word_offset,
                             const int32_t* rows, const int32_t* cols, size_t n_mat,
                             const int64_t* out_offset, uint32_t* d_out_w, uint64_t* d_out_idx,
                             int persistent_wgs, int autotune) {
    if (n_mat == 0)
        fail(BMM_ERR_INVALID, "coset_min: empty batch");
    auto P = std::make_unique<BatchPlan>();
    P->n_mat = n_mat;
    P->d_words = d_words;
    P->d_out_w = d_out_w;
    P->d_out_idx = d_out_idx;
    P->rows.assign(rows, rows + n_mat);
    P->cols.assign(cols, cols + n_mat);
    P->word_off.assign(word_offset, word_offset + n_mat);
    P->out_off.assign(out_offset, out_offset + n_mat);
    std::vector<int64_t> sig;
    for (size_t m = 0; m < n_mat; ++m) {
        if (rows[m] < 2 || rows[m] > 31)
            fail(BMM_ERR_INVALID, "coset_min: rows must be in [2, 31]");
        if (cols[m] < 0 || word_offset[m] < 0 || out_offset[m] < 0)
            fail(BMM_ERR_INVALID, "coset_min: negative size or offset");
        P->n_slots = std::max<int64_t>(P->n_slots, out_offset[m] + rows[m]);
        P->subset_evals += (double)rows[m] * std::ldexp(1.0, rows[m] - 1);
        P->transform_elems += std::ldexp(1.0, rows[m]);
        sig.push_back(rows[m]);
        sig.push_back(cols[m]);
    }
    BMM_HIP(hipGetDevice(&P->device));
    // Small batches: the single-launch form (no tuning needed).
    int64_t fused_blocks = 0;
    bool narrow = true;
    for (size_t m = 0; m < n_mat; ++m) {
        fused_blocks += int64_t(1) << (rows[m] - std::min(rows[m], kFusedL));
        narrow = narrow && cols[m] <= kPatRegs * 256;
    }
    if (narrow && fused_blocks <= kFusedMaxBlocks) {
        plan_upload(*P, kFusedL);
        fill_u32(P->done.get<uint32_t>(1), 1, 0u, nullptr);
        P->fused = true;
        P->grid = (int)P->total_blocks;
        P->chunk = 1;
        return P.release();
    }
    auto it = tune_cache().find(sig);
    if (autotune && it == tune_cache().end()) {
        plan_autotune(*P);
        tune_cache()[sig] = {P->lmax, P->grid, P->chunk};
    } else {
        int lmax = default_block_bits(*P), grid = persistent_wgs ? persistent_wgs
                                                                 : default_persistent_wgs();
        int chunk = 1;
        if (it != tune_cache().end() && persistent_wgs == 0)
            std::tie(lmax, grid, chunk) = it->second;
        plan_upload(*P, lmax);
        P->grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, P->full_blocks));
        if (it == tune_cache().end() || persistent_wgs != 0)
            chunk = P->full_blocks > (int64_t)P->grid * 64 ? 8 : 1;
        P->chunk = chunk;
    }
    return P.release();
}

void batch_plan_run(BatchPlan* P, hipStream_t stream, bool use_graph) {
    // A single-launch plan gains nothing from a one-node graph, whose launch
    // costs more than the kernel's own (about 7 us per pass for 20 x 128).
    if (!use_graph || P->fused) {
        plan_enqueue(*P, stream, true);
        return;
    }
    if (!P->exec) {
        BMM_HIP(hipStreamCreateWithFlags(&P->cap_stream, hipStreamNonBlocking));
        BMM_HIP(hipStreamBeginCapture(P->cap_stream, hipStreamCaptureModeThreadLocal));
        plan_enqueue(*P, P->cap_stream, false);
        BMM_HIP(hipStreamEndCapture(P->cap_stream, &P->graph));
        BMM_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    }
    ScopedKernelTimer timer(stream);
    BMM_HIP(hipGraphLaunch(P->exec, stream));
}

void batch_plan_set_shape(BatchPlan* P, int max_block_bits, int persistent_wgs, int chunk) {
    if (max_block_bits != 12 && max_block_bits != 13 && max_block_bits != kPkL)
        fail(BMM_ERR_INVALID, "batch_plan_set_shape: max_block_bits must be 12, 13 or 14");
    if (max_block_bits == kPkL &&
        *std::max_element(P->cols.begin(), P->cols.end()) > kPkMaxTotal)
        fail(BMM_ERR_INVALID,
             "batch_plan_set_shape: 14-bit (int16) blocks need at most 32767 columns");
    if (persistent_wgs < 0 || chunk < 0)
        fail(BMM_ERR_INVALID, "batch_plan_set_shape: negative launch size");
    P->fused = false;
    if (P->exec) {
        BMM_HIP(hipStreamSynchronize(P->cap_stream));
        BMM_HIP(hipGraphExecDestroy(P->exec));
        BMM_HIP(hipGraphDestroy(P->graph));
        P->exec = nullptr;
        P->graph = nullptr;
    }
    plan_upload(*P, max_block_bits);
    const int grid = persistent_wgs ? persistent_wgs : default_persistent_wgs();
    P->grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, P->full_blocks));
    P->chunk = chunk ? chunk : (P->full_blocks > (int64_t)P->grid * 64 ? 8 : 1);
}

void batch_plan_destroy(BatchPlan* P) { delete P; }

void batch_plan_info(const BatchPlan* P, bmm_batch_info* out) {
    out->n_matrices = P->n_mat;
    out->n_slots = (size_t)P->n_slots;
    out->total_blocks = (size_t)P->total_blocks;
    out->persistent_wgs = P->grid;
    out->chunk = P->chunk;
    out->max_block_bits = P->lmax;
    out->subset_evals = P->subset_evals;
    out->transform_elements = P->transform_elems;
}

// One-shot form used by the plain entry points.
void coset_min_batch_dev(const uint64_t* d_rows, const int64_t* word_offset, const int32_t* rows,
                         const int32_t* cols, size_t n_mat, const int64_t* out_offset,
                         uint32_t* d_out_w, uint64_t* d_out_idx, int persistent_wgs,
                         hipStream_t stream) {
    std::unique_ptr<BatchPlan> P(batch_plan_create(d_rows, word_offset, rows, cols, n_mat,
                                                   out_offset, d_out_w, d_out_idx,
                                                   persistent_wgs, 0));
    batch_plan_run(P.get(), stream, false);
    BMM_HIP(hipStreamSynchronize(stream));
}

} // namespace bmmhip
