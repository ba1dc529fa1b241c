// Exact top-k searches over a row window on gfx950: FwhtSearch
// (src/search/fwht_search.cpp:34-132), the MITM seam over a CompactSplitWindow
// (run_cuda_mitm_fwht_search, src/search/cuda_mitm_fwht_launch.cu:460-496)
// and the brute-force seam for widths the Gray kernel does not hold in
// registers.
//
// All three reduce to one engine: a multiset of (pattern, multiplicity) with
// weight(m) = (total - H[m]) / 2, H = FWHT of the pattern histogram.  Masks are
// split m = mh << L | ml; each persistent workgroup takes blocks mh from a
// device counter, builds g_mh in LDS (sign (-1)^popcount((p >> L) & mh) per
// pattern), transforms 2^L entries on chip and offers the packed keys
// (weight << t | mask) to its exact top-K (topk.hpp).  A tree of parallel
// merges (16 lists per workgroup and level) folds the per-workgroup lists.
// The split point does not change the
// result: the top-K under the strict (weight, mask) order is unique.
#include "common.hpp"
#include "fwht_block.hpp"
#include "fwht_pk16.hpp"
#include "topk.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace bmmhip {

// g[pl] = sum_i mult[i] (-1)^popcount((pat[i] >> L) & mh) at pl = pat[i] mod 2^L.
template <int L>
__device__ inline void block_build_g64(int32_t* s, const uint64_t* __restrict__ pat,
                                       const int32_t* __restrict__ mult, int n, uint64_t mh) {
    constexpr int N = 1 << L;
    for (int i = threadIdx.x; i < fwht_lds_words(L); i += blockDim.x)
        s[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const uint64_t p = pat[i];
        const int32_t m = mult[i];
        const int neg = __builtin_popcountll((p >> L) & mh) & 1;
        atomicAdd(&s[fwht_pad((int)(p & (N - 1)))], neg ? -m : m);
    }
}

template <int L, int CAP>
__global__ __launch_bounds__(256) void fwht_topk_kernel(const uint64_t* __restrict__ pat,
                                                        const int32_t* __restrict__ mult,
                                                        int n_pat, int t, int32_t total,
                                                        uint64_t n_blocks, int K,
                                                        unsigned long long* __restrict__ counter,
                                                        const unsigned long long* __restrict__ bound,
                                                        uint64_t* __restrict__ lists) {
    __shared__ int32_t s[fwht_lds_words(L)];
    __shared__ TopkShared<CAP> T;
    __shared__ unsigned long long blk_id;
    constexpr int N = 1 << L;
    constexpr int E = N >= 256 ? N / 256 : 1;
    topk_init(T, bound ? *bound : ~0ull);
    for (;;) {
        if (threadIdx.x == 0)
            blk_id = atomicAdd(counter, 1ull);
        __syncthreads();
        const uint64_t mh = blk_id;
        if (mh >= n_blocks)
            break;
        block_build_g64<L>(s, pat, mult, n_pat, mh);
        block_fwht<L>(s);
        // 32-bit prefilter: a key below the list threshold has weight at most
        // thr >> t, i.e. H >= total - 2 (thr >> t).  The threshold only
        // decreases, so reading it here is conservative.
        const uint64_t thr = T.thr;
        const int64_t wthr = thr == ~0ull ? INT64_MAX : (int64_t)(thr >> t);
        const int32_t hmin = wthr >= (int64_t)total ? INT32_MIN
                                                    : (int32_t)((int64_t)total - 2 * wthr);
        // Entry ml = e * 256 + tid sits at padded index pad(tid) + 264 e.
        const int32_t* sp = s + fwht_pad((int)threadIdx.x);
        uint32_t valid = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int ml = e * 256 + (int)threadIdx.x;
            if (ml < N && (mh != 0 || ml != 0) && sp[e * 264] >= hmin)
                valid |= 1u << e;
        }
        // The transform stays in LDS until the next block's build, so a
        // candidate's key is recomputed from it instead of held in registers.
        topk_offer_fn<CAP, E>(T, valid, K, [&](int e) {
            const uint64_t w = (uint64_t)(uint32_t)(total - sp[e * 264]) >> 1;
            return (w << t) | (mh << L) | (uint64_t)(e * 256 + (int)threadIdx.x);
        });
    }
    topk_store(T, lists + (size_t)blockIdx.x * K, K);
}

// Packed pass over a block of 2^14 masks (windows of >= 14 rows and at most
// 32767 active columns): g_mh built into biased int16 halves (a low-half
// entry adds +-mult to the word, a high-half entry +-mult << 16; the bias
// keeps every half in range, so no carry crosses halves) and transformed with
// the packed radix rounds.  Each thread keeps, from its 64 final values, the
// keys of the entries with H >= hmin (weight at most the list threshold's) in
// up to kPkCand registers; the block falls back to the exact int32 pass when
// any thread has more (warm-up, or a block crowded with light masks).
constexpr int kPkCand = 4;

// A column pattern prepared for register-held builds: hi = pattern >> 14
// (windows of at most 46 rows), av = mult << 16 | LDS byte address of its
// word | 1 for the high half (addresses are multiples of 4); av = 0 pads.
__device__ inline void pk_prep_entry(uint64_t p, int32_t m, uint32_t& hi, uint32_t& av) {
    const uint32_t pl = (uint32_t)p & ((1u << kPkL) - 1u);
    hi = (uint32_t)(p >> kPkL);
    av = ((uint32_t)m << 16) | ((uint32_t)pk_pad((int)(pl & ((1u << kPkW) - 1u))) * 4u) |
         (pl >> kPkW);
}

template <int R>
__device__ inline void pk_build_regs(uint32_t* s, const uint32_t (&hi)[R], const uint32_t (&av)[R],
                                     uint32_t mh) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const uint32_t a = av[k];
        if (a != 0u) {
            const uint32_t m = a >> 16;
            const uint32_t add = (a & 1u) ? (m << 16) : m;
            const int neg = __builtin_popcount(hi[k] & mh) & 1;
            atomicAdd(reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(s) + (a & 0xfffcu)),
                      neg ? 0u - add : add);
        }
    }
}

__device__ inline void pk_build_mem64(uint32_t* s, const uint64_t* __restrict__ pat,
                                      const int32_t* __restrict__ mult, int n, uint64_t mh) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const uint64_t p = pat[i];
        const uint32_t pl = (uint32_t)p & ((1u << kPkL) - 1u);
        const uint32_t m = (uint32_t)mult[i];
        const uint32_t add = (pl >> kPkW) ? (m << 16) : m;
        const int neg = __builtin_popcountll((p >> kPkL) & mh) & 1;
        atomicAdd(&s[pk_pad((int)(pl & ((1u << kPkW) - 1u)))], neg ? 0u - add : add);
    }
}

// One packed block of 2^14 masks: every stage but those on word bits 5-9
// runs in LDS; each thread then holds one group of 32 words over bits 5-9,
// whose outputs are signed sums x_0 +- x_1 ... +- x_31, so
// x_0 + sum_{e >= 1} |x_e| bounds them (at most the active column count: no
// int16 overflow).  A thread whose bound is below hmin has no candidate and
// stops; the others finish the five stages in registers and take the entries
// with H >= hmin.  Returns whether any thread had more than kPkCand.
template <typename Build>
__device__ inline bool pk_topk_block(uint32_t* s, Build&& build, uint64_t mh, int t, int32_t total,
                                     int32_t hmin, uint64_t (&cand)[kPkCand], int& nc) {
    int4* s4 = reinterpret_cast<int4*>(s);
    const int b = (int)kPkBias;
    for (int i = threadIdx.x; i < (pk_lds_words() + 3) / 4; i += blockDim.x)
        s4[i] = make_int4(b, b, b, b);
    __syncthreads();
    build();
    nc = 0;
    bool over = false;
    const uint64_t base = mh << kPkL;
    auto take = [&](int32_t h, uint32_t ml) {
        if (h >= hmin && (mh != 0 || ml != 0)) {
            if (nc < kPkCand) {
                const uint64_t key = ((uint64_t)(uint32_t)(total - h) >> 1 << t) | base | ml;
#pragma unroll
                for (int k = 0; k < kPkCand; ++k)   // static register slots
                    if (k == nc)
                        cand[k] = key;
                ++nc;
            } else {
                over = true;
            }
        }
    };
    __syncthreads();
    pk_round<0, 5, true>(s);
    __syncthreads();
    pk_round_top(s);
    __syncthreads();
    {
        // Thread t's group over word bits 5-9: words (t >> 5) << 10 | e << 5
        // | (t & 31), e < 32.
        constexpr int ST = pk_pad_stride<5>();
        const uint32_t w0 = ((threadIdx.x >> 5) << 10) | (threadIdx.x & 31u);
        const uint32_t* p = s + pk_pad((int)w0);
        uint32_t v[32];
#pragma unroll
        for (int e = 0; e < 32; ++e)
            v[e] = p[e * ST];
        // x_0 + sum |x_e| bounds every output of the group (four partial
        // sums: independent chains).
        s16x2 acc[4] = {pk_as(v[0]), {0, 0}, {0, 0}, {0, 0}};
#pragma unroll
        for (int e = 1; e < 32; ++e) {
            const s16x2 x = pk_as(v[e]);
            acc[e & 3] = acc[e & 3] + pk_max(x, -x);
        }
        const s16x2 bound = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        if (max(pk_lo(bound), pk_hi(bound)) >= hmin) {
            pk_butterflies(v);
            // The group's largest output (a packed max tree) decides whether
            // any entry needs the per-entry test.
            s16x2 m[16];
#pragma unroll
            for (int e = 0; e < 16; ++e)
                m[e] = pk_max(pk_as(v[e]), pk_as(v[e + 16]));
#pragma unroll
            for (int h = 8; h > 0; h >>= 1)
#pragma unroll
                for (int e = 0; e < h; ++e)
                    m[e] = pk_max(m[e], m[e + h]);
            if (max(pk_lo(m[0]), pk_hi(m[0])) >= hmin) {
#pragma unroll
                for (int e = 0; e < 32; ++e) {
                    const s16x2 x = pk_as(v[e]);
                    if (max(pk_lo(x), pk_hi(x)) >= hmin) {
                        const uint32_t w = w0 | ((uint32_t)e << 5);
                        take(pk_lo(x), w);
                        take(pk_hi(x), w + (1u << kPkW));
                    }
                }
            }
        }
    }
    return __syncthreads_or(over) != 0;
}

// Blocks of 2^14 masks, taken `chunk` at a time from the work counter (up to
// kPkChunk, fewer when each workgroup gets only a few blocks): the
// packed pass offers its few candidates directly (nothing when no thread has
// one); a block with more candidates than the registers hold is built again
// as two exact int32 blocks of 2^13 and offered in full.
constexpr int kPkChunk = 16;

// R > 0: each thread holds patterns tid + 256 k (k < R) in registers across
// its blocks (windows of at most 46 rows and 256 R patterns); R = 0 reads
// them from memory for every block.
// Waves per SIMD the packed top-k kernel is compiled for: four with a
// 256-key list workspace (39 KiB of LDS: 4 workgroups per CU), three
// otherwise.
template <int CAP> constexpr int kTopkPkWaves = CAP <= 256 ? 4 : 3;

template <int CAP, int R>
__global__ __launch_bounds__(256, kTopkPkWaves<CAP>) void fwht_topk_pk_kernel(const uint64_t* __restrict__ pat,
                                                           const int32_t* __restrict__ mult,
                                                           int n_pat, int t, int32_t total,
                                                           uint64_t n_blocks, int K, int chunk,
                                                           unsigned long long* __restrict__ counter,
                                                           const unsigned long long* __restrict__ bound,
                                                           uint64_t* __restrict__ lists) {
    constexpr int L = 13;
    constexpr int kWords = pk_lds_words() > fwht_lds_words(L) ? pk_lds_words() : fwht_lds_words(L);
    __shared__ __align__(16) int32_t s[kWords + 4];
    __shared__ TopkShared<CAP> T;
    __shared__ unsigned long long blk_id;
    constexpr int E = (1 << L) / 256;
    constexpr int RR = R > 0 ? R : 1;
    uint32_t hi[RR], av[RR];
#pragma unroll
    for (int k = 0; k < RR; ++k) {
        const int i = (int)threadIdx.x + 256 * k;
        hi[k] = 0u;
        av[k] = 0u;
        if (R > 0 && i < n_pat)
            pk_prep_entry(pat[i], mult[i], hi[k], av[k]);
    }
    topk_init(T, bound ? *bound : ~0ull);
    uint64_t next = 0, stop = 0;
    for (;;) {
        if (next == stop) {
            if (threadIdx.x == 0)
                blk_id = atomicAdd(counter, (unsigned long long)chunk);
            __syncthreads();
            next = blk_id;
            __syncthreads();
            if (next >= n_blocks)
                break;
            stop = min(next + (uint64_t)chunk, n_blocks);
        }
        const uint64_t mh14 = next++;
        __syncthreads();   // the previous block's LDS reads are done
        const uint64_t thr = T.thr;
        const int64_t wthr = thr == ~0ull ? INT64_MAX : (int64_t)(thr >> t);
        const int32_t hmin = wthr >= (int64_t)total ? INT32_MIN
                                                    : (int32_t)((int64_t)total - 2 * wthr);
        uint64_t cand[kPkCand];
        int nc = 0;
        uint32_t* su = reinterpret_cast<uint32_t*>(s);
        bool over;
        if constexpr (R > 0)
            over = pk_topk_block(
                su, [&] { pk_build_regs<RR>(su, hi, av, (uint32_t)mh14); }, mh14, t, total, hmin,
                cand, nc);
        else
            over = pk_topk_block(
                su, [&] { pk_build_mem64(su, pat, mult, n_pat, mh14); }, mh14, t, total, hmin,
                cand, nc);
        if (!over) {
            if (__syncthreads_or(nc > 0))
                topk_offer_fn<CAP, kPkCand>(T, (1u << nc) - 1u, K,
                                            [&](int e) { return cand[e]; });
            continue;
        }
        for (int j = 0; j < 2; ++j) {
            const uint64_t mh = 2 * mh14 + j;
            __syncthreads();
            block_build_g64<L>(s, pat, mult, n_pat, mh);
            block_fwht<L>(s);
            const uint64_t thr2 = T.thr;
            const int64_t wthr2 = thr2 == ~0ull ? INT64_MAX : (int64_t)(thr2 >> t);
            const int32_t hmin2 = wthr2 >= (int64_t)total ? INT32_MIN
                                                         : (int32_t)((int64_t)total - 2 * wthr2);
            const int32_t* sp = s + fwht_pad((int)threadIdx.x);
            // Offered 8 entries per thread at a time (fewer live registers).
            for (int c = 0; c < E; c += 8) {
                uint32_t valid = 0;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int ml = (c + e) * 256 + (int)threadIdx.x;
                    if ((mh != 0 || ml != 0) && sp[(c + e) * 264] >= hmin2)
                        valid |= 1u << e;
                }
                topk_offer_fn<CAP, 8>(T, valid, K, [&](int e) {
                    const uint64_t w = (uint64_t)(uint32_t)(total - sp[(c + e) * 264]) >> 1;
                    return (w << t) | (mh << L) | (uint64_t)((c + e) * 256 + (int)threadIdx.x);
                });
            }
        }
    }
    topk_store(T, lists + (size_t)blockIdx.x * K, K);
}

// A first pass for searches whose workgroups see few blocks each (a window
// of 22-30 rows over a persistent grid): every workgroup's list would start
// empty and its first blocks would all be crowded.  Instead, the sampled
// blocks 0, stride, 2 stride, ... are transformed (one per workgroup) and
// their masks counted by weight bin (bin = weight >> shift, so the bins span
// [0, total]); from the summed counts, with b the first bin at which the
// running count reaches K, at least K masks have weight below (b + 1) <<
// shift, so every key of the exact top-K is below ((b + 1) << shift) << t.
// A workgroup counts only the kBoundSpan bins from its block's lightest mask
// up (8-bit counters packed in one register per thread, summed per
// wavefront): fewer counts only weaken the bound, never invalidate it.  Each
// workgroup stores its counts in its own slot (same-address device atomics
// from many workgroups serialize at memory), and topk_bound_finish_kernel
// sums the slots.
constexpr int kBoundBins = 1024;
constexpr int kBoundSpan = 8;
constexpr int kBoundSample = 1024;   // sampled blocks of 2^14 masks at most

struct TopkBoundSlot {
    int32_t b0;                    // bin of the block's lightest mask
    uint32_t count[kBoundSpan];    // masks in bins b0 .. b0 + kBoundSpan - 1
};

struct TopkBoundState {   // header zeroed before every search
    unsigned long long counter;   // work queue of the main pass
    unsigned long long bound;     // written by topk_bound_finish_kernel
    TopkBoundSlot slot[kBoundSample];
};

// Counts one block's masks into its slot: h[i] for the NV values of this
// thread with ok[i] set (mask 0 excluded by the caller).
template <int NV>
__device__ inline void bound_count_block(const int32_t (&h)[NV], const bool (&ok)[NV],
                                         int32_t total, int shift,
                                         TopkBoundSlot& __restrict__ slot) {
    __shared__ uint32_t lh[4][kBoundSpan];
    __shared__ int32_t red[4];
    const int tid = (int)threadIdx.x;
    int32_t hmax = INT32_MIN;
#pragma unroll
    for (int i = 0; i < NV; ++i)
        if (ok[i])
            hmax = max(hmax, h[i]);
    hmax = wave_max_i32(hmax);
    if (tid % kWave == 0)
        red[tid / kWave] = hmax;
    __syncthreads();
    hmax = max(max(red[0], red[1]), max(red[2], red[3]));
    // A block with no mask (only mask 0) counts nothing.
    const int32_t b0 = hmax == INT32_MIN ? kBoundBins : ((total - hmax) >> 1) >> shift;
    // Counter d (bits 8d..8d+7) counts the thread's masks in bin b0 + d; at
    // most 64 masks per thread, so no counter overflows.
    static_assert(NV <= 255, "8-bit counters");
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int32_t d = (((total - h[i]) >> 1) >> shift) - b0;
        if (ok[i] && d < kBoundSpan)
            acc += 1ull << (8 * d);
    }
#pragma unroll
    for (int d = 0; d < kBoundSpan; ++d) {
        const uint32_t c = wave_sum_u32((uint32_t)(acc >> (8 * d)) & 0xffu);
        if (tid % kWave == 0)
            lh[tid / kWave][d] = c;
    }
    __syncthreads();
    if (tid == 0)
        slot.b0 = b0;
    if (tid < kBoundSpan)
        slot.count[tid] = lh[0][tid] + lh[1][tid] + lh[2][tid] + lh[3][tid];
}

// Packed blocks of 2^14 masks (the packed top-k path).
__global__ __launch_bounds__(256) void topk_bound_pk_kernel(const uint64_t* __restrict__ pat,
                                                            const int32_t* __restrict__ mult,
                                                            int n_pat, int32_t total,
                                                            uint64_t stride, int shift,
                                                            TopkBoundState* __restrict__ st) {
    __shared__ __align__(16) uint32_t s[pk_lds_words() + 4];
    const int tid = (int)threadIdx.x;
    const uint64_t mh14 = (uint64_t)blockIdx.x * stride;
    int4* s4 = reinterpret_cast<int4*>(s);
    const int b = (int)kPkBias;
    for (int i = tid; i < (pk_lds_words() + 3) / 4; i += blockDim.x)
        s4[i] = make_int4(b, b, b, b);
    __syncthreads();
    pk_build_mem64(s, pat, mult, n_pat, mh14);
    __syncthreads();
    pk_round<0, 5, true>(s);
    __syncthreads();
    pk_round_top(s);
    __syncthreads();
    // Thread t's group over word bits 5-9, finished in registers: entries
    // w and w + 2^13 for w = (t >> 5) << 10 | e << 5 | (t & 31).
    uint32_t v[32];
    pk_mid_group_visit(s, [&](const uint32_t(&x)[32]) {
#pragma unroll
        for (int e = 0; e < 32; ++e)
            v[e] = x[e];
    });
    pk_butterflies(v);
    int32_t h[64];
    bool ok[64];
#pragma unroll
    for (int e = 0; e < 32; ++e) {
        const s16x2 x = pk_as(v[e]);
        h[2 * e] = pk_lo(x);
        h[2 * e + 1] = pk_hi(x);
        ok[2 * e] = !(mh14 == 0 && tid == 0 && e == 0);   // mask 0 is never a result
        ok[2 * e + 1] = true;
    }
    bound_count_block(h, ok, total, shift, st->slot[blockIdx.x]);
}

// int32 blocks of 2^L masks (the generic top-k path).
template <int L>
__global__ __launch_bounds__(256) void topk_bound_i32_kernel(const uint64_t* __restrict__ pat,
                                                             const int32_t* __restrict__ mult,
                                                             int n_pat, int32_t total,
                                                             uint64_t stride, int shift,
                                                             TopkBoundState* __restrict__ st) {
    __shared__ int32_t s[fwht_lds_words(L)];
    constexpr int N = 1 << L;
    constexpr int E = N >= 256 ? N / 256 : 1;
    const int tid = (int)threadIdx.x;
    const uint64_t mh = (uint64_t)blockIdx.x * stride;
    block_build_g64<L>(s, pat, mult, n_pat, mh);
    block_fwht<L>(s);
    // Entry ml = e * 256 + tid sits at padded index pad(tid) + 264 e.
    const int32_t* sp = s + fwht_pad(tid);
    int32_t h[E];
    bool ok[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int ml = e * 256 + tid;
        ok[e] = ml < N && (mh != 0 || ml != 0);
        h[e] = ml < N ? sp[e * 264] : 0;
    }
    bound_count_block(h, ok, total, shift, st->slot[blockIdx.x]);
}

// One workgroup: sums the n_slots slots into bins and writes the key bound
// (~0 when fewer than K masks were counted).
__global__ __launch_bounds__(256) void topk_bound_finish_kernel(int n_slots, int shift, int t,
                                                                int K,
                                                                TopkBoundState* __restrict__ st) {
    __shared__ uint32_t hist[kBoundBins];
    __shared__ uint32_t scan[256];
    __shared__ int32_t first_bin;
    const int tid = (int)threadIdx.x;
    for (int i = tid; i < kBoundBins; i += blockDim.x)
        hist[i] = 0u;
    if (tid == 0)
        first_bin = INT32_MAX;
    __syncthreads();
    // Thread t takes slots t + 256 j (j < 4): all loads in flight at once.
    constexpr int SPT = kBoundSample / 256;
    int32_t b0[SPT];
    uint32_t cnt[SPT][kBoundSpan];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int i = tid + 256 * j;
        b0[j] = kBoundBins;   // nothing counted
        if (i < n_slots) {
            b0[j] = st->slot[i].b0;
#pragma unroll
            for (int d = 0; d < kBoundSpan; ++d)
                cnt[j][d] = st->slot[i].count[d];
        }
    }
#pragma unroll
    for (int j = 0; j < SPT; ++j)
#pragma unroll
        for (int d = 0; d < kBoundSpan; ++d)
            if (b0[j] + d < kBoundBins && cnt[j][d] != 0u)
                atomicAdd(&hist[b0[j] + d], cnt[j][d]);
    __syncthreads();
    // Running counts, four bins per thread.
    constexpr int PER = kBoundBins / 256;
    uint32_t c[PER], sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        c[i] = hist[tid * PER + i];
        sum += c[i];
    }
    scan[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {   // inclusive scan of the thread sums
        const uint32_t add = tid >= off ? scan[tid - off] : 0u;
        __syncthreads();
        scan[tid] += add;
        __syncthreads();
    }
    uint32_t run = scan[tid] - sum;   // count before this thread's bins
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        run += c[i];
        if (run >= (uint32_t)K) {
            atomicMin(&first_bin, tid * PER + i);
            break;
        }
    }
    __syncthreads();
    if (tid == 0) {
        const int32_t bin = first_bin;
        st->bound = bin == INT32_MAX
                        ? ~0ull
                        : (unsigned long long)((uint64_t)(bin + 1) << shift) << t;
    }
}

// u64 column patterns of a t-row window (t <= 64) over `cols` columns.
__global__ void column_patterns64_kernel(const uint64_t* __restrict__ rows, int t, int words,
                                         int cols, uint64_t* __restrict__ pat,
                                         int32_t* __restrict__ mult) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols)
        return;
    const int wi = c >> 6, bit = c & 63;
    uint64_t p = 0;
    for (int i = 0; i < t; ++i)
        p |= ((rows[(size_t)i * words + wi] >> bit) & 1ull) << i;
    pat[c] = p;
    mult[c] = p != 0 ? 1 : 0;
}

namespace {
struct SearchWorkspace {
    DevBuf rows, pat, mult, counter, lists, lists2, out, bound;
    HostBuf stage, keys;   // page-locked: inputs sent and the list received asynchronously
};

// Folds n_lists lists of K keys into one by a tree of parallel merges;
// returns the buffer holding the final list.
constexpr int kMergeFan = 16;

template <int CAP>
uint64_t* merge_lists(uint64_t* a, uint64_t* b, int n_lists, int K, hipStream_t s) {
    // Groups of at most kMergeSortedKeys keys (at a power-of-two list stride)
    // merge by pairwise rounds, up to 64 lists per group; longer lists by
    // offers to a top-K list.
    int lk = 0;
    while ((1 << lk) < K)
        ++lk;
    const int sorted_fan = std::min(64, kMergeSortedKeys >> lk);
    while (n_lists > 1) {
        const int fan = sorted_fan >= 2 ? sorted_fan : kMergeFan;
        const int groups = (n_lists + fan - 1) / fan;
        if (sorted_fan >= 2)
            hipLaunchKernelGGL(topk_merge_sorted_kernel, dim3(groups), dim3(kMergeSortedThreads),
                               0, s, a, n_lists, fan, K, lk, b);
        else
            hipLaunchKernelGGL(topk_merge_group_kernel<CAP>, dim3(groups), dim3(256), 0, s, a,
                               n_lists, fan, K, b);
        BMM_HIP(hipGetLastError());
        std::swap(a, b);
        n_lists = groups;
    }
    return a;
}
SearchWorkspace& search_ws() {
    thread_local SearchWorkspace ws;
    return ws;
}

constexpr int kTopkMaxL = 13;

// The int32 block pass, after the sampled bound pass when the blocks are few
// per workgroup (or all in the sample).
template <int L, int CAP>
void launch_fwht_topk(const uint64_t* pat, const int32_t* mult, int n_pat, int t, int32_t total,
                      uint64_t n_blocks, int K, unsigned long long* counter, uint64_t* lists,
                      int grid, TopkBoundState* st, hipStream_t s) {
    ScopedKernelTimer timer(s);
    const unsigned long long* bound = nullptr;
    if (st) {
        const uint64_t sample = std::min<uint64_t>(n_blocks, kBoundSample);
        int shift = 0;
        while ((total >> shift) >= kBoundBins)
            ++shift;
        hipLaunchKernelGGL((topk_bound_i32_kernel<L>), dim3((unsigned)sample), dim3(256), 0, s,
                           pat, mult, n_pat, total, n_blocks / sample, shift, st);
        BMM_HIP(hipGetLastError());
        hipLaunchKernelGGL(topk_bound_finish_kernel, dim3(1), dim3(256), 0, s, (int)sample, shift,
                           t, K, st);
        BMM_HIP(hipGetLastError());
        bound = &st->bound;
    }
    hipLaunchKernelGGL((fwht_topk_kernel<L, CAP>), dim3(grid), dim3(256), 0, s, pat, mult, n_pat,
                       t, total, n_blocks, K, counter, bound, lists);
    BMM_HIP(hipGetLastError());
}

template <int CAP>
void dispatch_fwht_topk(int L, const uint64_t* pat, const int32_t* mult, int n_pat, int t,
                        int32_t total, uint64_t n_blocks, int K, unsigned long long* counter,
                        uint64_t* lists, int grid, TopkBoundState* st, hipStream_t s) {
    switch (L) {
#define BMM_L(LL)                                                                              \
    case LL:                                                                                   \
        launch_fwht_topk<LL, CAP>(pat, mult, n_pat, t, total, n_blocks, K, counter, lists, grid, \
                                  st, s);                                                      \
        break;
        BMM_L(1) BMM_L(2) BMM_L(3) BMM_L(4) BMM_L(5) BMM_L(6) BMM_L(7) BMM_L(8) BMM_L(9)
        BMM_L(10) BMM_L(11) BMM_L(12) BMM_L(13)
#undef BMM_L
    default: fail(BMM_ERR_INVALID, "fwht_topk: bad split");
    }
}
} // namespace

int default_persistent_wgs();

// Device-resident engine: d_pat/d_mult hold n_pat entries on the device.
// Returns up to k (weight, mask) entries in ascending order.
std::vector<bmm_topk_entry> fwht_topk_engine(const uint64_t* d_pat, const int32_t* d_mult,
                                             int n_pat, int t, int64_t total, size_t k,
                                             hipStream_t stream) {
    if (t < 1 || t > 48)
        fail(BMM_ERR_INVALID, "fwht_topk: window rows must be in [1, 48]");
    if (k == 0)
        return {};
    if (k > BMM_TOPK_MAX)
        fail(BMM_ERR_INVALID, "fwht_topk: max_candidates exceeds the GPU top-k capacity");
    if (total < 0 || (uint64_t)total >= ((1ull << (64 - t)) - 1))
        fail(BMM_ERR_OVERFLOW, "fwht_topk: total weight too large for packed keys");
    const uint64_t n_masks = 1ull << t;
    const int K = (int)std::min<uint64_t>(k, n_masks - 1);
    const int L = std::min(t, kTopkMaxL);
    const uint64_t n_blocks = n_masks >> L;
    const int grid = (int)std::max<uint64_t>(
        1, std::min<uint64_t>(n_blocks, (uint64_t)default_persistent_wgs() / (K > 512 ? 4 : 1)));
    SearchWorkspace& ws = search_ws();
    unsigned long long* d_counter = ws.counter.get<unsigned long long>(1);
    uint64_t* d_lists = ws.lists.get<uint64_t>((size_t)grid * K);
    uint64_t* d_lists2 = ws.lists2.get<uint64_t>((size_t)((grid + kMergeFan - 1) / kMergeFan) * K);
    uint64_t* d_out;
    if (t >= kPkL && total <= kPkMaxTotal && (n_masks >> kPkL) >= 256) {
        // Packed blocks of 2^14 masks offering their few candidates, exact
        // int32 rescan of the crowded ones.  (Fewer than 256 such blocks
        // keep the int32 blocks of 2^13: twice the workgroups.)
        const uint64_t n14 = n_masks >> kPkL;
        const int grid14 = (int)std::max<uint64_t>(1, std::min<uint64_t>(n14, (uint64_t)grid));
        // Patterns held in registers: 4 or 16 per thread (windows of at
        // most 46 rows), else read from memory per block.
        const int R = t > 46 ? 0 : n_pat <= 256 * 4 ? 4 : n_pat <= 256 * 16 ? 16 : 0;
        // Few blocks per workgroup: a sampled first pass bounds the keys.
        const unsigned long long* d_bound = nullptr;
        unsigned long long* d_queue = d_counter;
        int grid_pk = grid14;
        if (n14 <= (uint64_t)grid14 * 64) {
            TopkBoundState* st = ws.bound.get<TopkBoundState>(1);
            // Only the header needs clearing: every slot read is written.
            fill_u64(reinterpret_cast<uint64_t*>(st), 2, 0, stream);
            const uint64_t sample = std::min<uint64_t>(n14, kBoundSample);
            int shift = 0;
            while ((total >> shift) >= kBoundBins)
                ++shift;
            ScopedKernelTimer timer(stream);
            hipLaunchKernelGGL(topk_bound_pk_kernel, dim3((unsigned)sample), dim3(256), 0, stream,
                               d_pat, d_mult, n_pat, (int32_t)total, n14 / sample, shift, st);
            BMM_HIP(hipGetLastError());
            hipLaunchKernelGGL(topk_bound_finish_kernel, dim3(1), dim3(256), 0, stream,
                               (int)sample, shift, t, K, st);
            BMM_HIP(hipGetLastError());
            d_bound = &st->bound;
            d_queue = &st->counter;
            // Bounded lists need no warm-up: one workgroup per resident slot
            // (fewer lists to merge).
            const int per_cu = K <= 128 && R != 16 ? 4 : K <= 512 ? 3 : 1;
            grid_pk = (int)std::min<uint64_t>(
                (uint64_t)grid14, (uint64_t)(default_persistent_wgs() / 8 * per_cu));
        }
        {
            ScopedKernelTimer timer(stream);
            // Lists of up to 128 keys use a 256-key workspace (at least
            // 128 staging slots; larger offers take several rounds): four
            // workgroups per CU instead of three (not with 16 registers of
            // patterns, which would spill at that occupancy).
            auto kern = K <= 128 && R != 16 ? (R == 4 ? fwht_topk_pk_kernel<256, 4>
                                                      : fwht_topk_pk_kernel<256, 0>)
                        : K <= 512 ? (R == 4    ? fwht_topk_pk_kernel<2048, 4>
                                      : R == 16 ? fwht_topk_pk_kernel<2048, 16>
                                                : fwht_topk_pk_kernel<2048, 0>)
                                   : (R == 4    ? fwht_topk_pk_kernel<8192, 4>
                                      : R == 16 ? fwht_topk_pk_kernel<8192, 16>
                                                : fwht_topk_pk_kernel<8192, 0>);
            // Bounded lists: one contiguous run of blocks per workgroup (one
            // queue atomic each); otherwise up to kPkChunk blocks per atomic.
            const int chunk =
                d_bound ? (int)((n14 + (uint64_t)grid_pk - 1) / (uint64_t)grid_pk)
                        : (int)std::min<uint64_t>(
                              kPkChunk, std::max<uint64_t>(1, n14 / ((uint64_t)grid_pk * 16)));
            if (!d_bound)   // the bounded pass takes the counter cleared with its header
                fill_u64(reinterpret_cast<uint64_t*>(d_counter), 1, 0, stream);
            hipLaunchKernelGGL(kern, dim3(grid_pk), dim3(256), 0, stream, d_pat, d_mult, n_pat, t,
                               (int32_t)total, n14, K, chunk, d_queue, d_bound, d_lists);
            BMM_HIP(hipGetLastError());
        }
        d_out = K <= 512 ? merge_lists<2048>(d_lists, d_lists2, grid_pk, K, stream)
                         : merge_lists<8192>(d_lists, d_lists2, grid_pk, K, stream);
    } else {
        // Few blocks per workgroup (or all of them in the sample): a sampled
        // first pass bounds the keys, as for the packed blocks.
        TopkBoundState* st = nullptr;
        unsigned long long* d_queue = d_counter;
        if (n_blocks <= (uint64_t)grid * 64) {
            st = ws.bound.get<TopkBoundState>(1);
            fill_u64(reinterpret_cast<uint64_t*>(st), 2, 0, stream);
            d_queue = &st->counter;
        } else {
            fill_u64(reinterpret_cast<uint64_t*>(d_counter), 1, 0, stream);
        }
        if (K <= 512) {
            dispatch_fwht_topk<2048>(L, d_pat, d_mult, n_pat, t, (int32_t)total, n_blocks, K,
                                     d_queue, d_lists, grid, st, stream);
            d_out = merge_lists<2048>(d_lists, d_lists2, grid, K, stream);
        } else {
            dispatch_fwht_topk<8192>(L, d_pat, d_mult, n_pat, t, (int32_t)total, n_blocks, K,
                                     d_queue, d_lists, grid, st, stream);
            d_out = merge_lists<8192>(d_lists, d_lists2, grid, K, stream);
        }
    }
    uint64_t* keys = ws.keys.get<uint64_t>((size_t)K);
    BMM_HIP(hipMemcpyAsync(keys, d_out, (size_t)K * 8, hipMemcpyDeviceToHost, stream));
    BMM_HIP(hipStreamSynchronize(stream));
    std::vector<bmm_topk_entry> out;
    const uint64_t mmask = t >= 64 ? ~0ull : ((1ull << t) - 1);
    for (int i = 0; i < K; ++i) {
        const uint64_t key = keys[i];
        if (key == ~0ull)
            break;
        out.push_back(bmm_topk_entry{key & mmask, (uint32_t)(key >> t), 0});
    }
    return out;
}

// FwhtSearch over host rows: patterns of the first `cols` columns (tail masked
// as the reference does, src/search/fwht_search.cpp:64-65).  The rows go up
// in one asynchronous copy from page-locked memory; the active-column count
// (the number of nonzero patterns) is taken on the host from the rows' OR.
std::vector<bmm_topk_entry> fwht_topk_rows(const uint64_t* rows, size_t t, size_t words,
                                           size_t cols, size_t k) {
    if (cols > (size_t)words * 64)
        fail(BMM_ERR_INVALID, "fwht_topk: cols exceeds words_per_row * 64");
    if (cols > (size_t)INT32_MAX)
        fail(BMM_ERR_OVERFLOW, "fwht_topk: total weight exceeds int32 range");
    SearchWorkspace& ws = search_ws();
    uint64_t* d_rows = ws.rows.get<uint64_t>(t * words);
    uint64_t* d_pat = ws.pat.get<uint64_t>(std::max<size_t>(cols, 1));
    int32_t* d_mult = ws.mult.get<int32_t>(std::max<size_t>(cols, 1));
    int64_t total = 0;
    if (cols) {
        uint64_t* h_rows = ws.stage.get<uint64_t>(t * words);
        std::memcpy(h_rows, rows, t * words * 8);
        BMM_HIP(hipMemcpyAsync(d_rows, h_rows, t * words * 8, hipMemcpyHostToDevice, nullptr));
        for (size_t wi = 0; wi * 64 < cols; ++wi) {
            uint64_t any = 0;
            for (size_t i = 0; i < t; ++i)
                any |= rows[i * words + wi];
            if (cols - wi * 64 < 64)
                any &= (1ull << (cols - wi * 64)) - 1;
            total += __builtin_popcountll(any);
        }
        hipLaunchKernelGGL(column_patterns64_kernel, dim3((unsigned)((cols + 255) / 256)),
                           dim3(256), 0, nullptr, d_rows, (int)t, (int)words, (int)cols, d_pat,
                           d_mult);
        BMM_HIP(hipGetLastError());
    }
    return fwht_topk_engine(d_pat, d_mult, (int)cols, (int)t, total, k, nullptr);
}

// The split form (patterns q << low_bits | r with multiplicities): patterns
// and multiplicities are formed on the host and sent in one asynchronous copy.
std::vector<bmm_topk_entry> mitm_topk_split(const uint64_t* q, const uint64_t* r,
                                            const int32_t* mult, size_t unique, size_t low_bits,
                                            size_t high_bits, int32_t total, size_t k) {
    SearchWorkspace& ws = search_ws();
    const size_t n = std::max<size_t>(unique, 1);
    const size_t words = n + (n + 1) / 2;   // patterns, then the int32 multiplicities
    uint64_t* d_in = ws.pat.get<uint64_t>(words);
    if (unique) {
        uint64_t* h_in = ws.stage.get<uint64_t>(words);
        for (size_t i = 0; i < unique; ++i)
            h_in[i] = (low_bits >= 64 ? 0ull : (q[i] << low_bits)) | r[i];
        std::memcpy(h_in + n, mult, unique * 4);
        BMM_HIP(hipMemcpyAsync(d_in, h_in, words * 8, hipMemcpyHostToDevice, nullptr));
    }
    return fwht_topk_engine(d_in, reinterpret_cast<const int32_t*>(d_in + n), (int)unique,
                            (int)(low_bits + high_bits), total, k, nullptr);
}

} // namespace bmmhip
