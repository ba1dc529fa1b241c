// Exact workgroup-level top-K over packed keys (weight << t | mask).
//
// Keys are unique (one per mask), so "the K smallest keys" is the reference
// ordering by (weight, mask) (src/search/fwht_search.cpp:15-22,
// include/bmmpy/search/cuda_launch_topk.cuh:16-19) with no tie ambiguity.
//
// Each workgroup keeps a sorted list of at most K keys in LDS and a threshold
// thr (the K-th key once the list is full).  Threads offer their keys from
// registers; only keys below thr are staged behind the list (LDS atomic slot
// reservation), and a staged batch is merged by ranks (a few keys) or by one
// bitonic sort of list + batch.  If a batch does not fit, the staged subset
// still tightens thr (the K-th smallest of any superset of the list bounds
// the true K-th key), and the not-yet-staged keys are re-filtered against
// it, so the loop ends with the exact top-K of everything offered.  After
// warm-up almost no key passes the filter and an offer costs one compare per
// key plus two barriers.
//
// An external bound `ext` (topk_init) caps the threshold while the list is
// not full: only keys below it are ever taken.  It must exceed the K-th
// smallest key of the whole search (for instance: K masks of the search are
// known to have keys below it); then every key of the exact top-K passes it
// and the result is unchanged.
#pragma once

#include "common.hpp"

namespace bmmhip {

// Staged batches of at most this many keys are merged into the list by
// ranks (two barriers) instead of a bitonic sort of list + batch.
constexpr uint32_t kTopkRankBatch = 32;

template <int CAP> struct TopkShared {
    uint64_t buf[CAP];   // [0, n_list) sorted list, then staged keys
    uint64_t thr;
    uint64_t ext;        // external bound (~0: none)
    uint32_t n_list;
    uint32_t n_req;      // keys requested in the current batch
};

template <int CAP> __device__ inline void topk_init(TopkShared<CAP>& S, uint64_t ext = ~0ull) {
    if (threadIdx.x == 0) {
        S.thr = ext;
        S.ext = ext;
        S.n_list = 0;
        S.n_req = 0;
    }
    __syncthreads();
}

// Bitonic sort of S.buf[0, n), n a power of two (<= CAP).
template <int CAP> __device__ inline void topk_sort(TopkShared<CAP>& S, int n) {
    for (int k = 2; k <= n; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = S.buf[i], b = S.buf[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        S.buf[i] = b;
                        S.buf[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
}

// Offers up to E keys per thread: bit e of `valid` marks key_of(e) as
// present (key_of computes the packed key, so callers can keep narrower
// values in registers).  Must be called by every thread of the workgroup.
template <int CAP, int E, typename KeyOf>
__device__ inline void topk_offer_fn(TopkShared<CAP>& S, uint32_t valid, int K, KeyOf&& key_of) {
    uint32_t inserted = 0;
    for (;;) {
        __syncthreads();
        const uint64_t thr = S.thr;
        const uint32_t base = S.n_list;
        uint32_t pend = 0;
#pragma unroll
        for (int e = 0; e < E; ++e)
            if (((valid & ~inserted) >> e) & 1u)
                if (key_of(e) < thr)
                    pend |= 1u << e;
        __syncthreads();
        if (threadIdx.x == 0)
            S.n_req = 0;
        __syncthreads();
        const uint32_t cnt = __builtin_popcount(pend);
        if (cnt) {
            const uint32_t cap = CAP - base;
            uint32_t pos = atomicAdd(&S.n_req, cnt);
#pragma unroll
            for (int e = 0; e < E; ++e)
                if ((pend >> e) & 1u) {
                    if (pos < cap) {
                        S.buf[base + pos] = key_of(e);
                        inserted |= 1u << e;
                    }
                    ++pos;
                }
        }
        __syncthreads();
        const uint32_t req = S.n_req;
        if (req == 0)
            return;
        const uint32_t took = min(req, (uint32_t)CAP - base);
        const uint32_t n = base + took;
        if (took <= kTopkRankBatch && n <= 2u * blockDim.x) {
            // A few staged keys: each key's place in the merged order is its
            // rank (keys are distinct), i.e. the keys below it in the list
            // (its own index, or a binary search) plus those staged below it.
            uint64_t x[2];
            uint32_t pos[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const uint32_t i = threadIdx.x + r * blockDim.x;
                pos[r] = ~0u;
                if (i < n) {
                    x[r] = S.buf[i];
                    uint32_t p = 0;
                    for (uint32_t j = 0; j < took; ++j)
                        p += S.buf[base + j] < x[r] ? 1u : 0u;
                    if (i < base) {
                        p += i;
                    } else {
                        uint32_t lo = 0, hi = base;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (S.buf[mid] < x[r])
                                lo = mid + 1;
                            else
                                hi = mid;
                        }
                        p += lo;
                    }
                    pos[r] = p;
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 2; ++r)
                if (pos[r] < (uint32_t)K)
                    S.buf[pos[r]] = x[r];
            __syncthreads();
        } else {
            int p2 = 1;
            while (p2 < (int)n)
                p2 <<= 1;
            for (int i = (int)n + threadIdx.x; i < p2; i += blockDim.x)
                S.buf[i] = ~0ull;
            __syncthreads();
            topk_sort(S, p2);
        }
        if (threadIdx.x == 0) {
            const uint32_t nl = min(n, (uint32_t)K);
            S.n_list = nl;
            S.thr = nl == (uint32_t)K ? S.buf[K - 1] : S.ext;
        }
        if (took == req)
            return;
    }
}

template <int CAP, int E>
__device__ inline void topk_offer(TopkShared<CAP>& S, const uint64_t (&keys)[E], uint32_t valid,
                                  int K) {
    topk_offer_fn<CAP, E>(S, valid, K, [&](int e) { return keys[e]; });
}

// Writes the list (padded with ~0) to out[0, K).
template <int CAP>
__device__ inline void topk_store(TopkShared<CAP>& S, uint64_t* out, int K) {
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += blockDim.x)
        out[i] = i < (int)S.n_list ? S.buf[i] : ~0ull;
}

// Tree level of the merge: workgroup g folds lists [g F, min((g + 1) F,
// n_lists)) of K keys each into the exact K smallest, out[g K, (g + 1) K).
template <int CAP>
__global__ __launch_bounds__(256) void topk_merge_group_kernel(const uint64_t* __restrict__ lists,
                                                               int n_lists, int F, int K,
                                                               uint64_t* __restrict__ out) {
    __shared__ TopkShared<CAP> S;
    topk_init(S);
    // Keys are loaded 2048 at a time (eight loads in flight per thread) and
    // offered 512 at a time: the list fills from the first offers and its
    // threshold then filters the rest, instead of sorting 2048 keys at once.
    constexpr int E = 8, EO = 2;
    const int g = blockIdx.x;
    const int first = g * F;
    const size_t n = (size_t)min(F, n_lists - first) * K;
    const uint64_t* keys = lists + (size_t)first * K;
    for (size_t base = 0; base < n; base += (size_t)E * blockDim.x) {
        uint64_t k[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const size_t i = base + (size_t)e * blockDim.x + threadIdx.x;
            k[e] = i < n ? keys[i] : ~0ull;
        }
#pragma unroll
        for (int o = 0; o < E; o += EO) {
            uint64_t ko[EO];
            uint32_t valid = 0;
#pragma unroll
            for (int e = 0; e < EO; ++e) {
                ko[e] = k[o + e];
                if (ko[e] != ~0ull)
                    valid |= 1u << e;
            }
            if (__syncthreads_or(valid != 0u))   // skip offers with nothing to take
                topk_offer<CAP, EO>(S, ko, valid, K);
        }
    }
    topk_store(S, out + (size_t)g * K, K);
}

// The same fold for groups of F lists of K keys (each sorted ascending with
// ~0 after its valid prefix) that fit kMergeSortedKeys keys at a power-of-two
// list stride KP >= K in LDS, by pairwise merge rounds: in a round, lists 2p
// and 2p + 1 become list p, each valid key x at index j of its list going to
// index j + (keys of the partner list below x) (a binary search by
// power-of-two steps; keys are distinct), kept when below K.  log2(F) rounds
// of two barriers each instead of sorting; 1024 threads, four keys each.
constexpr int kMergeSortedKeys = 4096;
constexpr int kMergeSortedThreads = 1024;

__global__ __launch_bounds__(kMergeSortedThreads) void topk_merge_sorted_kernel(
    const uint64_t* __restrict__ lists, int n_lists, int F, int K, int lk,
    uint64_t* __restrict__ out) {
    __shared__ uint64_t a[kMergeSortedKeys];
    __shared__ uint64_t b[kMergeSortedKeys];
    constexpr int PER = kMergeSortedKeys / kMergeSortedThreads;   // i = tid + 1024 e
    const int KP = 1 << lk;   // list stride in LDS
    const int tid = (int)threadIdx.x;
    const int g = blockIdx.x;
    const int first = g * F;
    int m = min(F, n_lists - first);
    const uint64_t* keys = lists + (size_t)first * K;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int i = tid + kMergeSortedThreads * e;
        const int l = i >> lk, j = i & (KP - 1);
        if (i < m * KP)
            a[i] = j < K ? keys[(size_t)l * K + j] : ~0ull;
    }
    uint64_t* src = a;
    uint64_t* dst = b;
    while (m > 1) {
        const int mo = (m + 1) / 2;
        const int n = m * KP;
        __syncthreads();   // src written
        // Every read is unconditional at an in-range index (i clamped), and
        // its result is applied under the condition afterwards.
        uint64_t x[PER];
        int cnt[PER], qoff[PER];
        bool live[PER], paired[PER];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int i = min(tid + kMergeSortedThreads * e, n - 1);
            x[e] = src[i];
            live[e] = tid + kMergeSortedThreads * e < n && x[e] != ~0ull;
            const int q = (i >> lk) ^ 1;   // partner list
            paired[e] = q < m;
            qoff[e] = paired[e] ? q * KP : 0;
            cnt[e] = 0;
        }
        for (int step = KP; step > 0; step >>= 1) {
            uint64_t v[PER];
#pragma unroll
            for (int e = 0; e < PER; ++e)
                v[e] = src[qoff[e] + min(cnt[e] + step, KP) - 1];
#pragma unroll
            for (int e = 0; e < PER; ++e)
                if (paired[e] && cnt[e] + step <= KP && v[e] < x[e])
                    cnt[e] += step;
        }
        __syncthreads();   // every read of src is done (dst was the previous src)
        for (int i = tid; i < mo * KP; i += kMergeSortedThreads)
            dst[i] = ~0ull;
        __syncthreads();
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int i = tid + kMergeSortedThreads * e;
            const int r = (i & (KP - 1)) + cnt[e];
            if (live[e] && r < K)
                dst[((i >> lk) >> 1) * KP + r] = x[e];
        }
        uint64_t* tmp = src;
        src = dst;
        dst = tmp;
        m = mo;
    }
    __syncthreads();
    for (int i = tid; i < K; i += kMergeSortedThreads)
        out[(size_t)g * K + i] = src[i];
}

} // namespace bmmhip
