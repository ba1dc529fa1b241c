// Exact workgroup-level top-K over packed keys (weight << t | mask).
//
// Keys are unique (one per mask), so "the K smallest keys" is the reference
// ordering by (weight, mask) (src/search/fwht_search.cpp:15-22,
// include/bmmpy/search/cuda_launch_topk.cuh:16-19) with no tie ambiguity.
//
// Each workgroup keeps a sorted list of at most K keys in LDS and a threshold
// thr (the K-th key once the list is full).  Threads offer their keys from
// registers; only keys below thr are staged behind the list (LDS atomic slot
// reservation), and a staged batch is merged by one bitonic sort of
// list + batch.  If a batch does not fit, the staged subset still tightens thr
// (the K-th smallest of any superset of the list bounds the true K-th key),
// and the not-yet-staged keys are re-filtered against it, so the loop ends
// with the exact top-K of everything offered.  After warm-up almost no key
// passes the filter and an offer costs one compare per key plus two barriers.
#pragma once

#include "common.hpp"

namespace bmmhip {

template <int CAP> struct TopkShared {
    uint64_t buf[CAP];   // [0, n_list) sorted list, then staged keys
    uint64_t thr;
    uint32_t n_list;
    uint32_t n_req;      // keys requested in the current batch
};

template <int CAP> __device__ inline void topk_init(TopkShared<CAP>& S) {
    if (threadIdx.x == 0) {
        S.thr = ~0ull;
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
        int p2 = 1;
        while (p2 < (int)n)
            p2 <<= 1;
        for (int i = (int)n + threadIdx.x; i < p2; i += blockDim.x)
            S.buf[i] = ~0ull;
        __syncthreads();
        topk_sort(S, p2);
        if (threadIdx.x == 0) {
            const uint32_t nl = min(n, (uint32_t)K);
            S.n_list = nl;
            S.thr = nl == (uint32_t)K ? S.buf[K - 1] : ~0ull;
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
    constexpr int E = 8;
    const int g = blockIdx.x;
    const int first = g * F;
    const size_t n = (size_t)min(F, n_lists - first) * K;
    const uint64_t* keys = lists + (size_t)first * K;
    for (size_t base = 0; base < n; base += (size_t)E * blockDim.x) {
        uint64_t k[E];
        uint32_t valid = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const size_t i = base + (size_t)e * blockDim.x + threadIdx.x;
            k[e] = i < n ? keys[i] : ~0ull;
            if (i < n && k[e] != ~0ull)
                valid |= 1u << e;
        }
        topk_offer<CAP, E>(S, k, valid, K);
    }
    topk_store(S, out + (size_t)g * K, K);
}

} // namespace bmmhip
