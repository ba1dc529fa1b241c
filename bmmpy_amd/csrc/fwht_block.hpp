// On-chip Walsh-Hadamard building blocks for one 256-thread workgroup.
//
// A block of 2^L int32 entries lives in LDS with one padding word per 32
// entries (entry x at x + x/32), so both the radix-32 gathers (stride 1 per
// thread, 32 entries apart per lane) and the strided gathers hit distinct
// banks.  block_fwht<L> applies the L radix-2 stages as rounds of up to five
// stages held in registers; since the per-bit butterflies commute, any round
// order gives the natural-order transform (src/math/fwht.cpp:28-57) with the
// same wrapping arithmetic (src/math/fwht_scalar.cpp:19-29).
#pragma once

#include "common.hpp"

namespace bmmhip {

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
__host__ __device__ constexpr int fwht_pad(int x) { return x + (x >> 5); }
__host__ __device__ constexpr int fwht_lds_words(int L) { return (1 << L) + ((1 << L) >> 5) + 1; }

// Padded distance between the entries one round combines.  Rounds start at
// LO in {0, 5, 10, 15}: for LO >= 5, adding e << LO moves the padded index by
// e * (2^LO + 2^(LO-5)); for LO = 0 the group base is a multiple of 32 (or 0)
// and e < 32, so the padded index moves by e.  One base address per group then
// serves all 2^R entries through immediate offsets.
template <int LO> __host__ __device__ constexpr int fwht_pad_stride() {
    static_assert(LO % 5 == 0, "rounds start at multiples of 5 bits");
    return (1 << LO) + ((1 << LO) >> 5);
}

template <int L, int LO, int R> __device__ inline void fwht_round(int32_t* s) {
    constexpr int G = 1 << (L - R);
    constexpr int E = 1 << R;
    constexpr int ST = fwht_pad_stride<LO>();
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        const int low = g & ((1 << LO) - 1);
        const int high = g >> LO;
        int32_t* p = s + fwht_pad((high << (LO + R)) | low);
        uint32_t v[E];
#pragma unroll
        for (int e = 0; e < E; ++e)
            v[e] = (uint32_t)p[e * ST];
#pragma unroll
        for (int len = 1; len < E; len <<= 1)
#pragma unroll
            for (int i = 0; i < E; ++i)
                if ((i & len) == 0) {
                    const uint32_t a = v[i], b = v[i + len];
                    v[i] = a + b;
                    v[i + len] = a - b;
                }
#pragma unroll
        for (int e = 0; e < E; ++e)
            p[e * ST] = (int32_t)v[e];
    }
}

// Transforms s[0 .. 2^L) in place; starts and ends with a workgroup barrier.
template <int L, int LO = 0> __device__ inline void block_fwht(int32_t* s) {
    if constexpr (LO < L) {
        constexpr int R = (L - LO) < 5 ? (L - LO) : 5;
        __syncthreads();
        fwht_round<L, LO, R>(s);
        block_fwht<L, LO + R>(s);
    } else {
        __syncthreads();
    }
}

// Like block_fwht, but the last round hands its results to visit(q, v)
// instead of storing them: v[e] (e < 2^R) is entry ml = (q + P * e) * 256 + t
// for thread t, with P = 2^(L-R-8) and q in [0, P).  Needs L >= 11 (the last
// round then spans >= 256 groups, one or more per thread, with stride >= 256);
// starts with a barrier and leaves LDS holding the pre-last-round values.
template <int L, int LO, typename Visit> __device__ inline void block_fwht_visit(int32_t* s,
                                                                                Visit&& visit) {
    constexpr int R = (L - LO) < 5 ? (L - LO) : 5;
    if constexpr (LO + R < L) {
        __syncthreads();
        fwht_round<L, LO, R>(s);
        block_fwht_visit<L, LO + R>(s, visit);
    } else {
        static_assert(LO >= 8, "fused last round needs a stride of at least 256");
        __syncthreads();
        constexpr int G = 1 << LO;          // groups in the last round
        constexpr int E = 1 << R;
        constexpr int PER = G / 256;        // groups per thread
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int32_t* p = s + fwht_pad((int)threadIdx.x + 256 * q);
            uint32_t v[E];
#pragma unroll
            for (int e = 0; e < E; ++e)
                v[e] = (uint32_t)p[e * fwht_pad_stride<LO>()];
#pragma unroll
            for (int len = 1; len < E; len <<= 1)
#pragma unroll
                for (int i = 0; i < E; ++i)
                    if ((i & len) == 0) {
                        const uint32_t a = v[i], b = v[i + len];
                        v[i] = a + b;
                        v[i + len] = a - b;
                    }
            visit(q, v);
        }
    }
}

// Subset maxima of x[0..N): returns the maximum of all entries and sets
// bm[j] = max of x[i] over the i with bit j set (j < log2 N), in
// 2N - log2(N) - 2 max operations; x is overwritten.
template <int N> __device__ inline int32_t subset_max(int32_t* x, int32_t* bm) {
    if constexpr (N == 1) {
        return x[0];
    } else {
        constexpr int H = N / 2;
        int32_t top = x[H];
#pragma unroll
        for (int i = H + 1; i < N; ++i)
            top = max(top, x[i]);
        bm[ilog2c(H)] = top;
#pragma unroll
        for (int i = 0; i < H; ++i)
            x[i] = max(x[i], x[i + H]);
        return subset_max<H>(x, bm);
    }
}

// g[pl] = sum over active columns c of (-1)^popcount((pat[c] >> L) & mh) at
// pl = pat[c] & (2^L - 1).  Column patterns carry row r at bit r.
template <int L>
__device__ inline void block_build_g(int32_t* s, const uint32_t* __restrict__ pat, int cols,
                                     uint32_t mh) {
    constexpr int N = 1 << L;
    for (int i = threadIdx.x; i < fwht_lds_words(L); i += blockDim.x)
        s[i] = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
        const uint32_t p = pat[c];
        if (p == 0u)
            continue;
        const int neg = __builtin_popcount((p >> L) & mh) & 1;
        atomicAdd(&s[fwht_pad((int)(p & (N - 1)))], neg ? -1 : 1);
    }
}

} // namespace bmmhip
