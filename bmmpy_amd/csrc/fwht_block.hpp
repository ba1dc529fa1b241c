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

__host__ __device__ constexpr int fwht_pad(int x) { return x + (x >> 5); }
__host__ __device__ constexpr int fwht_lds_words(int L) { return (1 << L) + ((1 << L) >> 5) + 1; }

template <int L, int LO, int R> __device__ inline void fwht_round(int32_t* s) {
    constexpr int G = 1 << (L - R);
    constexpr int E = 1 << R;
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        const int low = g & ((1 << LO) - 1);
        const int high = g >> LO;
        const int base = (high << (LO + R)) | low;
        uint32_t v[E];
#pragma unroll
        for (int e = 0; e < E; ++e)
            v[e] = (uint32_t)s[fwht_pad(base + (e << LO))];
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
            s[fwht_pad(base + (e << LO))] = (int32_t)v[e];
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

// Like block_fwht, but the last round hands its results to
// visit(index, value) instead of storing them: thread t sees the entries
// ml = index * 256 + t, index in [0, 2^L / 256).  Needs L >= 11 (the last
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
            const int g = (int)threadIdx.x + 256 * q;
            uint32_t v[E];
#pragma unroll
            for (int e = 0; e < E; ++e)
                v[e] = (uint32_t)s[fwht_pad(g + (e << LO))];
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
                visit(q + PER * e, (int32_t)v[e]);
        }
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
