// Packed-int16 Walsh-Hadamard blocks for one 256-thread workgroup.
//
// A block of 2^14 entries lives in LDS as 2^13 words: word w holds entry w in
// its low half and entry w + 2^13 in its high half (padded: 4 words after
// every 32, see pk_pad).  Every radix-2 stage on entry bits 0..12
// applies the same butterfly to both halves, so one v_pk_add_u16 /
// v_pk_sub_i16 pair does two butterflies; the stage on entry bit 13 combines
// the two halves of a word, (lo, hi) -> (lo + hi, lo - hi), in one
// v_pk_mad_u16.  Arithmetic wraps modulo 2^16, which is exact whenever every
// partial sum fits in int16: a partial transform of a column-pattern
// histogram is a signed sum of counts, bounded by the number of active
// columns, so the block is exact for at most 32767 active columns
// (the int16 limit of src/search/cuda_mitm_fwht_search.cpp:118-162).
//
// Build: the histogram is accumulated into biased halves (every half starts at
// 0x8000), so a +-1 update of one half never borrows from or carries into the
// other.  The bias needs no removal: the first stage adds or subtracts two
// biased halves, and 0x8000 + 0x8000 = 0 and 0x8000 - 0x8000 = 0 modulo 2^16.
#pragma once

#include "fwht_block.hpp"

namespace bmmhip {

typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ inline s16x2 pk_as(uint32_t u) { return __builtin_bit_cast(s16x2, u); }
__device__ inline uint32_t pk_bits(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ inline int32_t pk_lo(s16x2 v) { return (int32_t)v.x; }
__device__ inline int32_t pk_hi(s16x2 v) { return (int32_t)v.y; }
__device__ inline s16x2 pk_max(s16x2 a, s16x2 b) { return __builtin_elementwise_max(a, b); }

constexpr int kPkL = 14;                 // entries per block: 2^14
constexpr int kPkW = kPkL - 1;           // word index bits
constexpr uint32_t kPkBias = 0x80008000u;
constexpr int kPkMaxTotal = 32767;

// LDS layout of a packed block: 4 pad words after every 32 (rows of 36
// words, 16-byte aligned).  The first round's 32 consecutive words per
// thread are then eight 16-byte accesses; lanes 36 words apart start on
// distinct bank groups, and 8-byte accesses of lanes 2 words apart within a
// row cover distinct bank pairs.
__host__ __device__ constexpr int pk_pad(int x) { return x + ((x >> 5) << 2); }
template <int LO> __host__ __device__ constexpr int pk_pad_stride() {
    static_assert(LO == 0 || LO >= 5, "strides inside a 32-word row are unpadded");
    return pk_pad(1 << LO);
}
__host__ __device__ constexpr int pk_lds_words() { return pk_pad(1 << kPkW); }

template <int E> __device__ inline void pk_butterflies(uint32_t (&v)[E]) {
#pragma unroll
    for (int len = 1; len < E; len <<= 1)
#pragma unroll
        for (int i = 0; i < E; ++i)
            if ((i & len) == 0) {
                const s16x2 a = pk_as(v[i]), b = pk_as(v[i + len]);
                v[i] = pk_bits(a + b);
                v[i + len] = pk_bits(a - b);
            }
}

// Stage on entry bit 13: (lo, hi) -> (lo + hi, lo - hi).
__device__ inline uint32_t pk_half_stage(uint32_t x) {
    const s16x2 a = pk_as(x);
    const s16x2 hh = __builtin_shufflevector(a, a, 1, 1);
    const s16x2 ll = __builtin_shufflevector(a, a, 0, 0);
    const s16x2 sg = {1, -1};
    return pk_bits(hh * sg + ll);
}

// Bias-filled block, then g[pl] += sign per cached column pattern (pl = the
// low 14 bits, sign = parity of (pattern >> 14) & mh).  Patterns 0 are
// inactive columns.
template <int NREG>
__device__ inline void pk_build(uint32_t* s, const uint32_t (&preg)[NREG], uint32_t mh) {
    int4* s4 = reinterpret_cast<int4*>(s);
    const int b = (int)kPkBias;
    for (int i = threadIdx.x; i < (pk_lds_words() + 3) / 4; i += blockDim.x)
        s4[i] = make_int4(b, b, b, b);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NREG; ++k) {
        const uint32_t p = preg[k];
        if (p != 0u) {
            const uint32_t pl = p & ((1u << kPkL) - 1u);
            const uint32_t one = (pl >> kPkW) ? 0x10000u : 1u;
            const int neg = __builtin_popcount((p >> kPkL) & mh) & 1;
            atomicAdd(&s[pk_pad((int)(pl & ((1u << kPkW) - 1u)))], neg ? 0u - one : one);
        }
    }
}

// Same from patterns in memory (wider matrices).
__device__ inline void pk_build_mem(uint32_t* s, const uint32_t* __restrict__ pat, int cols,
                                    uint32_t mh) {
    int4* s4 = reinterpret_cast<int4*>(s);
    const int b = (int)kPkBias;
    for (int i = threadIdx.x; i < (pk_lds_words() + 3) / 4; i += blockDim.x)
        s4[i] = make_int4(b, b, b, b);
    __syncthreads();
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
        const uint32_t p = pat[c];
        if (p == 0u)
            continue;
        const uint32_t pl = p & ((1u << kPkL) - 1u);
        const uint32_t one = (pl >> kPkW) ? 0x10000u : 1u;
        const int neg = __builtin_popcount((p >> kPkL) & mh) & 1;
        atomicAdd(&s[pk_pad((int)(pl & ((1u << kPkW) - 1u)))], neg ? 0u - one : one);
    }
}

// One radix-2^R round over word bits [LO, LO + R): one group per thread
// (2^(13 - R) = 256 groups for R = 5).  The first round also applies the
// bit-13 stage (which cancels the build bias).
template <int LO, int R, bool FIRST> __device__ inline void pk_round(uint32_t* s) {
    static_assert((1 << (kPkW - R)) == 256, "one group per thread");
    constexpr int E = 1 << R;
    constexpr int ST = pk_pad_stride<LO>();
    const int g = (int)threadIdx.x;
    const int low = g & ((1 << LO) - 1);
    const int high = g >> LO;
    uint32_t* p = s + pk_pad((high << (LO + R)) | low);
    uint32_t v[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
        v[e] = p[e * ST];
    if constexpr (FIRST) {
#pragma unroll
        for (int e = 0; e < E; ++e)
            v[e] = pk_half_stage(v[e]);
    }
    pk_butterflies(v);
#pragma unroll
    for (int e = 0; e < E; ++e)
        p[e * ST] = v[e];
}

// The first two rounds (word bits 0-9, with the bit-13 stage).  Starts and
// ends with a barrier; LDS then holds the pre-last-round words.
__device__ inline void pk_block_fwht_rounds(uint32_t* s) {
    __syncthreads();
    pk_round<0, 5, true>(s);
    __syncthreads();
    pk_round<5, 5, false>(s);
    __syncthreads();
}

// The last round's words of thread t, word w = (q + 4 e) * 256 + t for q < 4
// (entries w in the low half and w + 2^13 in the high half), handed to
// visit(q, v) after the last three stages (BUTTERFLY) or as loaded (the
// pre-last-round words, whose group of 8 combines into v's 8 outputs).
template <bool BUTTERFLY, typename Visit>
__device__ inline void pk_last_round_visit(const uint32_t* s, Visit&& visit) {
    constexpr int ST = pk_pad_stride<10>();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t* p = s + pk_pad((int)threadIdx.x + 256 * q);
        uint32_t v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e)
            v[e] = p[e * ST];
        if constexpr (BUTTERFLY)
            pk_butterflies(v);
        visit(q, v);
    }
}

// The stages on word bits 10-12 in place: thread t takes the 4 groups of 8
// words c + 1024 e (e < 8) for c = 2 t + j + 512 h (j, h < 2).  The words
// c, c + 1 of one e are adjacent and 8-byte aligned, so they go as one 8-byte
// access (the stride between e is too wide for paired offsets), and 16 lanes
// of one access cover distinct bank pairs.
__device__ inline void pk_round_top(uint32_t* s) {
    constexpr int ST = pk_pad_stride<10>();
#pragma unroll
    for (int h = 0; h < 2; ++h) {   // two groups at a time (register budget)
        uint32_t* p = s + pk_pad(2 * (int)threadIdx.x + 512 * h);
        uint32_t v[2][8];
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                v[j][e] = p[e * ST + j];
#pragma unroll
        for (int j = 0; j < 2; ++j)
            pk_butterflies(v[j]);
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                p[e * ST + j] = v[j][e];
    }
}

// Thread t's group over word bits 5-9 (words (t >> 5) << 10 | e << 5 | (t & 31),
// e < 32), as loaded.
template <typename Visit> __device__ inline void pk_mid_group_visit(const uint32_t* s, Visit&& visit) {
    constexpr int ST = pk_pad_stride<5>();
    const int g = (int)threadIdx.x;
    const uint32_t* p = s + pk_pad(((g >> 5) << 10) | (g & 31));
    uint32_t v[32];
#pragma unroll
    for (int e = 0; e < 32; ++e)
        v[e] = p[e * ST];
    visit(v);
}

// Transforms the block; the last round goes by thread t's words c + 1024 e
// for c = 2 t + j + 512 h (entry bit 0 = j, bits 1-8 = t, bit 9 = h, bits
// 10-12 = e), handed to visit(j | h << 1, v) with v[e] after the last three
// stages.  The words c, c + 1 of one e go as one 8-byte access.  Starts with
// a barrier; leaves LDS holding the pre-last-round words.
template <typename Visit> __device__ inline void pk_block_fwht_visit_paired(uint32_t* s,
                                                                           Visit&& visit) {
    pk_block_fwht_rounds(s);
    constexpr int ST = pk_pad_stride<10>();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t* p = s + pk_pad(2 * (int)threadIdx.x + 512 * h);
        uint32_t v0[8], v1[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            v0[e] = p[e * ST];
            v1[e] = p[e * ST + 1];
        }
        pk_butterflies(v0);
        visit(h << 1, v0);
        pk_butterflies(v1);
        visit(1 | h << 1, v1);
    }
}

// Transforms the block and hands the last round's words to visit(q, v)
// instead of storing them: v[e] (e < 8) is word w = (q + 4 e) * 256 + t for
// thread t, q < 4, i.e. entries w (low half) and w + 2^13 (high half).
// Starts with a barrier; leaves LDS holding the pre-last-round words.
template <typename Visit> __device__ inline void pk_block_fwht_visit(uint32_t* s, Visit&& visit) {
    pk_block_fwht_rounds(s);
    pk_last_round_visit<true>(s, visit);
}

// Packed subset maxima (subset_max in fwht_block.hpp, both halves at once).
template <int N> __device__ inline s16x2 pk_subset_max(s16x2* x, s16x2* bm) {
    if constexpr (N == 1) {
        return x[0];
    } else {
        constexpr int H = N / 2;
        s16x2 top = x[H];
#pragma unroll
        for (int i = H + 1; i < N; ++i)
            top = pk_max(top, x[i]);
        bm[ilog2c(H)] = top;
#pragma unroll
        for (int i = 0; i < H; ++i)
            x[i] = pk_max(x[i], x[i + H]);
        return pk_subset_max<H>(x, bm);
    }
}

} // namespace bmmhip
