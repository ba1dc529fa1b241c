// Single-target coset minimum over wide windows (the planted meet-in-the-middle
// workload): for target row t and window rows w_0..w_{n-1} (n <= 47), the
// minimum over subsets S of the window of wt(row_t ^ XOR_S) and the lowest
// such S (bit j = w_j).
//
// Signed histogram: every active column c (any of target/window set) has a
// window pattern p_c and a sign s_c = -1 if the target bit is set, else +1;
// b[p] = sum of s_c, total = active columns; then weight(m) = (total - H[m]) / 2
// with H = FWHT(b) (the FwhtSearch identity, src/search/fwht_search.cpp:105-112,
// applied to the window + target with the target bit forced).  The mask space
// is split m = mh << L | ml exactly as in the MITM search
// (src/search/mitm_fwht_search.cpp:164-265, which walks the low half and
// transforms the high half): one workgroup block enumerates one mh, builds
// g_mh[pl] = sum_i mult_i (-1)^popcount(ph_i & mh) in LDS from the unique
// (pattern, signed multiplicity) list and transforms the 2^L low half on chip.
// Each block folds (weight << n | mask) into the target's key with one
// atomicMin, so the final key carries the minimum weight and its lowest mask.
#include "common.hpp"
#include "fwht_block.hpp"
#include "fwht_pk16.hpp"

#include <algorithm>
#include <vector>

namespace bmmhip {

struct WinTarget {
    int64_t pat_off;
    int32_t n_pat;
    int32_t bits;     // window size n
    int32_t total;
    int32_t L;        // (total + 1) << L < 2^31 (packed block keys below)
    int64_t first_block;  // number of this target's block 0 in its kernel's queue
    int32_t slot;     // output index
    int32_t top;      // packed targets: 1 when pattern >> 14 has 33 bits (47 rows)
    int64_t lo_off;   // packed targets: WinPat8 list of the low-half entries
    int64_t hi_off;   //   and of the high-half entries,
    int32_t n_lo;     //   each padded to a multiple of 256 entries
    int32_t n_hi;
};

constexpr int kWinMaxL = 13;

// A window entry prepared for packed-int16 blocks.  The sign for block mh is
// the parity of (pattern >> 14) & mh: `hi` keeps the low 32 bits of
// pattern >> 14, and av bit 0 its bit 32 (47-row windows only; the LDS byte
// address is a multiple of 4).  av = mult << 16 | LDS byte address | that bit:
// a low-half entry adds mult (sign-extended) to the word, a high-half entry
// mult << 16.  Padding entries add 0 at distinct addresses.
struct WinPat8 {
    uint32_t hi;
    uint32_t av;
};

// Entries per thread and half held in registers across a target's blocks
// (lists of at most 256 * 10 entries per half); longer lists are read from
// memory for every block.
constexpr int kWinRegSlots = 10;

template <int L>
__device__ inline void block_build_g_signed(int32_t* s, const uint64_t* __restrict__ pat,
                                            const int32_t* __restrict__ mult, int n,
                                            uint64_t mh) {
    for (int i = threadIdx.x; i < fwht_lds_words(L); i += blockDim.x)
        s[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const uint64_t p = pat[i];
        const int32_t m = mult[i];
        const int neg = __builtin_popcountll((p >> L) & mh) & 1;
        atomicAdd(&s[fwht_pad((int)(p & ((1u << L) - 1u)))], neg ? -m : m);
    }
}

// One block mh: transform, then the block's best element as the packed key
// K = H * 2^L - ml (maximum H first, then lowest ml; |H| <= total keeps it in
// int32), and one atomicMin of (weight << n | mh << L | ml) into the key.
template <int L>
__device__ void win_block(int32_t* s, const WinTarget& T, const uint64_t* pats,
                          const int32_t* mults, uint64_t mh, unsigned long long* key) {
    block_build_g_signed<L>(s, pats + T.pat_off, mults + T.pat_off, T.n_pat, mh);
    constexpr int N = 1 << L;
    constexpr int NE = N >= 256 ? N / 256 : 1;   // register values per thread
    int32_t best = INT32_MIN;
    // v[e] is register index q + PER * e, i.e. ml = (q + PER * e) * 256 + tid.
    auto visit = [&](int q, const auto& v) {
        constexpr int E = sizeof(v) / sizeof(v[0]);
        constexpr int PER = NE / E;
#pragma unroll
        for (int e = 0; e < E; ++e)
            best = max(best, (int32_t)(v[e] << L) - 256 * (q + PER * e));
    };
    if constexpr (L >= 11) {
        block_fwht_visit<L, 0>(s, visit);
    } else {
        block_fwht<L>(s);
        if ((int)threadIdx.x < N) {
            uint32_t v[NE];
#pragma unroll
            for (int e = 0; e < NE; ++e)
                v[e] = (uint32_t)s[fwht_pad(e * 256 + (int)threadIdx.x)];
            visit(0, v);
        }
    }
    if (best != INT32_MIN)
        best -= (int)threadIdx.x;
    best = wave_max_i32(best);
    __shared__ int32_t wbest[4];
    if (threadIdx.x % kWave == 0)
        wbest[threadIdx.x / kWave] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int32_t k = max(max(wbest[0], wbest[1]), max(wbest[2], wbest[3]));
        // k = H * 2^L - ml with 0 <= ml < 2^L: H = ceil(k / 2^L).
        const int32_t h = -((-k) >> L);
        const uint32_t ml = (uint32_t)((h << L) - k);
        const uint32_t w = (uint32_t)(T.total - h) >> 1;
        const unsigned long long nk = ((unsigned long long)w << T.bits) | (mh << L) | ml;
        if (nk < __atomic_load_n(key, __ATOMIC_RELAXED))  // keys only decrease
            atomicMin(key, nk);
    }
    __syncthreads();
}

// One entry of a packed block build: sign = parity(hi & mh) (+ av bit 0 &
// mh bit 32 when TOP), applied as (v ^ m) - m with m = 0 or -1.
template <bool TOP, bool HIGH>
__device__ inline void win_add(char* sb, uint32_t hi, uint32_t av, uint32_t mh, uint32_t mh_top) {
    const uint32_t par = __builtin_popcount(hi & mh) + (TOP ? (av & mh_top) : 0u);
    const int32_t p = (int32_t)(par & 1u);
    const int32_t v = HIGH ? (int32_t)(av & 0xffff0000u) : ((int32_t)av >> 16);
    atomicAdd(reinterpret_cast<uint32_t*>(sb + (av & 0xfffcu)), (uint32_t)((v ^ -p) + p));
}

// The entries of one padded list from memory, four loads in flight.
template <bool TOP, bool HIGH>
__device__ inline void win_add_list(char* sb, const WinPat8* __restrict__ p, int n, uint32_t mh,
                                    uint32_t mh_top) {
    const WinPat8* q = p + threadIdx.x;
    int i = 0;
    for (; i + 1024 <= n; i += 1024) {
        const WinPat8 P0 = q[i], P1 = q[i + 256], P2 = q[i + 512], P3 = q[i + 768];
        win_add<TOP, HIGH>(sb, P0.hi, P0.av, mh, mh_top);
        win_add<TOP, HIGH>(sb, P1.hi, P1.av, mh, mh_top);
        win_add<TOP, HIGH>(sb, P2.hi, P2.av, mh, mh_top);
        win_add<TOP, HIGH>(sb, P3.hi, P3.av, mh, mh_top);
    }
    for (; i < n; i += 256)
        win_add<TOP, HIGH>(sb, q[i].hi, q[i].av, mh, mh_top);
}

// A target's entries held in registers: slot k < n_lo / 256 of the low half
// is entry lo_off + 256 k + t, likewise the high half at kWinRegSlots + k.
struct WinRegs {
    uint32_t hi[2 * kWinRegSlots];
    uint32_t av[2 * kWinRegSlots];
    bool held;
};

__device__ inline void win_load_regs(WinRegs& R, const WinTarget& T,
                                     const WinPat8* __restrict__ w8) {
    R.held = T.n_lo <= 256 * kWinRegSlots && T.n_hi <= 256 * kWinRegSlots;
    if (!R.held)
        return;
#pragma unroll
    for (int k = 0; k < kWinRegSlots; ++k) {
        if (256 * k < T.n_lo) {
            const WinPat8 P = w8[T.lo_off + 256 * k + threadIdx.x];
            R.hi[k] = P.hi;
            R.av[k] = P.av;
        }
        if (256 * k < T.n_hi) {
            const WinPat8 P = w8[T.hi_off + 256 * k + threadIdx.x];
            R.hi[kWinRegSlots + k] = P.hi;
            R.av[kWinRegSlots + k] = P.av;
        }
    }
}

template <bool TOP>
__device__ inline void win_add_all(char* sb, const WinTarget& T, const WinRegs& R,
                                   const WinPat8* __restrict__ w8, uint32_t mh, uint32_t mh_top) {
    if (R.held) {
#pragma unroll
        for (int k = 0; k < kWinRegSlots; ++k)
            if (256 * k < T.n_lo)
                win_add<TOP, false>(sb, R.hi[k], R.av[k], mh, mh_top);
#pragma unroll
        for (int k = 0; k < kWinRegSlots; ++k)
            if (256 * k < T.n_hi)
                win_add<TOP, true>(sb, R.hi[kWinRegSlots + k], R.av[kWinRegSlots + k], mh, mh_top);
    } else {
        win_add_list<TOP, false>(sb, w8 + T.lo_off, T.n_lo, mh, mh_top);
        win_add_list<TOP, true>(sb, w8 + T.hi_off, T.n_hi, mh, mh_top);
    }
}

// Bias-filled block plus the signed entries.  `after_clear` runs once every
// thread is past the clear barrier (so past all reads of the previous
// block's shared state).
template <typename AfterClear>
__device__ inline void win_build_pk16(uint32_t* s, const WinTarget& T, const WinRegs& R,
                                      const WinPat8* __restrict__ w8, uint64_t mh,
                                      AfterClear&& after_clear) {
    int4* s4 = reinterpret_cast<int4*>(s);
    const int b = (int)kPkBias;
    for (int i = threadIdx.x; i < (pk_lds_words() + 3) / 4; i += blockDim.x)
        s4[i] = make_int4(b, b, b, b);
    __syncthreads();
    after_clear();
    char* sb = reinterpret_cast<char*>(s);
    if (T.top)
        win_add_all<true>(sb, T, R, w8, (uint32_t)mh, (uint32_t)(mh >> 32) & 1u);
    else
        win_add_all<false>(sb, T, R, w8, (uint32_t)mh, 0u);
}

// Packed-int16 block of 2^14 masks (windows of >= 14 rows, at most 32767
// active columns).  Word w = (q + 4 e) * 256 + t holds masks w and
// w + 2^13.
//
// The block's best key is at least (w << n | mh << 14) with
// w = (total - max H) / 2; when that floor is not below the target's current
// key (a relaxed read: keys only decrease, so a stale value only sends more
// blocks on), no mask of the block can win and the block ends.  The test
// runs first on an upper bound of max H: the stages on word bits 0-4 (with
// the bit-13 stage) and 10-12 are done, and the 32 words x_0..x_31 of a
// group over bits 5-9, each a signed sum over its own set of column
// patterns, would combine into outputs x_0 + sum_e +-x_e <=
// x_0 + sum_{e >= 1} |x_e| (at most total, so no int16 overflow).  A block the
// bound cannot settle finishes those stages in registers and is tested on its
// exact max H.  A block that still can win is rebuilt and transformed with
// the exact keyed scan: every entry folds into the key
// H * 2^16 | (65535 - 4 m') with m' = ml - t, so one max keeps the largest H
// and the lowest ml among the thread's entries.
__device__ void win_block_pk16(uint32_t* s, const WinTarget& T, const WinRegs& R,
                               const WinPat8* __restrict__ w8, uint64_t mh,
                               unsigned long long* key) {
    __shared__ int32_t wbest[4], wexact[4];
    __shared__ unsigned long long s_key;
    // The target's key is read (relaxed) right after the clear barrier and
    // only used after the transform, so the load's latency overlaps the build.
    unsigned long long key_seen = 0;
    win_build_pk16(s, T, R, w8, mh, [&] {
        if (threadIdx.x == 0)
            key_seen = __atomic_load_n(key, __ATOMIC_RELAXED);
    });
    __syncthreads();
    pk_round<0, 5, true>(s);
    __syncthreads();
    pk_round_top(s);
    __syncthreads();
    if (threadIdx.x == 0)
        s_key = key_seen;
    // Block maximum of one value per thread; every thread takes the same
    // decision from the same shared values.
    auto cannot_win = [&](int32_t h, int32_t* ws) {
        h = wave_max_i32(h);
        if (threadIdx.x % kWave == 0)
            ws[threadIdx.x / kWave] = h;
        __syncthreads();
        const int32_t hb = max(max(ws[0], ws[1]), max(ws[2], ws[3]));
        const uint32_t w = (uint32_t)(T.total - hb) >> 1;
        const unsigned long long floor_key = ((unsigned long long)w << T.bits) | (mh << kPkL);
        return !(floor_key < s_key);
    };
    int32_t hbound = 0;
    pk_mid_group_visit(s, [&](const uint32_t (&v)[32]) {
        s16x2 b = pk_as(v[0]);
#pragma unroll
        for (int e = 1; e < 32; ++e) {
            const s16x2 x = pk_as(v[e]);
            b = b + pk_max(x, -x);
        }
        hbound = max(pk_lo(b), pk_hi(b));
    });
    if (cannot_win(hbound, wbest))
        return;
    int32_t hexact = 0;
    pk_mid_group_visit(s, [&](const uint32_t (&v)[32]) {
        uint32_t y[32];
#pragma unroll
        for (int e = 0; e < 32; ++e)
            y[e] = v[e];
        pk_butterflies(y);
        s16x2 m = pk_as(y[0]);
#pragma unroll
        for (int e = 1; e < 32; ++e)
            m = pk_max(m, pk_as(y[e]));
        hexact = max(pk_lo(m), pk_hi(m));
    });
    if (cannot_win(hexact, wexact))
        return;
    win_build_pk16(s, T, R, w8, mh, [] {});
    int32_t best = INT32_MIN;
    pk_block_fwht_visit(s, [&](int q, const uint32_t (&v)[8]) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t m_lo = (uint32_t)(256 * (q + 4 * e));
            const uint32_t c_lo = 65535u - 4u * m_lo;
            const uint32_t c_hi = 65535u - 4u * (m_lo + (1u << kPkW));
            best = max(best, (int32_t)((v[e] << 16) | c_lo));
            best = max(best, (int32_t)((v[e] & 0xffff0000u) | c_hi));
        }
    });
    // Thread key -> (H, ml) -> block key H * 2^16 | (65535 - 4 ml).
    {
        const int32_t h = best >> 16;
        const uint32_t ml = ((65535u - ((uint32_t)best & 0xffffu)) >> 2) + threadIdx.x;
        best = (int32_t)(((uint32_t)h << 16) | (65535u - 4u * ml));
    }
    best = wave_max_i32(best);
    if (threadIdx.x % kWave == 0)
        wbest[threadIdx.x / kWave] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int32_t k = max(max(wbest[0], wbest[1]), max(wbest[2], wbest[3]));
        const int32_t h = k >> 16;
        const uint32_t ml = (65535u - ((uint32_t)k & 0xffffu)) >> 2;
        const uint32_t w = (uint32_t)(T.total - h) >> 1;
        const unsigned long long nk = ((unsigned long long)w << T.bits) | (mh << kPkL) | ml;
        if (nk < __atomic_load_n(key, __ATOMIC_RELAXED))  // keys only decrease
            atomicMin(key, nk);
    }
    __syncthreads();
}

// Blocks of the targets whose blocks are 2^L wide (the persistent queue of
// one kernel); targets in queue order, found by scanning forward.
template <int L>
__global__ __launch_bounds__(256, 4) void win_full_kernel(const WinTarget* __restrict__ tgts,
                                                       int n_tgt, int64_t total_blocks,
                                                       const uint64_t* __restrict__ pats,
                                                       const int32_t* __restrict__ mults,
                                                       const WinPat8* __restrict__ w8,
                                                       int chunk,
                                                       unsigned long long* __restrict__ counter,
                                                       unsigned long long* __restrict__ keys) {
    constexpr int kWords = L == kPkL ? pk_lds_words() : fwht_lds_words(L);
    __shared__ __align__(16) int32_t s[kWords + 4];
    __shared__ long long first;
    int ti = 0;
    [[maybe_unused]] WinRegs R;
    [[maybe_unused]] int held_ti = -1;
    for (;;) {
        if (threadIdx.x == 0)
            first = (long long)atomicAdd(counter, (unsigned long long)chunk);
        __syncthreads();
        const int64_t g0 = first;
        __syncthreads();
        if (g0 >= total_blocks)
            break;
        const int64_t g1 = min(g0 + (int64_t)chunk, total_blocks);
        for (int64_t g = g0; g < g1; ++g) {
            while (ti + 1 < n_tgt && tgts[ti + 1].first_block <= g)
                ++ti;
            const WinTarget T = tgts[ti];
            const uint64_t mh = (uint64_t)(g - T.first_block);
            if constexpr (L == kPkL) {
                if (held_ti != ti) {   // a new target: reload its register-held entries
                    win_load_regs(R, T, w8);
                    held_ti = ti;
                }
                win_block_pk16(reinterpret_cast<uint32_t*>(s), T, R, w8, mh, &keys[T.slot]);
            } else
                win_block<L>(s, T, pats, mults, mh, &keys[T.slot]);
        }
    }
}

// Targets with narrower blocks (small windows or very many columns).
__global__ __launch_bounds__(256) void win_min_kernel(const WinTarget* __restrict__ tgts,
                                                      int n_tgt, int64_t total_blocks,
                                                      const uint64_t* __restrict__ pats,
                                                      const int32_t* __restrict__ mults,
                                                      const WinPat8* __restrict__ /*w8*/,
                                                      int chunk,
                                                      unsigned long long* __restrict__ counter,
                                                      unsigned long long* __restrict__ keys) {
    __shared__ __align__(16) int32_t s[fwht_lds_words(kWinMaxL) + 4];
    __shared__ long long first;
    int ti = 0;
    for (;;) {
        if (threadIdx.x == 0)
            first = (long long)atomicAdd(counter, (unsigned long long)chunk);
        __syncthreads();
        const int64_t g0 = first;
        __syncthreads();
        if (g0 >= total_blocks)
            break;
        const int64_t g1 = min(g0 + (int64_t)chunk, total_blocks);
        for (int64_t g = g0; g < g1; ++g) {
            while (ti + 1 < n_tgt && tgts[ti + 1].first_block <= g)
                ++ti;
            const WinTarget T = tgts[ti];
            const uint64_t mh = (uint64_t)(g - T.first_block);
            switch (T.L) {
#define BMM_CASE(LL)                                                                           \
    case LL: win_block<LL>(s, T, pats, mults, mh, &keys[T.slot]); break;
                BMM_CASE(1) BMM_CASE(2) BMM_CASE(3) BMM_CASE(4) BMM_CASE(5) BMM_CASE(6)
                BMM_CASE(7) BMM_CASE(8) BMM_CASE(9) BMM_CASE(10) BMM_CASE(11) BMM_CASE(12)
#undef BMM_CASE
            default: break;
            }
        }
    }
}

namespace {
struct WinWorkspace {
    DevBuf tgts, pats, mults, w8, keys, counter;
};
WinWorkspace& win_ws() {
    thread_local WinWorkspace ws;
    return ws;
}
} // namespace

int default_persistent_wgs();

// Host: unique (pattern, signed multiplicity) per target, then one persistent
// launch over every target's blocks.  Results: weight and lowest window mask.
void window_coset_min(const uint64_t* rows, size_t n_rows, size_t words, size_t cols,
                      const int32_t* window_rows, size_t width, const int32_t* targets,
                      size_t n_targets, uint32_t* out_w, uint64_t* out_mask) {
    if (width == 0 || width > 47)
        fail(BMM_ERR_INVALID, "window_coset_min: window width must be in [1, 47]");
    if (cols > words * 64)
        fail(BMM_ERR_INVALID, "window_coset_min: cols exceeds words_per_row * 64");
    if (cols >= (1ull << (63 - width)))
        fail(BMM_ERR_OVERFLOW, "window_coset_min: too many columns for packed keys");
    std::vector<WinTarget> tg(n_targets);
    std::vector<uint64_t> all_p;
    std::vector<int32_t> all_m;
    std::vector<WinPat8> all_w8;  // per-target lists for packed-int16 targets
    const uint64_t tail = cols % 64 ? ((1ull << (cols % 64)) - 1) : ~0ull;
    for (size_t i = 0; i < n_targets; ++i) {
        const int t = targets[i];
        if (t < 0 || (size_t)t >= n_rows)
            fail(BMM_ERR_RANGE, "window_coset_min: target row out of range");
        const int32_t* wr = window_rows + i * width;
        for (size_t j = 0; j < width; ++j) {
            if (wr[j] < 0 || (size_t)wr[j] >= n_rows)
                fail(BMM_ERR_RANGE, "window_coset_min: window row out of range");
            if (wr[j] == t)
                fail(BMM_ERR_INVALID, "window_coset_min: target inside its own window");
        }
        // (pattern << 1 | target bit) per active column, sorted to aggregate.
        std::vector<uint64_t> keyed;
        int32_t total = 0;
        for (size_t wi = 0; wi * 64 < cols; ++wi) {
            uint64_t active = rows[(size_t)t * words + wi];
            for (size_t j = 0; j < width; ++j)
                active |= rows[(size_t)wr[j] * words + wi];
            if ((wi + 1) * 64 >= cols)
                active &= tail;
            while (active) {
                const int bit = __builtin_ctzll(active);
                uint64_t p = 0;
                for (size_t j = 0; j < width; ++j)
                    p |= ((rows[(size_t)wr[j] * words + wi] >> bit) & 1ull) << j;
                const uint64_t tb = (rows[(size_t)t * words + wi] >> bit) & 1ull;
                keyed.push_back((p << 1) | tb);
                ++total;
                active &= active - 1;
            }
        }
        std::sort(keyed.begin(), keyed.end());
        WinTarget& T = tg[i];
        T.pat_off = (int64_t)all_p.size();
        for (size_t a = 0; a < keyed.size();) {
            const uint64_t p = keyed[a] >> 1;
            int32_t m = 0;
            size_t b = a;
            for (; b < keyed.size() && (keyed[b] >> 1) == p; ++b)
                m += (keyed[b] & 1) ? -1 : 1;
            if (m != 0) {
                all_p.push_back(p);
                all_m.push_back(m);
            }
            a = b;
        }
        T.n_pat = (int32_t)(all_p.size() - T.pat_off);
        // Packed lists: low-half entries, then high-half entries, each padded
        // to a multiple of 256 with zero entries at distinct words.
        T.top = width > (size_t)kPkL + 32 ? 1 : 0;
        T.lo_off = T.hi_off = 0;
        T.n_lo = T.n_hi = 0;
        if (width >= (size_t)kPkL && total <= kPkMaxTotal) {
            for (uint32_t half = 0; half < 2; ++half) {
                const int64_t off = (int64_t)all_w8.size();
                for (int64_t k = T.pat_off; k < (int64_t)all_p.size(); ++k) {
                    const uint32_t pl = (uint32_t)all_p[k] & ((1u << kPkL) - 1u);
                    if ((pl >> kPkW) != half)
                        continue;
                    const uint64_t hi = all_p[k] >> kPkL;
                    const uint32_t byte = 4u * (uint32_t)pk_pad((int)(pl & ((1u << kPkW) - 1u)));
                    all_w8.push_back(WinPat8{
                        (uint32_t)hi, ((uint32_t)(uint16_t)(int16_t)all_m[k] << 16) | byte |
                                          (uint32_t)((hi >> 32) & 1u)});
                }
                while ((all_w8.size() - off) % 256)
                    all_w8.push_back(WinPat8{
                        0u, 4u * (uint32_t)pk_pad((int)((all_w8.size() - off) &
                                                          ((1u << kPkW) - 1u)))});
                (half ? T.hi_off : T.lo_off) = off;
                (half ? T.n_hi : T.n_lo) = (int32_t)(all_w8.size() - off);
            }
        }
        T.bits = (int32_t)width;
        T.total = total;
        int L = (int)std::min<size_t>(width, kWinMaxL);
        while (L > 1 && ((int64_t)total + 1) << L >= (int64_t(1) << 31))
            --L;
        if (width >= (size_t)kPkL && total <= kPkMaxTotal)
            L = kPkL;
        T.L = L;
        T.slot = (int32_t)i;
    }
    if (n_targets == 0)
        return;
    // Three queues, one kernel each: packed 2^14 blocks, int32 2^13 blocks,
    // narrower blocks; block numbers restart per queue.
    auto queue_of = [](const WinTarget& T) { return T.L == kPkL ? 0 : T.L == kWinMaxL ? 1 : 2; };
    std::stable_sort(tg.begin(), tg.end(), [&](const WinTarget& a, const WinTarget& b) {
        return queue_of(a) < queue_of(b);
    });
    size_t n_in[3] = {0, 0, 0};
    int64_t q_blocks[3] = {0, 0, 0};
    for (WinTarget& T : tg) {
        const int qi = queue_of(T);
        ++n_in[qi];
        T.first_block = q_blocks[qi];
        q_blocks[qi] += int64_t(1) << (T.bits - T.L);
    }
    WinWorkspace& ws = win_ws();
    WinTarget* d_t = ws.tgts.get<WinTarget>(n_targets);
    uint64_t* d_p = ws.pats.get<uint64_t>(std::max<size_t>(all_p.size(), 1));
    int32_t* d_m = ws.mults.get<int32_t>(std::max<size_t>(all_m.size(), 1));
    WinPat8* d_w8 = ws.w8.get<WinPat8>(std::max<size_t>(all_w8.size(), 1));
    unsigned long long* d_k = ws.keys.get<unsigned long long>(n_targets);
    unsigned long long* d_c = ws.counter.get<unsigned long long>(3);
    BMM_HIP(hipMemcpy(d_t, tg.data(), sizeof(WinTarget) * n_targets, hipMemcpyHostToDevice));
    if (!all_p.empty()) {
        BMM_HIP(hipMemcpy(d_p, all_p.data(), all_p.size() * 8, hipMemcpyHostToDevice));
        BMM_HIP(hipMemcpy(d_m, all_m.data(), all_m.size() * 4, hipMemcpyHostToDevice));
        BMM_HIP(hipMemcpy(d_w8, all_w8.data(), all_w8.size() * sizeof(WinPat8),
                          hipMemcpyHostToDevice));
    }
    fill_u64(reinterpret_cast<uint64_t*>(d_k), n_targets, ~0ull, nullptr);
    fill_u64(reinterpret_cast<uint64_t*>(d_c), 3, 0ull, nullptr);
    auto shape = [](int64_t blocks, int& grid, int& chunk) {
        grid = (int)std::max<int64_t>(1, std::min<int64_t>(blocks, default_persistent_wgs()));
        chunk = blocks > (int64_t)grid * 64 ? 16 : 1;
    };
    int grid, chunk;
    const WinTarget* d_q[3] = {d_t, d_t + n_in[0], d_t + n_in[0] + n_in[1]};
    {
        // The timed kernel is the one holding the most blocks.
        const int timed = (int)(std::max_element(q_blocks, q_blocks + 3) - q_blocks);
        for (int qi = 0; qi < 3; ++qi) {
            if (q_blocks[qi] == 0)
                continue;
            shape(q_blocks[qi], grid, chunk);
            ScopedKernelTimer timer(nullptr, qi == timed);
            auto kern = qi == 0 ? win_full_kernel<kPkL>
                        : qi == 1 ? win_full_kernel<kWinMaxL> : win_min_kernel;
            hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, nullptr, d_q[qi], (int)n_in[qi],
                               q_blocks[qi], d_p, d_m, d_w8, chunk, d_c + qi, d_k);
            BMM_HIP(hipGetLastError());
        }
    }
    std::vector<unsigned long long> keys(n_targets);
    BMM_HIP(hipMemcpy(keys.data(), d_k, n_targets * 8, hipMemcpyDeviceToHost));
    const uint64_t mmask = (1ull << width) - 1;
    for (size_t i = 0; i < n_targets; ++i) {
        out_w[i] = (uint32_t)(keys[i] >> width);
        out_mask[i] = keys[i] & mmask;
    }
}

} // namespace bmmhip
