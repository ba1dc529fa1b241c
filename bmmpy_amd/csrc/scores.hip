// calc_scores_and_order (src/math/fwht.cpp:59-100) on the GPU: from a
// spectrum h of m entries and a column count n, the score of mask u is
// s[u] = clamp((n - h[u]) >> 1, 0, n) (the row-combination weight when h is
// the transform of the column-pattern histogram), s[0] = 0; cnt[v] counts
// the masks u = 1 .. m-1 of score v, and order lists those masks sorted by
// score, stably (ascending u within a score); off[v] ends as the end of
// bucket v in order (the reference's offsets after its scatter).
//
// Scores of at most 10 bits (n < 1024) take the one-pass tile order (the
// section "one-pass order" below): 2^14-mask tiles, each sorted stably in
// LDS and written run by run from a scanned per-tile score histogram, 16 B
// of traffic per mask in all (2^28 masks, n = 1023: 1.19 ms of kernels,
// against 1.94 ms for the radix passes).
//
// Wider scores take an LSD radix sort over digits of the score:
// ceil(bits(n) / 8) passes with the bits split evenly.  Two passes pack the
// mask and the score bits above the first digit into one 32-bit word between
// them when both fit, which saves 8 bytes of traffic per mask.  The scores
// kernel also writes the first pass's digit histogram per tile of 4096
// masks; later passes count their digits in a histogram kernel.  On the
// first pass the masks themselves are the values.  The digit-major table
// (digit, tile) is scanned exclusively on the device, which gives each tile
// the global start of each digit's run.  The scatter stages a tile in LDS:
// each wave ranks its 512 contiguous masks (8 slices of 64; equal digits
// matched with one ballot per digit bit, running per-wave digit counts in
// LDS), the tile is reordered by digit in LDS (one word per entry where the
// value can be rebuilt from it), and the runs are written out with
// consecutive lanes on consecutive addresses.  Tiles of 4096 keep the
// scatter at 57-73 VGPRs and 26-43 KiB of LDS, so 3-4 workgroups share a CU.
#include "common.hpp"

#include <algorithm>
#include <array>
#include <type_traits>
#include <vector>

namespace bmmhip {

template <typename T> void fwht_array_dev(T* d, size_t n, hipStream_t stream);

// Device results of one scores-and-order call (this thread's workspace).
template <typename S> struct ScoresDev {
    S* s;
    int32_t* order;
    uint32_t* cnt;
    uint32_t* off;
};

namespace {

constexpr int kRsTile = 4096;   // masks per tile
constexpr int kRsWaves = 8;     // waves of a scatter workgroup
constexpr int kRsThreads = kRsWaves * kWave;
constexpr int kRsSlices = kRsTile / kRsThreads;   // 8 slices of 64 per wave
constexpr int kScanBlock = 1024;
constexpr int kLdsBuckets = 16384;
constexpr int kTileKeyBits = 10;      // one-pass order: scores below 2^10
constexpr int kOrderTileBits = 14;    // masks per tile sort (2^14: 66 KiB of LDS, two tiles per CU)
constexpr uint32_t kScoresGrid = 1024;

template <typename S>
struct Vec4;
template <>
struct Vec4<int32_t> {
    using T = int4;
};
template <>
struct Vec4<int16_t> {
    using T = short4;
};

// Tiles are ranges of kRsTile indices: masks u (the first pass, u = 1 .. m-1
// valid) or positions in the previous pass's output (later passes).  A
// workgroup reads 4 consecutive entries per lane and vector load.
//
// Scores of masks 1 .. m-1, s[0] = 0, the score histogram cnt (in LDS when
// lds_hist, then one atomic per bucket and workgroup) and the first radix
// pass's digit histogram per tile: hist0[d * n_tiles + tile] for the low
// dbits bits of the score.  Tiles go round-robin over the workgroups, so the
// grid sweeps the arrays front to back together.
template <typename S>
__global__ __launch_bounds__(256) void scores_kernel(const S* __restrict__ h, uint64_t m, int32_t n,
                                                     S* __restrict__ s_out,
                                                     unsigned int* __restrict__ cnt, int lds_hist,
                                                     unsigned int* __restrict__ overflow,
                                                     uint32_t n_tiles, int dbits,
                                                     uint32_t* __restrict__ hist0) {
    using V = typename Vec4<S>::T;
    constexpr int kVecs = kRsTile / (256 * 4);
    extern __shared__ unsigned int sh_cnt[];
    __shared__ unsigned int dig[256];
    const int buckets = n + 1;
    const uint32_t dmask = (1u << dbits) - 1u;
    if (lds_hist)
        for (int q = threadIdx.x; q < buckets; q += blockDim.x)
            sh_cnt[q] = 0;
    for (uint32_t v = blockIdx.x; v < n_tiles; v += gridDim.x) {
        // The grid sweeps windows of gridDim.x tiles front to back; inside a
        // full window each XCD takes a contiguous range of tiles, so the
        // digit table's entries for neighbouring tiles (one cache line per 16
        // tiles and digit) are written from one L2.
        const uint32_t w0 = v - v % gridDim.x, b = v % gridDim.x;
        const uint32_t tile = gridDim.x % kXcds == 0 && w0 + gridDim.x <= n_tiles
                                  ? w0 + (b % kXcds) * (gridDim.x / kXcds) + b / kXcds
                                  : v;
        // Thread t alone reads dig[t] below and clears it here: no barrier
        // between the two.
        dig[threadIdx.x] = 0;
        __syncthreads();
        const uint64_t base = (uint64_t)tile * kRsTile + 4u * threadIdx.x;
        S hv[kVecs][4];
#pragma unroll
        for (int k = 0; k < kVecs; ++k) {
            const uint64_t u = base + (uint64_t)k * 1024;
            if (u + 4 <= m) {
                const V v = *reinterpret_cast<const V*>(h + u);
                hv[k][0] = v.x;
                hv[k][1] = v.y;
                hv[k][2] = v.z;
                hv[k][3] = v.w;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    hv[k][j] = u + j < m ? h[u + j] : S(0);
            }
        }
#pragma unroll
        for (int k = 0; k < kVecs; ++k) {
            const uint64_t u = base + (uint64_t)k * 1024;
            S sv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                // 64-bit difference: n - h[u] cannot overflow here.
                int64_t su = ((int64_t)n - (int64_t)hv[k][j]) >> 1;
                su = su < 0 ? 0 : (su > n ? n : su);
                const bool valid = u + j >= 1 && u + j < m;
                sv[j] = valid ? S(su) : S(0);   // s[0] = 0
                if (!valid)
                    continue;
                if (std::is_same<S, int16_t>::value && su > 32767)
                    atomicOr(overflow, 1u);
                atomicAdd(&dig[(uint32_t)su & dmask], 1u);
                if (lds_hist)
                    atomicAdd(&sh_cnt[su], 1u);
                else
                    atomicAdd(&cnt[su], 1u);
            }
            if (u + 4 <= m) {
                V v;
                v.x = sv[0];
                v.y = sv[1];
                v.z = sv[2];
                v.w = sv[3];
                *reinterpret_cast<V*>(s_out + u) = v;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (u + j < m)
                        s_out[u + j] = sv[j];
            }
        }
        __syncthreads();
        if (threadIdx.x <= dmask)
            hist0[(size_t)threadIdx.x * n_tiles + tile] = dig[threadIdx.x];
    }
    if (lds_hist) {
        __syncthreads();
        for (int q = threadIdx.x; q < buckets; q += blockDim.x)
            if (sh_cnt[q])
                atomicAdd(&cnt[q], sh_cnt[q]);
    }
}

// hist[d * n_tiles + tile] = entries of the tile (positions below N) whose
// digit is d.
__global__ __launch_bounds__(256) void radix_hist_kernel(const int32_t* __restrict__ keys,
                                                         uint32_t N, int shift, int dbits,
                                                         uint32_t n_tiles,
                                                         uint32_t* __restrict__ hist) {
    constexpr int kVecs = kRsTile / (256 * 4);
    __shared__ unsigned int c[256];
    const uint32_t dmask = (1u << dbits) - 1u;
    c[threadIdx.x] = 0;
    const uint64_t base = (uint64_t)blockIdx.x * kRsTile + 4u * threadIdx.x;
    int32_t kv[kVecs][4];
#pragma unroll
    for (int k = 0; k < kVecs; ++k) {
        const uint64_t p = base + (uint64_t)k * 1024;
        if (p + 4 <= N) {
            const int4 v = *reinterpret_cast<const int4*>(keys + p);
            kv[k][0] = v.x;
            kv[k][1] = v.y;
            kv[k][2] = v.z;
            kv[k][3] = v.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                kv[k][j] = p + j < N ? keys[p + j] : 0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kVecs; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (base + (uint64_t)k * 1024 + j < N)
                atomicAdd(&c[((uint32_t)kv[k][j] >> shift) & dmask], 1u);
    __syncthreads();
    if (threadIdx.x <= dmask)
        hist[(size_t)threadIdx.x * n_tiles + blockIdx.x] = c[threadIdx.x];
}

// Where a scatter pass takes its values from: the entry index i (the masks,
// on the first pass), an array, or the low vbits bits of the key (a packed
// word mask | rest-of-score << vbits, written by a packing first pass).
enum : int { kValIndex = 0, kValArray = 1, kValInKey = 2 };
constexpr int kRsTileBits = 12;   // log2 kRsTile
static_assert((1 << kRsTileBits) == kRsTile, "tile size");

// Stable scatter of one digit pass over the entries lo <= i < hi (tile t:
// i in [t kRsTile, (t + 1) kRsTile)); start = the scanned digit-major table.
// keys_out == nullptr writes values only (the last pass, or a packing pass).
// pack_out: the value written is val | (key >> (shift + dbits)) << vbits,
// the score bits later passes need stored above the mask.
//
// ONE: the tile is staged in LDS as one word per entry, which halves the
// workgroup's LDS (more workgroups per CU hide the load and LDS latency):
// the key itself when the value is part of it, or the entry's offset in the
// tile | key << 12 for index values (keys below 2^20).
template <typename KI, int VMODE, bool ONE>
__global__ __launch_bounds__(kRsThreads, ONE ? 8 : 4) void radix_scatter_kernel(
    const KI* __restrict__ keys_in, const int32_t* __restrict__ vals_in, uint64_t lo, uint64_t hi,
    int shift, int dbits, const uint32_t* __restrict__ start, uint32_t n_tiles,
    int32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out, int vbits, int pack_out) {
    static_assert(!ONE || VMODE != kValArray, "array values need their own stage");
    __shared__ int32_t skey[kRsTile];
    __shared__ int32_t sval[ONE ? 1 : kRsTile];
    __shared__ uint32_t run[kRsWaves][256];   // per-wave digit counts, then wave offsets
    __shared__ uint32_t lstart[256];          // tile-local start of each digit
    __shared__ uint32_t gstart[256];          // global start of each digit's run
    __shared__ uint32_t wsum[kRsWaves];
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const uint32_t dmask = (1u << dbits) - 1u;
    const uint32_t vmask = (uint32_t)((1ull << vbits) - 1ull);
    // Neighbouring tiles write neighbouring parts of each digit's region:
    // they go to one XCD.
    const uint32_t tile = xcd_tile(blockIdx.x, n_tiles);
    const uint64_t tile0 = (uint64_t)tile * kRsTile;
    for (int i = tid; i < kRsWaves * 256; i += kRsThreads)
        (&run[0][0])[i] = 0;
    if (tid < 256)
        gstart[tid] = (uint32_t)tid <= dmask ? start[(size_t)tid * n_tiles + tile] : 0u;
    // Wave w owns the tile's entries [512 w, 512 w + 512), slice k of
    // them at lanes 0..63; offsets in the tile are 32-bit, valid in
    // [r_lo, r_hi).
    const uint32_t rbase = (uint32_t)wave * (kRsSlices * kWave) + lane;
    const uint32_t r_lo = lo > tile0 ? (uint32_t)(lo - tile0) : 0u;
    const uint32_t r_hi = hi - tile0 < (uint64_t)kRsTile ? (uint32_t)(hi - tile0) : kRsTile;
    int32_t key[kRsSlices], val[kRsSlices];
    uint32_t pos[kRsSlices];
#pragma unroll
    for (int k = 0; k < kRsSlices; ++k) {
        const uint32_t r = rbase + k * kWave;
        key[k] = r >= r_lo && r < r_hi ? (int32_t)keys_in[tile0 + r] : 0;
    }
    // The value source is a template parameter: a select per slice inside
    // the unrolled loops measured 1.7x slower.
    if constexpr (!ONE) {
        if constexpr (VMODE == kValArray) {
#pragma unroll
            for (int k = 0; k < kRsSlices; ++k) {
                const uint32_t r = rbase + k * kWave;
                val[k] = r >= r_lo && r < r_hi ? vals_in[tile0 + r] : 0;
            }
        } else if constexpr (VMODE == kValIndex) {
#pragma unroll
            for (int k = 0; k < kRsSlices; ++k)
                val[k] = (int32_t)((uint32_t)tile0 + rbase + k * kWave);
        } else if constexpr (VMODE == kValInKey) {
#pragma unroll
            for (int k = 0; k < kRsSlices; ++k)
                val[k] = (int32_t)((uint32_t)key[k] & vmask);
        }
        if (pack_out) {
#pragma unroll
            for (int k = 0; k < kRsSlices; ++k)
                val[k] = (int32_t)((uint32_t)val[k] |
                                   (((uint32_t)key[k] >> (shift + dbits)) << vbits));
        }
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kRsSlices; ++k) {
        const uint32_t r = rbase + k * kWave;
        const bool valid = r >= r_lo && r < r_hi;
        const uint32_t d = ((uint32_t)key[k] >> shift) & dmask;
        const uint64_t match = match_digit(d, dbits, valid);
        const uint32_t rank = (uint32_t)__popcll(match & below);
        // All lanes read the wave's count before the one lane per digit
        // updates it; a wave's LDS operations complete in issue order.
        const uint32_t before = run[wave][d];
        pos[k] = before + rank;
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0)
            run[wave][d] = before + (uint32_t)__popcll(match);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // Threads 0..255, digit tid: exclusive offsets of the waves, tile count
    // c, then the exclusive scan of c over the digits.
    uint32_t c = 0, x = 0;
    if (tid < 256) {
#pragma unroll
        for (int w = 0; w < kRsWaves; ++w) {
            const uint32_t r = run[w][tid];
            run[w][tid] = c;
            c += r;
        }
        x = c;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, kWave);
            if (lane >= o)
                x += y;
        }
        if (lane == kWave - 1)
            wsum[wave] = x;
    }
    __syncthreads();
    if (tid < 256) {
        uint32_t ex = x - c;
        for (int q = 0; q < wave; ++q)
            ex += wsum[q];
        lstart[tid] = ex;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRsSlices; ++k) {
        const uint32_t r = rbase + k * kWave;
        if (r >= r_lo && r < r_hi) {
            const uint32_t d = ((uint32_t)key[k] >> shift) & dmask;
            const uint32_t slot = lstart[d] + run[wave][d] + pos[k];
            if constexpr (!ONE) {
                skey[slot] = key[k];
                sval[slot] = val[k];
            } else if constexpr (VMODE == kValIndex) {
                skey[slot] = (int32_t)(r | ((uint32_t)key[k] << kRsTileBits));
            } else {
                skey[slot] = key[k];
            }
        }
    }
    __syncthreads();
    // Valid entries of the tile: slots 0 .. tile_n-1.
    const uint32_t tile_n = r_hi - r_lo;
    // Slot tid + kRsThreads q goes to the start of its digit's run plus its
    // offset in the run: consecutive lanes write consecutive addresses.
#pragma unroll
    for (int q = 0; q < kRsSlices; ++q) {
        const uint32_t j = tid + (uint32_t)kRsThreads * q;
        if (j >= tile_n)
            continue;
        const uint32_t w = (uint32_t)skey[j];
        const uint32_t k = ONE && VMODE == kValIndex ? w >> kRsTileBits : w;   // the key
        const uint32_t d = (k >> shift) & dmask;
        const uint32_t g = gstart[d] + (j - lstart[d]);
        uint32_t v;
        if constexpr (!ONE) {
            v = (uint32_t)sval[j];
        } else if constexpr (VMODE == kValIndex) {
            v = (uint32_t)tile0 + (w & (kRsTile - 1));
            if (pack_out)
                v |= (k >> (shift + dbits)) << vbits;
        } else {
            v = w & vmask;
        }
        if (keys_out)
            keys_out[g] = (int32_t)k;
        vals_out[g] = (int32_t)v;
    }
}

// Scan of blocks of 1024 entries in place (exclusive or inclusive);
// sums[b] = total of block b.
__global__ __launch_bounds__(256) void scan_block_kernel(uint32_t* __restrict__ data, uint64_t L,
                                                         uint32_t* __restrict__ sums,
                                                         int inclusive) {
    __shared__ uint32_t wsum[4];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint64_t b0 = (uint64_t)blockIdx.x * kScanBlock + threadIdx.x * 4u;
    uint32_t v[4], t = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = b0 + q < L ? data[b0 + q] : 0u;
        t += v[q];
    }
    uint32_t x = t;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, kWave);
        if (lane >= o)
            x += y;
    }
    if (lane == kWave - 1)
        wsum[wave] = x;
    __syncthreads();
    uint32_t run = x - t;
    for (int q = 0; q < wave; ++q)
        run += wsum[q];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t before = run;
        run += v[q];
        if (b0 + q < L)
            data[b0 + q] = inclusive ? run : before;
    }
    if (threadIdx.x == blockDim.x - 1)
        sums[blockIdx.x] = run;
}

__global__ __launch_bounds__(256) void scan_add_kernel(uint32_t* __restrict__ data, uint64_t L,
                                                       const uint32_t* __restrict__ sums) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < L && i >= (uint64_t)kScanBlock)
        data[i] += sums[i / kScanBlock];
}

// ---------------------------------------------------- one-pass order ----
// Scores of at most 10 bits (n < 1024) are ordered in one pass over tiles
// of 2^TB masks: the scores kernel writes each tile's score histogram as a
// row of 16-bit counts (counts[tile * Kp + k]; a tile holds at most 2^TB <=
// 2^16 masks of a score), a column scan in chunks of tiles turns them into
// the 32-bit global start of every (score, tile) run, and each tile
// is sorted stably in LDS (two digit passes on the packed word
// score << TB | offset; the offsets never leave LDS) and written out run by
// run.  Per mask the
// call then reads h and the scores once and writes the scores and the order
// once (16 B), against 28 B for two global radix passes.  Mask 0 (score 0,
// the first entry of tile 0) and the offsets past m (read as score 0, last
// among the tile's score-0 entries) are sorted with the rest but not written.

template <typename S, int TB>
__global__ __launch_bounds__(256) void scores_tile_kernel(const S* __restrict__ h, uint64_t m,
                                                          int32_t n, S* __restrict__ s_out,
                                                          unsigned int* __restrict__ cnt,
                                                          unsigned int* __restrict__ overflow,
                                                          uint32_t n_tiles,
                                                          uint16_t* __restrict__ counts) {
    static_assert(TB <= 16, "a tile's count of one score must fit 16 bits");
    using V = typename Vec4<S>::T;
    constexpr int kVecs = 4;                          // 4 vectors of 4 entries per thread
    constexpr int kSubs = (1 << TB) / (256 * 4 * kVecs);
    extern __shared__ unsigned int sh[];
    const int K = n + 1;
    unsigned int* th = sh;          // this tile's counts
    unsigned int* tot = sh + K;     // the workgroup's counts over all its tiles
    for (int q = threadIdx.x; q < K; q += blockDim.x)
        tot[q] = 0;
    for (uint32_t v = blockIdx.x; v < n_tiles; v += gridDim.x) {
        // Windows of gridDim.x tiles, front to back.
        const uint32_t tile = v;
        // Thread t reads and clears the same entries (q = t mod 256): the
        // barrier after the clear is the only one the table write needs.
        for (int q = threadIdx.x; q < K; q += blockDim.x)
            th[q] = 0;
        __syncthreads();
        for (int sub = 0; sub < kSubs; ++sub) {
            const uint64_t base =
                ((uint64_t)tile << TB) + (uint64_t)sub * (256 * 4 * kVecs) + 4u * threadIdx.x;
            S hv[kVecs][4];
#pragma unroll
            for (int k = 0; k < kVecs; ++k) {
                const uint64_t u = base + (uint64_t)k * 1024;
                if (u + 4 <= m) {
                    const V x = *reinterpret_cast<const V*>(h + u);
                    hv[k][0] = x.x;
                    hv[k][1] = x.y;
                    hv[k][2] = x.z;
                    hv[k][3] = x.w;
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        hv[k][j] = u + j < m ? h[u + j] : S(0);
                }
            }
#pragma unroll
            for (int k = 0; k < kVecs; ++k) {
                const uint64_t u = base + (uint64_t)k * 1024;
                S sv[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    int64_t su = ((int64_t)n - (int64_t)hv[k][j]) >> 1;
                    su = su < 0 ? 0 : (su > n ? n : su);
                    const bool valid = u + j >= 1 && u + j < m;
                    sv[j] = valid ? S(su) : S(0);   // s[0] = 0
                    if (!valid)
                        continue;
                    if (std::is_same<S, int16_t>::value && su > 32767)
                        atomicOr(overflow, 1u);
                    atomicAdd(&th[su], 1u);
                }
                if (u + 4 <= m) {
                    V x;
                    x.x = sv[0];
                    x.y = sv[1];
                    x.z = sv[2];
                    x.w = sv[3];
                    *reinterpret_cast<V*>(s_out + u) = x;
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (u + j < m)
                            s_out[u + j] = sv[j];
                }
            }
        }
        __syncthreads();
        const int Kp = (K + 3) & ~3;
        for (int q = threadIdx.x; q < Kp; q += blockDim.x) {
            const unsigned int c = q < K ? th[q] : 0u;
            counts[(size_t)tile * Kp + q] = (uint16_t)c;
            if (q < K)
                tot[q] += c;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < K; q += blockDim.x)
        if (tot[q])
            atomicAdd(&cnt[q], tot[q]);
}

// LDS layouts of a tile sort: data entry i at i + i/32 (rows of 32 with a
// pad word, so a thread reading 32 consecutive entries and its neighbours
// hit distinct banks); digit counters as rows of T thread columns, two
// 16-bit counts per word (digits 2r and 2r + 1 in row r), one pad word per
// 16 columns.
// Column scan of the counts (rows = tiles, Kp = K rounded up to 4 entries
// per row, 4 keys per thread as one 8-byte access) in chunks of kScanRows
// rows: col_sum_kernel totals each chunk's columns, col_base_kernel turns the
// chunk totals into chunk starts (key-major: all of key k before key k + 1),
// and col_start_kernel walks each chunk's rows, writing every count's run
// start (32-bit, one 16-byte access per 4 keys) to the start table.  Rows are
// read 16 at a time so the loads overlap.
constexpr int kScanRows = 128;
constexpr int kScanBatch = 16;

// Four 16-bit counts: one 8-byte access, then as 32-bit values (zero when
// the row is not taken).
__device__ inline uint2 load_counts2(const uint16_t* p) {
    return *reinterpret_cast<const uint2*>(p);
}
__device__ inline uint4 unpack_counts4(uint2 v, bool take) {
    if (!take)
        v = make_uint2(0, 0);
    return make_uint4(v.x & 0xffffu, v.x >> 16, v.y & 0xffffu, v.y >> 16);
}

__global__ __launch_bounds__(256) void col_sum_kernel(const uint16_t* __restrict__ counts, int Kp,
                                                      uint32_t n_rows,
                                                      uint32_t* __restrict__ csum) {
    const uint32_t r0 = blockIdx.x * kScanRows, r1 = min(r0 + kScanRows, n_rows);
    const int k4 = (int)threadIdx.x * 4;
    if (k4 >= Kp)
        return;
    uint4 acc = make_uint4(0, 0, 0, 0);
    for (uint32_t r = r0; r < r1; r += kScanBatch) {
        // Unconditional loads at clamped rows (all 16 in flight), rows past
        // the chunk's end masked afterwards.
        uint2 v[kScanBatch];
#pragma unroll
        for (int i = 0; i < kScanBatch; ++i)
            v[i] = load_counts2(counts + (size_t)min(r + i, r1 - 1) * Kp + k4);
#pragma unroll
        for (int i = 0; i < kScanBatch; ++i) {
            const uint4 c = unpack_counts4(v[i], r + i < r1);
            acc.x += c.x;
            acc.y += c.y;
            acc.z += c.z;
            acc.w += c.w;
        }
    }
    *reinterpret_cast<uint4*>(csum + (size_t)blockIdx.x * Kp + k4) = acc;
}

// One workgroup of 1024 threads: thread (seg, g) holds keys 4g .. 4g+3 and
// a quarter of the chunks.  Each sums its quarter, the quarters' sums give
// each thread its prefix and the key totals, a scan over the keys gives the
// key starts, and each thread writes its chunks' starts.
__global__ __launch_bounds__(1024) void col_base_kernel(uint32_t* __restrict__ csum, int Kp,
                                                        uint32_t n_chunks) {
    __shared__ uint4 qsum[4][256];
    __shared__ uint32_t kbase[1024];
    __shared__ uint32_t wsum[4];
    const int g = threadIdx.x % 256, seg = threadIdx.x / 256;
    const int k4 = 4 * g;
    const uint32_t per = (n_chunks + 3) / 4;
    const uint32_t c0 = min((uint32_t)seg * per, n_chunks), c1 = min(c0 + per, n_chunks);
    uint4 tot = make_uint4(0, 0, 0, 0);
    if (k4 < Kp)
        for (uint32_t c = c0; c < c1; c += kScanBatch) {
            uint4 v[kScanBatch];
#pragma unroll
            for (int i = 0; i < kScanBatch; ++i)
                v[i] = c + i < c1 ? *reinterpret_cast<const uint4*>(csum + (size_t)(c + i) * Kp + k4)
                                  : make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < kScanBatch; ++i) {
                tot.x += v[i].x;
                tot.y += v[i].y;
                tot.z += v[i].z;
                tot.w += v[i].w;
            }
        }
    qsum[seg][g] = tot;
    __syncthreads();
    uint4 pre = make_uint4(0, 0, 0, 0), all = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = qsum[q][g];
        if (q < seg) {
            pre.x += v.x;
            pre.y += v.y;
            pre.z += v.z;
            pre.w += v.w;
        }
        all.x += v.x;
        all.y += v.y;
        all.z += v.z;
        all.w += v.w;
    }
    // Exclusive scan of the key totals, 4 per thread, over the threads of
    // the first quarter (g = tid for them; the others compute the same
    // values and drop them).
    const uint32_t t4 = all.x + all.y + all.z + all.w;
    const int wave = g / kWave, lane = g % kWave;
    uint32_t inc = t4;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, kWave);
        if (lane >= o)
            inc += y;
    }
    if (seg == 0 && lane == kWave - 1)
        wsum[wave] = inc;
    __syncthreads();
    if (seg == 0) {
        uint32_t b = inc - t4;
        for (int w = 0; w < wave; ++w)
            b += wsum[w];
        kbase[k4] = b;
        kbase[k4 + 1] = b + all.x;
        kbase[k4 + 2] = b + all.x + all.y;
        kbase[k4 + 3] = b + all.x + all.y + all.z;
    }
    __syncthreads();
    if (k4 >= Kp)
        return;
    uint4 run = make_uint4(kbase[k4] + pre.x, kbase[k4 + 1] + pre.y, kbase[k4 + 2] + pre.z,
                           kbase[k4 + 3] + pre.w);
    for (uint32_t c = c0; c < c1; c += kScanBatch) {
        uint4 v[kScanBatch];
#pragma unroll
        for (int i = 0; i < kScanBatch; ++i)
            v[i] = c + i < c1 ? *reinterpret_cast<const uint4*>(csum + (size_t)(c + i) * Kp + k4)
                              : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < kScanBatch; ++i)
            if (c + i < c1) {
                *reinterpret_cast<uint4*>(csum + (size_t)(c + i) * Kp + k4) = run;
                run.x += v[i].x;
                run.y += v[i].y;
                run.z += v[i].z;
                run.w += v[i].w;
            }
    }
}

__global__ __launch_bounds__(256) void col_start_kernel(const uint16_t* __restrict__ counts,
                                                        int Kp, uint32_t n_rows,
                                                        const uint32_t* __restrict__ csum,
                                                        uint32_t* __restrict__ start) {
    const uint32_t r0 = blockIdx.x * kScanRows, r1 = min(r0 + kScanRows, n_rows);
    const int k4 = (int)threadIdx.x * 4;
    if (k4 >= Kp)
        return;
    uint4 run = *reinterpret_cast<const uint4*>(csum + (size_t)blockIdx.x * Kp + k4);
    for (uint32_t r = r0; r < r1; r += kScanBatch) {
        uint2 v[kScanBatch];
#pragma unroll
        for (int i = 0; i < kScanBatch; ++i)
            v[i] = load_counts2(counts + (size_t)min(r + i, r1 - 1) * Kp + k4);
#pragma unroll
        for (int i = 0; i < kScanBatch; ++i)
            if (r + i < r1) {
                *reinterpret_cast<uint4*>(start + (size_t)(r + i) * Kp + k4) = run;
                const uint4 c = unpack_counts4(v[i], true);
                run.x += c.x;
                run.y += c.y;
                run.z += c.z;
                run.w += c.w;
            }
    }
}

typedef int v4i32 __attribute__((ext_vector_type(4)));
// LDS layouts of a tile sort: data entry i at tile_pad(i) = i + i/32 where
// threads read 32 consecutive entries each (one pad word per 32 puts the
// threads of a wave on distinct banks), at i where they read in stripes;
// digit counters as rows of T thread columns (row = digit) with one pad word
// per 16 columns.
__device__ inline uint32_t tile_pad(uint32_t i) { return i + (i >> 5); }
template <int T> __device__ inline int cnt_at(int r, int c) { return r * (T + T / 16) + c + (c >> 4); }
constexpr int kMaxDigitBits = 5;

// One stable digit pass over a tile of E T entries held E per thread in
// blocked order (thread t: entries t E .. t E + E - 1): entry j of thread t
// goes to rank = (entries of smaller digit) + (entries of its digit in
// threads below t) + (its digit among x[0 .. j-1]), stored at
// lds[tile_pad(rank)] when PAD (the next pass reads blocked) and at
// lds[rank] otherwise.  Each thread counts its digits in its own counter
// column; a raking scan (thread q owns R consecutive counter words of one
// row, the row's threads scan their totals with lane shuffles) and a scan
// of the digit totals turn the counts into starts, which the thread then
// advances for its entries.  Counter updates are LDS atomics on the
// thread's own column: uncontended, completed in issue order, so a
// thread's updates go out back to back and the returned values are its
// ranks.  The counters share LDS with the data: starts and ends with a
// barrier.
template <int E, int T, int DB, bool PAD>
__device__ inline void tile_rank_pass(const uint32_t (&x)[E], int shift, uint32_t* lds,
                                      uint32_t* rowtot, uint32_t* dbase) {
    static_assert(DB >= 1 && DB <= kMaxDigitBits, "digits of 1 to 5 bits");
    // Counter rows: one per digit, at least T / 64 so that a row's raking
    // threads (G of them) share a wave.
    constexpr int R = (1 << DB) > T / kWave ? (1 << DB) : T / kWave;
    constexpr int G = T / R;
    static_assert(G >= 1 && G <= kWave && (G & (G - 1)) == 0, "a row's threads share a wave");
    constexpr uint32_t dmask = (1u << DB) - 1u;
    const int t = threadIdx.x;
    __syncthreads();   // the data in lds may still be read
#pragma unroll
    for (int r = 0; r < R; ++r)
        lds[cnt_at<T>(r, t)] = 0;
    // Each thread counts in its own column: no barrier before the counts.
#pragma unroll
    for (int j = 0; j < E; ++j)
        atomicAdd(&lds[cnt_at<T>((int)((x[j] >> shift) & dmask), t)], 1u);
    __syncthreads();
    const int row = t / G, g = t % G, c0 = R * g;
    uint32_t v[R];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const uint32_t c = lds[cnt_at<T>(row, c0 + k)];
        v[k] = sum;
        sum += c;
    }
    uint32_t inc = sum;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, G);
        if (g >= o)
            inc += y;
    }
    const uint32_t ex = inc - sum;
    if (g == G - 1)
        rowtot[row] = inc;
    __syncthreads();
    if (t < kWave) {
        // lane = digit: exclusive start over the digit totals
        const uint32_t c = t <= (int)dmask ? rowtot[t] : 0u;
        uint32_t x2 = c;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t y = __shfl_up(x2, o, kWave);
            if (t >= o)
                x2 += y;
        }
        dbase[t] = x2 - c;
    }
    __syncthreads();
    const uint32_t base = row <= (int)dmask ? dbase[row] : 0u;
#pragma unroll
    for (int k = 0; k < R; ++k)
        lds[cnt_at<T>(row, c0 + k)] = v[k] + ex + base;
    __syncthreads();
    // A fresh copy of the shift: the digits are recomputed here rather than
    // kept live (in registers) from the counting sweep.
    int sh2 = shift;
    asm volatile("" : "+s"(sh2));
    uint32_t pos[E];
#pragma unroll
    for (int j = 0; j < E; ++j)
        pos[j] = atomicAdd(&lds[cnt_at<T>((int)((x[j] >> sh2) & dmask), t)], 1u);
    __syncthreads();   // every start read before the data overwrites the counters
#pragma unroll
    for (int j = 0; j < E; ++j)
        lds[PAD ? tile_pad(pos[j]) : pos[j]] = x[j];
}

// The order of one tile of 2^TB masks from the scores s (KB-bit keys);
// start is the scanned table: start[tile * Kp + k] is the global position
// of the first mask of key k from this tile (Kp = K rounded up to 4).
template <typename S, int TB, int KB>
__global__ __launch_bounds__((1 << TB) / 32) void order_tile_kernel(
    const S* __restrict__ s, uint64_t m, int K, const uint32_t* __restrict__ start,
    uint32_t n_tiles, int32_t* __restrict__ order) {
    constexpr int E = 32;
    constexpr int T = (1 << TB) / E;
    constexpr uint32_t kOff = (1u << TB) - 1u;
    constexpr int D1 = (KB + 1) / 2, D2 = KB - D1;
    constexpr int kRows = (1 << kMaxDigitBits) > T / kWave ? (1 << kMaxDigitBits) : T / kWave;
    constexpr int kData = (1 << TB) + (1 << TB) / 32;
    constexpr int kCnt = kRows * (T + T / 16);
    __shared__ uint32_t lds[kData > kCnt ? kData : kCnt];
    __shared__ uint32_t gst[1 << kTileKeyBits];   // per key: global start - local start
    __shared__ uint32_t rowtot[kRows];
    __shared__ uint32_t dbase[kWave];
    const int tid = threadIdx.x;
    const uint32_t tile = xcd_tile(blockIdx.x, n_tiles);
    const uint64_t tile0 = (uint64_t)tile << TB;
    // Offsets lo <= p < lim hold masks 1 .. m-1; a buffer resource over the
    // tile keeps the offsets 32-bit.  Offsets past m count as score 0: they
    // come last among the tile's score-0 entries and are not written, nor is
    // mask 0 (score 0, the first entry of tile 0).
    const uint32_t lim = (uint32_t)std::min<uint64_t>((uint64_t)1 << TB, m - tile0);
    const uint32_t lo = tile0 == 0 ? 1u : 0u;
    // The range check is per dword: the range is rounded up to whole dwords
    // (the scores array has room for it) and entries at or past lim are
    // replaced below.
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<S*>(s + tile0), 0, (int)((lim * sizeof(S) + 3) & ~3u), 0x00020000);
    const bool whole = lo == 0 && lim == (1u << TB);
    const uint32_t p0 = (uint32_t)tid * E;
    uint32_t x[E];
    if constexpr (sizeof(S) == 4) {
#pragma unroll
        for (int q = 0; q < E / 4; ++q) {
            const v4i32 a = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)((p0 + 4 * q) * 4), 0, 0);
            x[4 * q] = (uint32_t)a[0];
            x[4 * q + 1] = (uint32_t)a[1];
            x[4 * q + 2] = (uint32_t)a[2];
            x[4 * q + 3] = (uint32_t)a[3];
        }
    } else {
#pragma unroll
        for (int q = 0; q < E / 8; ++q) {
            const v4i32 a = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)((p0 + 8 * q) * 2), 0, 0);
#pragma unroll
            for (int e = 0; e < 8; ++e)
                x[8 * q + e] = ((uint32_t)a[e >> 1] >> ((e & 1) * 16)) & 0xffffu;
        }
    }
    if (whole) {
#pragma unroll
        for (int j = 0; j < E; ++j)
            x[j] = (x[j] << TB) | (p0 + (uint32_t)j);
    } else {
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const uint32_t p = p0 + (uint32_t)j;
            x[j] = ((p < lim ? x[j] : 0u) << TB) | p;
        }
    }
    // Global starts of the keys (the loads overlap the first digit pass).
    for (int k = tid; k < K; k += T)
        gst[k] = start[(size_t)tile * ((K + 3) & ~3) + k];
    if constexpr (D2 > 0) {
        tile_rank_pass<E, T, D1, true>(x, TB, lds, rowtot, dbase);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < E; ++j)
            x[j] = lds[tile_pad(p0 + (uint32_t)j)];
        tile_rank_pass<E, T, D2, false>(x, TB + D1, lds, rowtot, dbase);
    } else {
        tile_rank_pass<E, T, D1, false>(x, TB, lds, rowtot, dbase);
    }
    __syncthreads();
    // The first entry of each key's run turns its global start into global -
    // local; in tile 0, mask 0 heads the run of score 0.
    uint32_t w[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const uint32_t i = (uint32_t)tid + (uint32_t)T * q;
        w[q] = lds[i];
        const uint32_t k = w[q] >> TB;
        if (i == 0 || (lds[i - 1] >> TB) != k)
            gst[k] -= i + (k == 0 ? lo : 0u);
    }
    __syncthreads();
    // Entry i of the sorted tile: consecutive lanes write consecutive
    // positions of a run.
    if (whole) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const uint32_t i = (uint32_t)tid + (uint32_t)T * q;
            order[gst[w[q] >> TB] + i] = (int32_t)(tile0 + (w[q] & kOff));
        }
    } else {
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const uint32_t i = (uint32_t)tid + (uint32_t)T * q;
            const uint32_t p = w[q] & kOff;
            if (p >= lo && p < lim)
                order[gst[w[q] >> TB] + i] = (int32_t)(tile0 + p);
        }
    }
}

struct ScoresWorkspace {
    DevBuf h, s, cnt, off, overflow, hist, order, table, counts, csum;
    DevBuf keys[2], vals[2];
    std::array<DevBuf, 6> levels;
};
ScoresWorkspace& scores_ws() {
    thread_local ScoresWorkspace ws;
    return ws;
}

// In-place scan of L entries on the device (totals below 2^32).
void scan_u32(uint32_t* d, uint64_t L, bool inclusive, std::array<DevBuf, 6>& levels, size_t depth,
              hipStream_t s) {
    const uint64_t nb = (L + kScanBlock - 1) / kScanBlock;
    if (depth >= levels.size())
        fail(BMM_ERR_OVERFLOW, "scan: input too long");
    uint32_t* sums = levels[depth].get<uint32_t>((size_t)nb);
    hipLaunchKernelGGL(scan_block_kernel, dim3((unsigned)nb), dim3(256), 0, s, d, L, sums,
                       inclusive ? 1 : 0);
    BMM_HIP(hipGetLastError());
    if (nb > 1) {
        scan_u32(sums, nb, false, levels, depth + 1, s);
        hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, s, d,
                           L, sums);
        BMM_HIP(hipGetLastError());
    }
}

} // namespace

// Device-resident core: d_h (m entries) on the device; the results stay in
// this thread's workspace (valid until the next call).
template <typename S>
ScoresDev<S> scores_order_dev(const S* d_h, size_t m, int32_t n) {
    ScoresWorkspace& ws = scores_ws();
    const size_t buckets = (size_t)n + 1;
    const uint64_t N = m - 1;   // masks 1 .. m-1
    // First pass: tiles over the masks 0 .. m-1; later passes: over the
    // positions 0 .. N-1.
    const uint32_t tiles_u = (uint32_t)((m + kRsTile - 1) / kRsTile);
    const uint32_t tiles_p = (uint32_t)((N + kRsTile - 1) / kRsTile);
    // Digits: ceil(bits / 8) passes, bits split evenly.  Two passes with
    // packing when the mask (vbits bits) and the score bits above the first
    // digit fit one 32-bit word: the first pass then writes
    // mask | (score >> d0) << vbits and the second reads both from it (8
    // bytes per mask fewer than separate keys and values).  The first digit
    // widens from half the bits up to 8 to make room.
    const int bits = n == 0 ? 1 : 32 - __builtin_clz((uint32_t)n);
    const int passes = (bits + 7) / 8;
    const int vbits = N > 0 ? 32 - __builtin_clz((uint32_t)(m - 1)) : 0;
    int dbits = (bits + passes - 1) / passes;
    bool packed = false;
    if (passes == 2)
        for (int d = dbits; d <= 8 && !packed; ++d)
            if (vbits + (bits - d) <= 32) {
                dbits = d;
                packed = true;
            }
    S* d_s = ws.s.get<S>(m + 2);   // two spare entries: whole-dword reads of the last tile
    uint32_t* d_cnt = ws.cnt.get<uint32_t>(buckets);
    uint32_t* d_off = ws.off.get<uint32_t>(buckets);
    uint32_t* d_ovf = ws.overflow.get<uint32_t>(1);
    // Scores of at most 10 bits take the one-pass tile order.
    const int kbits = n == 0 ? 1 : 32 - __builtin_clz((uint32_t)n);
    if (N > 0 && kbits <= kTileKeyBits && m <= ((uint64_t)1 << 31)) {
        constexpr int TB = kOrderTileBits;
        const uint32_t tiles = (uint32_t)((m + ((uint64_t)1 << TB) - 1) >> TB);
        uint32_t* d_tab = ws.table.get<uint32_t>(((buckets + 3) & ~(size_t)3) * tiles);
        uint16_t* d_tc = ws.counts.get<uint16_t>(((buckets + 3) & ~(size_t)3) * tiles);
        fill_u32(d_cnt, buckets, 0, nullptr);
        fill_u32(d_ovf, 1, 0, nullptr);
        {
            ScopedKernelTimer timer(nullptr);   // scores, counts, per-tile key histograms
            hipLaunchKernelGGL((scores_tile_kernel<S, TB>), dim3(std::min(tiles, kScoresGrid)),
                               dim3(256), 2 * buckets * 4, nullptr, d_h, (uint64_t)m, n, d_s,
                               d_cnt, d_ovf, tiles, d_tc);
            BMM_HIP(hipGetLastError());
        }
        uint32_t ovf = 0;
        BMM_HIP(hipMemcpy(&ovf, d_ovf, 4, hipMemcpyDeviceToHost));
        if (ovf)
            fail(BMM_ERR_OVERFLOW, "calc_scores_and_order: a score exceeds the int16 range");
        int32_t* d_order = ws.order.get<int32_t>((size_t)N);
        {
            ScopedKernelTimer timer(nullptr);   // scan of the table and the tile sorts
            const uint32_t chunks = (tiles + kScanRows - 1) / kScanRows;
            const int Kp = (int)((buckets + 3) & ~(size_t)3);
            uint32_t* d_cs = ws.csum.get<uint32_t>((size_t)chunks * Kp);
            hipLaunchKernelGGL(col_sum_kernel, dim3(chunks), dim3(256), 0, nullptr, d_tc, Kp,
                               tiles, d_cs);
            hipLaunchKernelGGL(col_base_kernel, dim3(1), dim3(1024), 0, nullptr, d_cs, Kp, chunks);
            hipLaunchKernelGGL(col_start_kernel, dim3(chunks), dim3(256), 0, nullptr, d_tc, Kp,
                               tiles, d_cs, d_tab);
            BMM_HIP(hipGetLastError());
            switch (kbits) {
#define BMM_KB(KB)                                                                             \
    case KB:                                                                                   \
        hipLaunchKernelGGL((order_tile_kernel<S, TB, KB>), dim3(tiles), dim3((1 << TB) / 32),  \
                           0, nullptr, d_s, (uint64_t)m, (int)buckets, d_tab, tiles, d_order); \
        break;
                BMM_KB(1) BMM_KB(2) BMM_KB(3) BMM_KB(4) BMM_KB(5) BMM_KB(6) BMM_KB(7) BMM_KB(8)
                BMM_KB(9) BMM_KB(10)
#undef BMM_KB
            default: fail(BMM_ERR_INVALID, "calc_scores_and_order: bad key width");
            }
            BMM_HIP(hipGetLastError());
        }
        BMM_HIP(hipMemcpy(d_off, d_cnt, buckets * 4, hipMemcpyDeviceToDevice));
        scan_u32(d_off, buckets, true, ws.levels, 0, nullptr);
        return ScoresDev<S>{d_s, d_order, d_cnt, d_off};
    }
    uint32_t* d_hist = ws.hist.get<uint32_t>((size_t)256 * std::max(tiles_u, tiles_p));
    fill_u32(d_cnt, buckets, 0, nullptr);
    fill_u32(d_ovf, 1, 0, nullptr);
    const bool lds = buckets <= (size_t)kLdsBuckets;
    const unsigned grid = std::min(tiles_u, 2048u);
    {
        ScopedKernelTimer timer(nullptr);   // scores, counts, first digit histogram
        hipLaunchKernelGGL(scores_kernel<S>, dim3(grid), dim3(256), lds ? buckets * 4 : 0, nullptr,
                           d_h, (uint64_t)m, n, d_s, d_cnt, lds ? 1 : 0, d_ovf, tiles_u, dbits,
                           d_hist);
        BMM_HIP(hipGetLastError());
    }
    uint32_t ovf = 0;
    BMM_HIP(hipMemcpy(&ovf, d_ovf, 4, hipMemcpyDeviceToHost));
    if (ovf)
        fail(BMM_ERR_OVERFLOW, "calc_scores_and_order: a score exceeds the int16 range");
    int32_t* d_order = ws.order.get<int32_t>(std::max<uint64_t>(N, 1));
    if (N > 0) {
        ScopedKernelTimer timer(nullptr);   // the radix passes
        int32_t* keys[2] = {ws.keys[0].get<int32_t>(passes > 1 ? (size_t)N : 1),
                            ws.keys[1].get<int32_t>(passes > 2 ? (size_t)N : 1)};
        int32_t* vals[2] = {ws.vals[0].get<int32_t>(passes > 1 && !packed ? (size_t)N : 1),
                            ws.vals[1].get<int32_t>(passes > 2 ? (size_t)N : 1)};
        const int32_t* kin = nullptr;
        const int32_t* vin = nullptr;
        for (int p = 0; p < passes; ++p) {
            // Packed: the second digit sits at bit vbits of the packed word.
            const int shift = packed ? (p == 0 ? 0 : vbits) : p * dbits;
            const int db = packed ? (p == 0 ? dbits : bits - dbits)
                                  : std::min(dbits, bits - shift);
            const bool last = p == passes - 1;
            const uint32_t tiles = p == 0 ? tiles_u : tiles_p;
            int32_t* kout = last || packed ? nullptr : keys[p & 1];
            int32_t* vout = last ? d_order : packed ? keys[0] : vals[p & 1];
            if (p > 0) {
                hipLaunchKernelGGL(radix_hist_kernel, dim3(tiles), dim3(256), 0, nullptr, kin,
                                   (uint32_t)N, shift, db, tiles, d_hist);
                BMM_HIP(hipGetLastError());
            }
            scan_u32(d_hist, ((uint64_t)1 << db) * tiles, false, ws.levels, 0, nullptr);
            // The scores of masks 1 .. m-1 with the masks as values (one LDS
            // word per entry for scores below 2^20), the packed words, or
            // keys and values.
            if (p == 0 && bits <= 32 - kRsTileBits)
                hipLaunchKernelGGL((radix_scatter_kernel<S, kValIndex, true>), dim3(tiles),
                                   dim3(kRsThreads), 0, nullptr, d_s, nullptr, (uint64_t)1,
                                   (uint64_t)m, shift, db, d_hist, tiles, kout, vout, vbits,
                                   packed ? 1 : 0);
            else if (p == 0)
                hipLaunchKernelGGL((radix_scatter_kernel<S, kValIndex, false>), dim3(tiles),
                                   dim3(kRsThreads), 0, nullptr, d_s, nullptr, (uint64_t)1,
                                   (uint64_t)m, shift, db, d_hist, tiles, kout, vout, vbits,
                                   packed ? 1 : 0);
            else if (packed)
                hipLaunchKernelGGL((radix_scatter_kernel<int32_t, kValInKey, true>), dim3(tiles),
                                   dim3(kRsThreads), 0, nullptr, kin, nullptr, (uint64_t)0, N,
                                   shift, db, d_hist, tiles, kout, vout, vbits, 0);
            else
                hipLaunchKernelGGL((radix_scatter_kernel<int32_t, kValArray, false>), dim3(tiles),
                                   dim3(kRsThreads), 0, nullptr, kin, vin, (uint64_t)0, N, shift,
                                   db, d_hist, tiles, kout, vout, vbits, 0);
            BMM_HIP(hipGetLastError());
            kin = packed ? vout : kout;
            vin = vout;
        }
    }
    BMM_HIP(hipMemcpy(d_off, d_cnt, buckets * 4, hipMemcpyDeviceToDevice));
    scan_u32(d_off, buckets, true, ws.levels, 0, nullptr);
    return ScoresDev<S>{d_s, d_order, d_cnt, d_off};
}

template <typename S>
void scores_order(const S* h, size_t m, int32_t n, S* s_by_mask, int32_t* order, int32_t* cnt,
                  int32_t* off) {
    ScoresWorkspace& ws = scores_ws();
    S* d_h = ws.h.get<S>(m);
    BMM_HIP(hipMemcpy(d_h, h, m * sizeof(S), hipMemcpyHostToDevice));
    const ScoresDev<S> r = scores_order_dev<S>(d_h, m, n);
    const size_t buckets = (size_t)n + 1;
    if (m > 1)
        BMM_HIP(hipMemcpy(order, r.order, (m - 1) * 4, hipMemcpyDeviceToHost));
    BMM_HIP(hipMemcpy(s_by_mask, r.s, m * sizeof(S), hipMemcpyDeviceToHost));
    BMM_HIP(hipMemcpy(cnt, r.cnt, buckets * 4, hipMemcpyDeviceToHost));
    BMM_HIP(hipMemcpy(off, r.off, buckets * 4, hipMemcpyDeviceToHost));
}

namespace {
// b[pattern] += multiplicity (the column-pattern histogram of a window).
__global__ void pattern_hist_kernel(const uint32_t* __restrict__ pat,
                                    const int32_t* __restrict__ mult, size_t U,
                                    int32_t* __restrict__ b) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < U)
        atomicAdd(&b[pat[i]], mult[i]);
}

struct RankWorkspace {
    DevBuf b, pat, mult;
};

RankWorkspace& rank_ws() {
    thread_local RankWorkspace ws;
    return ws;
}
} // namespace

// The first min(k, 2^t - 1) masks of a t-row window in (weight, mask)
// order, for any k: the column-pattern histogram (U distinct t-bit patterns
// with multiplicities, total = their sum) is built on the device, transformed
// in HBM (weight(m) = (total - H[m]) / 2, src/search/fwht_search.cpp:59-112),
// and calc_scores_and_order's stable radix order by score = weight (ties by
// ascending mask) is the reference's order (fwht_search.cpp:15-22).
std::vector<bmm_topk_entry> ranked_topk(const uint32_t* pat, const int32_t* mult, size_t U, int t,
                                        int32_t total, size_t k) {
    if (t < 1 || t > 30)
        fail(BMM_ERR_INVALID, "max_candidates above 4096 needs a window of 1 to 30 rows");
    if (total < 0 || total > (1 << 28))
        fail(BMM_ERR_INVALID, "max_candidates above 4096 needs a window of at most 2^28 columns");
    const size_t m = (size_t)1 << t;
    RankWorkspace& ws = rank_ws();
    int32_t* d_b = ws.b.get<int32_t>(m);
    fill_u32(reinterpret_cast<uint32_t*>(d_b), m, 0u, nullptr);
    if (U > 0) {
        uint32_t* d_pat = ws.pat.get<uint32_t>(U);
        int32_t* d_mult = ws.mult.get<int32_t>(U);
        BMM_HIP(hipMemcpy(d_pat, pat, U * 4, hipMemcpyHostToDevice));
        BMM_HIP(hipMemcpy(d_mult, mult, U * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(pattern_hist_kernel, dim3((unsigned)((U + 255) / 256)), dim3(256), 0,
                           nullptr, d_pat, d_mult, U, d_b);
        BMM_HIP(hipGetLastError());
    }
    fwht_array_dev<int32_t>(d_b, m, nullptr);
    const ScoresDev<int32_t> r = scores_order_dev<int32_t>(d_b, m, total);
    const size_t K = std::min(k, m - 1);
    std::vector<int32_t> order(K);
    std::vector<uint32_t> off((size_t)total + 1);
    if (K)
        BMM_HIP(hipMemcpy(order.data(), r.order, K * 4, hipMemcpyDeviceToHost));
    BMM_HIP(hipMemcpy(off.data(), r.off, off.size() * 4, hipMemcpyDeviceToHost));
    std::vector<bmm_topk_entry> out(K);
    uint32_t v = 0;
    for (size_t i = 0; i < K; ++i) {
        while (off[v] <= i)
            ++v;
        out[i].mask = (uint64_t)(uint32_t)order[i];
        out[i].weight = v;
        out[i].pad = 0;
    }
    return out;
}

template void scores_order<int32_t>(const int32_t*, size_t, int32_t, int32_t*, int32_t*, int32_t*,
                                    int32_t*);
template void scores_order<int16_t>(const int16_t*, size_t, int32_t, int16_t*, int32_t*, int32_t*,
                                    int32_t*);

} // namespace bmmhip
