// Island-model genetic algorithm over row-XOR sequences, fully on the GPU.
//
// Reference semantics kept (src/ga): exact fitness = total Hamming weight of
// the transformed matrix (genetic_algorithm_internal.hpp:85-103 sums candidate
// weights; here the individual *is* the transform, as a sequence of elementary
// row XORs, and fitness is the popcount after replaying it); elitism +
// tournament selection (genetic_algorithm_operators.cpp:39-56); mutation as
// row-combination moves (:100-129); islands exchanging their best every
// interval (island.cpp:149-228), here as a ring.
//
// Representation: island i holds a base matrix B_i (rows x words, HBM) and a
// population of op sequences (src, dst) relative to B_i, at most max_ops long;
// replaying ops on B_i gives the individual's matrix.  Every migrate_every
// generations the island folds its best sequence into B_i (so progress is not
// capped by max_ops), and each island adopts its ring predecessor's base when
// that one is lighter.
//
// Kernels per generation (captured in one hipGraph per buffer parity):
//   ga_eval_kernel  - one wavefront per (island, individual, 2048-column
//                     slice): the slice of B_i (rows x 64 u32) is staged in
//                     LDS, the ops are applied lane-parallel across the 64
//                     words, the popcount is atomically added to fitness;
//   ga_breed_kernel - one workgroup per island, one thread per child: best
//                     tracking, elitism, tournament, one-point crossover of the
//                     op lists, mutation (append / replace / delete an op).
// Proposals for new ops are uniform random pairs (plain) or drawn from a
// per-island list of pairs ranked by the proposal MLP (learned).
#include "common.hpp"
#include "fwht_pk16.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <vector>

namespace bmmhip {

struct Mlp;
Mlp* mlp_create(uint64_t seed);
void mlp_destroy(Mlp* M);
float* mlp_forward_dev(Mlp* M, const __bf16* d_x, int B, hipStream_t s);
float* mlp_score_pairs_dev(Mlp* M, const uint64_t* d_sketch, int rows, const uint16_t* d_pairs,
                           const int32_t* d_pair_isl, int n, hipStream_t s);
float* mlp_backward_dev(Mlp* M, const __bf16* d_x, const float* d_t, int B, int valid,
                        hipStream_t s);
void mlp_adam_dev(Mlp* M, float lr, hipStream_t s);

void replay_host(uint64_t* rows, size_t n_rows, size_t words, const uint16_t* ops, size_t n_ops,
                 uint64_t* weight);
int default_persistent_wgs();

__device__ inline uint64_t xs_next(uint64_t& s) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}

// ------------------------------------------------------------- eval ------
// rows <= 256; u32 words per row = W32 (multiple of 64).  One wavefront per
// (island, individual, 2048-column slice).  Exact fitness from the rows the
// op list touches: fit (initialised to the base weight) += sum over those rows
// of popcount(after) - popcount(before), so untouched rows cost nothing.  The
// touched rows of the slice are staged in LDS slots (row -> slot table), the
// ops applied lane-parallel over the slice's 64 words.
// One task (island, individual, slice) on one wavefront.  `slots` rows x 64
// words of LDS at `cols` hold the touched rows; returns false, with nothing
// computed, when the op list touches more rows than that (the task is then
// left to the large tier).
__device__ inline bool ga_eval_task(int task, int slots, uint32_t* __restrict__ cols,
                                    uint16_t* __restrict__ slot_of, uint16_t* __restrict__ row_of,
                                    const uint32_t* __restrict__ base,
                                    const uint16_t* __restrict__ ops,
                                    const int32_t* __restrict__ len, int pop, int rows, int W32,
                                    int max_ops, unsigned int* __restrict__ fit) {
    // LDS: slots x 64 words for the touched rows, then the ops padded to a
    // multiple of 4 (first as loaded, then as src slot | dst slot << 16).
    uint32_t* opsl = cols + slots * 64;
    const int slices = W32 / 64;
    const int slice = task % slices;
    const int ind = (task / slices) % pop;
    const int isl = task / (slices * pop);
    const int lane = threadIdx.x;
    const size_t id = (size_t)isl * pop + ind;
    const uint16_t* o = ops + id * max_ops;
    const int n = len[id];
    __syncthreads();   // the previous task's reads of LDS are done
    for (int r = lane; r < 256; r += kWave)
        slot_of[r] = 0;
    __syncthreads();
    for (int k = lane; k < n; k += kWave) {
        const uint32_t op = o[k];
        const int src = op & 0xff, dst = op >> 8;
        opsl[k] = op;
        if (src != dst && src < rows && dst < rows) {
            slot_of[src] = 1;
            slot_of[dst] = 1;
        }
    }
    for (int k = n + lane; k < ((n + 3) & ~3); k += kWave)
        opsl[k] = 0u;   // pad to a multiple of 4 with no-ops (src == dst)
    __syncthreads();
    // Slots in row order: lane l owns rows 4l .. 4l + 3.
    uint32_t f[4], cnt = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        f[q] = slot_of[4 * lane + q];
        cnt += f[q];
    }
    uint32_t incl = cnt;   // inclusive scan over the wavefront
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, kWave);
        if (lane >= off)
            incl += v;
    }
    const int n_touched = (int)__shfl(incl, kWave - 1, kWave);
    if (n_touched > slots)
        return false;
    uint32_t slot = incl - cnt;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = 4 * lane + q;
        if (f[q]) {
            slot_of[r] = (uint16_t)(slot + 1);
            row_of[slot] = (uint16_t)r;
            ++slot;
        }
    }
    __syncthreads();
    // Ops as slot pairs (no-ops as 0xffffffff).
    for (int k = lane; k < n; k += kWave) {
        const uint32_t op = opsl[k];
        const int src = op & 0xff, dst = op >> 8;
        opsl[k] = (src != dst && src < rows && dst < rows)
                      ? (uint32_t)(slot_of[src] - 1) | ((uint32_t)(slot_of[dst] - 1) << 16)
                      : 0xffffffffu;
    }
    for (int k = n + lane; k < ((n + 3) & ~3); k += kWave)
        opsl[k] = 0xffffffffu;
    const uint32_t* b = base + (size_t)isl * rows * W32 + (size_t)slice * 64 + lane;
    uint32_t before = 0;
    int j = 0;
    for (; j + 16 <= n_touched; j += 16) {   // sixteen row loads in flight
        uint32_t v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)
            v[q] = b[(size_t)row_of[j + q] * W32];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            cols[(j + q) * 64 + lane] = v[q];
            before += __builtin_popcount(v[q]);
        }
    }
    for (; j + 4 <= n_touched; j += 4) {
        uint32_t v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            v[q] = b[(size_t)row_of[j + q] * W32];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            cols[(j + q) * 64 + lane] = v[q];
            before += __builtin_popcount(v[q]);
        }
    }
    for (; j < n_touched; ++j) {
        const uint32_t v = b[(size_t)row_of[j] * W32];
        cols[j * 64 + lane] = v;
        before += __builtin_popcount(v);
    }
    __syncthreads();
    const uint4* op4 = reinterpret_cast<const uint4*>(opsl);
    for (int k = 0; k < (n + 3) >> 2; ++k) {
        const uint4 q = op4[k];   // four ops per LDS read
        const uint32_t qs[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (qs[e] != 0xffffffffu)
                cols[(qs[e] >> 16) * 64 + lane] ^= cols[(qs[e] & 0xffffu) * 64 + lane];
    }
    uint32_t after = 0;
    for (int t = 0; t < n_touched; ++t)
        after += __builtin_popcount(cols[t * 64 + lane]);
    const int32_t delta = (int32_t)wave_sum_u32(after) - (int32_t)wave_sum_u32(before);
    if (lane == 0)
        atomicAdd(&fit[id], (unsigned int)delta);   // wraps back to the exact total
    return true;
}

// Task order: workgroups go to the 8 XCDs in turn; each XCD takes a
// contiguous eighth of the tasks (whole islands when the island count is a
// multiple of 8), so the island bases it reads stay in its own L2.
__device__ inline int ga_eval_task_of(unsigned int i, unsigned int n) {
    return (n & 7u) == 0u ? (int)((i & 7u) * (n >> 3) + (i >> 3)) : (int)i;
}

// Tiers of row slots in LDS: the first tier (one wavefront per task, 32
// slots: many wavefronts per CU) takes every task whose ops touch at most 32
// rows and flags the others 1; each later tier takes the tasks flagged with
// its level and flags one level up those touching more rows than it holds;
// the last holds every row.  Early in an epoch op lists are long; once the
// information-set trials have lightened the bases they stay a few ops long.
constexpr int kGaEvalSlots0 = 32;
constexpr int kGaEvalTiers = 4;   // later tiers: 64, 128, 192 and all rows
constexpr int kGaEvalSlots[kGaEvalTiers] = {64, 128, 192, 256};
constexpr int kGaEvalPerCu[kGaEvalTiers] = {8, 4, 3, 2};   // wavefronts per CU (LDS)

__global__ __launch_bounds__(64) void ga_eval_kernel(const uint32_t* __restrict__ base,
                                                     const uint16_t* __restrict__ ops,
                                                     const int32_t* __restrict__ len, int pop,
                                                     int rows, int W32, int max_ops,
                                                     unsigned int* __restrict__ fit,
                                                     int slots, uint8_t* __restrict__ tier) {
    extern __shared__ __align__(16) uint32_t cols[];
    __shared__ uint16_t slot_of[256];    // 0: untouched, else slot + 1
    __shared__ uint16_t row_of[256];
    const int task = ga_eval_task_of(blockIdx.x, gridDim.x);
    const bool done = ga_eval_task(task, slots, cols, slot_of, row_of, base, ops, len, pop, rows,
                                   W32, max_ops, fit);
    if (threadIdx.x == 0)
        tier[task] = done ? 0 : 1;
}

// A later tier: a fixed grid of wavefronts reading the flags 64 at a time
// and evaluating the tasks flagged `level` with `slots` row slots (flagging
// level + 1 a task that touches more rows).
__global__ __launch_bounds__(64) void ga_eval_tier_kernel(const uint32_t* __restrict__ base,
                                                          const uint16_t* __restrict__ ops,
                                                          const int32_t* __restrict__ len,
                                                          int pop, int rows, int W32,
                                                          int max_ops, int n_tasks, int slots,
                                                          int level,
                                                          unsigned int* __restrict__ fit,
                                                          uint8_t* __restrict__ tier) {
    extern __shared__ __align__(16) uint32_t cols[];
    __shared__ uint16_t slot_of[256];
    __shared__ uint16_t row_of[256];
    // Wavefront b takes launch indices b, b + G, b + 2 G, ... (G = the grid):
    // the same XCD as the first tier gave those tasks, and long and short op
    // lists spread over all wavefronts.
    const int G = (int)gridDim.x;
    for (int k0 = 0; (int)blockIdx.x + k0 * G < n_tasks; k0 += kWave) {
        const int i = (int)blockIdx.x + (k0 + (int)threadIdx.x) * G;
        const int ti = i < n_tasks ? ga_eval_task_of((unsigned int)i, (unsigned int)n_tasks) : 0;
        uint64_t m = __ballot(i < n_tasks && tier[ti] == level);
        while (m) {
            const int lane_k = __builtin_ctzll(m);
            m &= m - 1;
            const int task = __shfl(ti, lane_k, kWave);
            if (!ga_eval_task(task, slots, cols, slot_of, row_of, base, ops, len, pop, rows, W32,
                              max_ops, fit) &&
                threadIdx.x == 0)
                tier[task] = (uint8_t)(level + 1);
        }
    }
}

// fit[island, individual] = the island base's total weight.
__global__ void ga_fit_init_kernel(unsigned int* __restrict__ fit,
                                   const unsigned int* __restrict__ base_total, int pop, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n)
        fit[i] = base_total[i / pop];
}

// ------------------------------------------------------------ breed ------
constexpr int kMaxParents = 4;
// Learned crossover: the start of a later parent's segment is the best of
// kCutTries uniform draws by the number of ops among its next kCutSpan that
// the proposal ranks as weight-reducing (the island's candidate pairs).
constexpr int kCutTries = 4;
constexpr int kCutSpan = 8;

struct GaDev {
    int islands, pop, rows, max_ops, tournament, elite, n_cand, learned;
    float p_append, p_replace, p_delete, p_learned;
    int mutations;
    int num_parents, num_offspring;
    float mutation_rate;         // > 0: the reference's row-XOR mutation, adapted per generation
    int catastrophe_threshold;   // 0: off
    float catastrophe_survival;
    int restart_after;           // 0: off
    float restart_p;
};

// Per-island counters of the reference GA (genetic_algorithm.hpp: _no_improvement,
// _catastrophe, the adapted mutation rate).
struct GaIslandState {
    unsigned int no_improvement;
    unsigned int catastrophe;
    float rate;
    unsigned int pad;
};

__device__ inline float unit24(uint64_t& rng) {
    return (float)(xs_next(rng) >> 40) * (1.0f / 16777216.0f);
}

__device__ inline uint16_t propose(const GaDev& P, uint64_t& rng, const uint16_t* cand) {
    // Learned: a ranked pair with probability p_learned, else uniform (the
    // plain proposal stays the fallback).
    if (P.learned && P.n_cand > 0 && unit24(rng) < P.p_learned)
        return cand[xs_next(rng) % (uint64_t)P.n_cand];
    const int r = P.rows;
    const int s = (int)(xs_next(rng) % (uint64_t)r);
    int d = (int)(xs_next(rng) % (uint64_t)(r - 1));
    if (d >= s)
        ++d;
    return (uint16_t)(s | (d << 8));
}

// A random individual: the identity followed by 3 * rows random row XORs
// r_i ^= r_j, i == j skipped (genetic_algorithm_population.cpp:20-36), as
// ops (src j, dst i), at most max_ops of them.
__device__ inline int random_sequence(const GaDev& P, uint64_t& rng, uint16_t* child) {
    int clen = 0;
    for (int k = 0; k < 3 * P.rows && clen < P.max_ops; ++k) {
        const int i = (int)(xs_next(rng) % (uint64_t)P.rows);
        const int j = (int)(xs_next(rng) % (uint64_t)P.rows);
        if (i != j)
            child[clen++] = (uint16_t)(j | (i << 8));
    }
    return clen;
}

// Tournament of P.tournament uniform draws, lightest wins, first drawn on ties
// (genetic_algorithm_operators.cpp:39-56).
__device__ inline int tournament(const GaDev& P, uint64_t& rng, const unsigned int* sf) {
    int best = (int)(xs_next(rng) % (uint64_t)P.pop);
    for (int k = 1; k < P.tournament; ++k) {
        const int c = (int)(xs_next(rng) % (uint64_t)P.pop);
        if (sf[c] < sf[best])
            best = c;
    }
    return best;
}

// One workgroup per island, one thread per child.  Order of one generation
// (the reference's step(), genetic_algorithm_lifecycle.cpp:112-200, read from
// the evaluation of the population that was just bred):
//  1. rank the evaluated population by (fitness, index); a lighter best than
//     the island's best is recorded and resets the stale counters, else they
//     grow;
//  2. catastrophe when its counter reaches the threshold
//     (genetic_algorithm_operators.cpp:15-37): the floor(pop * survival) best
//     survive, the rest are random individuals;
//  3. otherwise the mutation rate adapts (genetic_algorithm_core.cpp:33-41),
//     the elite_count best are copied, and the other children come in groups
//     of num_offspring that share num_parents tournament winners and their
//     cut points: child = parent 0 ops [0, c0) ++ middle parent p ops
//     [s_p, e_p) ++ last parent ops [s, len) (two parents: one-point
//     crossover of the op lists); after restart_after stale generations a
//     child is instead a random individual with probability restart_p;
//  4. each child is mutated on its own stream: mutation_rate > 0 appends
//     max(1, floor(rows * rate)) row XORs r_i ^= r_j with i, j uniform and
//     i == j skipped (:100-129; a full sequence takes the op at a random
//     position instead); otherwise `mutations` draws of append / replace /
//     delete.
// LDS: fitness[pop] u32, order[pop] u16, then per group and parent (index,
// start, end) u16.
__global__ __launch_bounds__(1024) void ga_breed_kernel(
    GaDev P, const unsigned int* __restrict__ fit, const uint16_t* __restrict__ ops_cur,
    const int32_t* __restrict__ len_cur, uint16_t* __restrict__ ops_next,
    int32_t* __restrict__ len_next, uint64_t* __restrict__ rng_state,
    unsigned int* __restrict__ best_w, uint16_t* __restrict__ best_ops,
    int32_t* __restrict__ best_len, const uint16_t* __restrict__ cand,
    unsigned long long* __restrict__ evals, unsigned long long* __restrict__ n_ops,
    GaIslandState* __restrict__ state) {
    extern __shared__ unsigned int sf[];
    uint16_t* order = reinterpret_cast<uint16_t*>(sf + P.pop);
    uint16_t* gsel = order + ((P.pop + 1) & ~1);
    __shared__ uint32_t good[65536 / 32];   // learned: bitmap of the ranked pairs
    __shared__ int s_mode;
    __shared__ unsigned int s_ops;
    __shared__ unsigned int s_noimp;
    __shared__ float s_rate;
    const int isl = blockIdx.x;
    const int t = threadIdx.x;
    const size_t base_id = (size_t)isl * P.pop;
    if (t == 0)
        s_ops = 0u;
    if (t < P.pop)
        sf[t] = fit[base_id + t];
    __syncthreads();
    {
        // Ops the evaluation replayed on this island (one device atomic per
        // island instead of one per individual).
        const uint32_t c = wave_sum_u32(t < P.pop ? (uint32_t)len_cur[base_id + t] : 0u);
        if (t % kWave == 0 && c != 0u)
            atomicAdd(&s_ops, c);
    }
    if (t < P.pop) {
        const unsigned int f = sf[t];
        int r = 0;
        for (int j = 0; j < P.pop; ++j) {
            const unsigned int g = sf[j];
            r += (g < f) || (g == f && j < t);
        }
        order[r] = (uint16_t)t;
    }
    __syncthreads();
    const int arg = order[0];
    const unsigned int wbest = sf[arg];
    const bool improved = wbest < best_w[isl];
    __syncthreads();
    if (improved) {
        const uint16_t* src = ops_cur + (base_id + arg) * P.max_ops;
        for (int k = t; k < P.max_ops; k += blockDim.x)
            best_ops[(size_t)isl * P.max_ops + k] = src[k];
    }
    if (t == 0) {
        GaIslandState s = state[isl];
        if (improved) {
            best_w[isl] = wbest;
            best_len[isl] = len_cur[base_id + arg];
            s.no_improvement = 0;
            s.catastrophe = 0;
        } else {
            ++s.no_improvement;
            ++s.catastrophe;
        }
        int mode = 0;
        if (P.catastrophe_threshold > 0 && s.catastrophe == (unsigned int)P.catastrophe_threshold) {
            mode = 1;
            s.catastrophe = 0;
        }
        if (P.mutation_rate > 0.f) {
            if (s.no_improvement > 10)
                s.rate = fminf(0.8f, s.rate * 1.5f);
            else if (s.no_improvement > 5)
                s.rate = fminf(0.6f, s.rate * 1.3f);
            else
                s.rate = fmaxf(0.3f, s.rate * 0.95f);
        }
        state[isl] = s;
        s_mode = mode;
        s_noimp = s.no_improvement;
        s_rate = s.rate;
        atomicAdd(evals, (unsigned long long)P.pop);
    }
    __syncthreads();
    if (t == 0 && s_ops != 0u)
        atomicAdd(n_ops, (unsigned long long)s_ops);
    __syncthreads();
    const int mode = s_mode;
    const bool active = t < P.pop;
    uint64_t rng = active ? rng_state[base_id + t] : 0;
    uint16_t* child = ops_next + (base_id + (active ? t : 0)) * P.max_ops;
    int clen = 0;
    int kind = 0;   // 0 copy of order[t], 1 random, 2 crossover child
    int src_copy = 0;
    if (active) {
        if (mode == 1) {
            const int survivors = (int)((float)P.pop * P.catastrophe_survival);
            kind = t < survivors ? 0 : 1;
            src_copy = t;
        } else if (t < P.elite) {
            kind = 0;
            src_copy = t;
        } else if (P.restart_after > 0 && s_noimp > (unsigned int)P.restart_after &&
                   unit24(rng) < P.restart_p) {
            kind = 1;
        } else {
            kind = 2;
        }
    }
    const bool learned_cut = P.learned && P.n_cand > 0;
    if (learned_cut) {
        for (int k = t; k < 65536 / 32; k += blockDim.x)
            good[k] = 0u;
        __syncthreads();
        for (int k = t; k < P.n_cand; k += blockDim.x) {
            const uint32_t op = cand[(size_t)isl * P.n_cand + k];
            atomicOr(&good[op >> 5], 1u << (op & 31u));
        }
        __syncthreads();
    }
    // Group leaders (the first non-elite child of each group, whatever its
    // own kind) draw the parents and cut points of their group.
    const int c = t - P.elite;
    const int g = c >= 0 ? c / P.num_offspring : 0;
    if (active && mode == 0 && c >= 0 && c % P.num_offspring == 0) {
        uint16_t* gs = gsel + (size_t)g * P.num_parents * 3;
        for (int p = 0; p < P.num_parents; ++p) {
            const int par = tournament(P, rng, sf);
            const int lp = len_cur[base_id + par];
            int s0 = 0, e0 = lp;
            if (P.num_parents == 1) {
                // one parent: the child is its copy (then mutated)
            } else if (p == 0) {
                e0 = (int)(xs_next(rng) % (uint64_t)(lp + 1));
            } else {
                if (learned_cut) {
                    const uint16_t* src = ops_cur + (base_id + par) * P.max_ops;
                    int best_sc = -1;
                    for (int j = 0; j < kCutTries; ++j) {
                        const int cs = (int)(xs_next(rng) % (uint64_t)(lp + 1));
                        int sc = 0;
                        for (int k = cs; k < min(cs + kCutSpan, lp); ++k)
                            sc += (int)((good[src[k] >> 5] >> (src[k] & 31u)) & 1u);
                        if (sc > best_sc) {
                            best_sc = sc;
                            s0 = cs;
                        }
                    }
                } else {
                    s0 = (int)(xs_next(rng) % (uint64_t)(lp + 1));
                }
                if (p + 1 < P.num_parents)
                    e0 = s0 + (int)(xs_next(rng) % (uint64_t)(lp - s0 + 1));
            }
            gs[p * 3] = (uint16_t)par;
            gs[p * 3 + 1] = (uint16_t)s0;
            gs[p * 3 + 2] = (uint16_t)e0;
        }
    }
    __syncthreads();
    if (!active)
        return;
    if (kind == 0) {
        const int pick = order[src_copy];
        const uint16_t* src = ops_cur + (base_id + pick) * P.max_ops;
        clen = len_cur[base_id + pick];
        for (int k = 0; k < clen; ++k)
            child[k] = src[k];
    } else if (kind == 1) {
        clen = random_sequence(P, rng, child);
    } else {
        const uint16_t* gs = gsel + (size_t)g * P.num_parents * 3;
        for (int p = 0; p < P.num_parents; ++p) {
            const uint16_t* src = ops_cur + (base_id + gs[p * 3]) * P.max_ops;
            for (int k = gs[p * 3 + 1]; k < gs[p * 3 + 2] && clen < P.max_ops; ++k)
                child[clen++] = src[k];
        }
        const uint16_t* ic = cand + (size_t)isl * P.n_cand;
        if (P.mutation_rate > 0.f) {
            const int count = max(1, (int)((float)P.rows * s_rate));
            for (int m = 0; m < count; ++m) {
                uint16_t op;
                if (P.learned && P.n_cand > 0 && unit24(rng) < P.p_learned) {
                    op = ic[xs_next(rng) % (uint64_t)P.n_cand];
                } else {
                    const int i = (int)(xs_next(rng) % (uint64_t)P.rows);
                    const int j = (int)(xs_next(rng) % (uint64_t)P.rows);
                    if (i == j)
                        continue;
                    op = (uint16_t)(j | (i << 8));
                }
                if (clen < P.max_ops)
                    child[clen++] = op;
                else
                    child[xs_next(rng) % (uint64_t)clen] = op;
            }
        } else {
            for (int m = 0; m < P.mutations; ++m) {
                const float u = unit24(rng);
                if (u < P.p_append) {
                    if (clen < P.max_ops)
                        child[clen++] = propose(P, rng, ic);
                } else if (u < P.p_append + P.p_replace) {
                    if (clen > 0)
                        child[xs_next(rng) % (uint64_t)clen] = propose(P, rng, ic);
                } else if (u < P.p_append + P.p_replace + P.p_delete) {
                    if (clen > 0) {
                        const int pos = (int)(xs_next(rng) % (uint64_t)clen);
                        for (int k = pos; k + 1 < clen; ++k)
                            child[k] = child[k + 1];
                        --clen;
                    }
                }
            }
        }
    }
    len_next[base_id + t] = clen;
    rng_state[base_id + t] = rng;
}

// ------------------------------------------------------ epoch kernels ----
// Fold each island's best sequence into its base (one thread per u64 word
// column; the ops of one island run in order on every column).
__global__ void ga_rebase_kernel(uint64_t* __restrict__ base, int rows, int words,
                                 const uint16_t* __restrict__ best_ops,
                                 const int32_t* __restrict__ best_len, int max_ops) {
    const int isl = blockIdx.y;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words)
        return;
    uint64_t* b = base + (size_t)isl * rows * words + w;
    const uint16_t* o = best_ops + (size_t)isl * max_ops;
    const int n = best_len[isl];
    for (int k = 0; k < n; ++k) {
        const int src = o[k] & 0xff, dst = o[k] >> 8;
        if (src != dst && src < rows && dst < rows)
            b[(size_t)dst * words] ^= b[(size_t)src * words];
    }
}

// After a rebase the best sequence is empty; individual 0 of the current
// population (the elite slot) becomes the empty sequence.
__global__ void ga_after_rebase_kernel(int32_t* __restrict__ best_len, int32_t* __restrict__ len,
                                       int pop, int islands) {
    const int isl = blockIdx.x * blockDim.x + threadIdx.x;
    if (isl < islands) {
        best_len[isl] = 0;
        len[(size_t)isl * pop] = 0;
    }
}

// Ring migration: island i adopts island i-1's base when it is lighter, if
// island i-1 exports (mig bit 0) and island i imports (mig bit 1).  The
// reference publishes each island's best alone (genetic_algorithm_lifecycle.cpp:
// 76-81) and an import replaces the best only when lighter (:83-110).
__global__ void ga_migrate_decide_kernel(const unsigned int* __restrict__ best_w, int islands,
                                         const int32_t* __restrict__ mig,
                                         int32_t* __restrict__ flag,
                                         unsigned int* __restrict__ w_snap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= islands)
        return;
    const int prev = (i + islands - 1) % islands;
    flag[i] = ((mig[prev] & 1) && (mig[i] & 2) && best_w[prev] < best_w[i]) ? 1 : 0;
    w_snap[i] = best_w[prev];
}

__global__ void ga_migrate_copy_kernel(const uint64_t* __restrict__ src,
                                       uint64_t* __restrict__ dst, size_t per_island,
                                       const int32_t* __restrict__ flag, int islands, int shift) {
    const int i = blockIdx.y;
    if (!flag[i])
        return;
    const int from = (i + islands - shift) % islands;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < per_island;
         k += (size_t)gridDim.x * blockDim.x)
        dst[(size_t)i * per_island + k] = src[(size_t)from * per_island + k];
}

__global__ void ga_migrate_weights_kernel(unsigned int* __restrict__ best_w,
                                          const unsigned int* __restrict__ w_snap,
                                          const int32_t* __restrict__ flag, int islands) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < islands && flag[i])
        best_w[i] = w_snap[i];
}

// Per island row: weight and 128-bucket sketch (bucket = one u64 word of the
// 64-word-per-2 layout: bucket b covers words [b*W/128, (b+1)*W/128)); bit set
// when the bucket holds more ones than the row's average bucket.
__global__ void ga_sketch_kernel(const uint64_t* __restrict__ base, int rows, int words,
                                 uint32_t* __restrict__ roww, uint64_t* __restrict__ sketch) {
    const int row = blockIdx.x;   // global row index = isl * rows + r
    const uint64_t* b = base + (size_t)row * words;
    __shared__ uint32_t cnt[128];
    __shared__ uint32_t total;
    if (threadIdx.x == 0)
        total = 0;
    if (threadIdx.x < 128)
        cnt[threadIdx.x] = 0;
    __syncthreads();
    const int per = words >= 128 ? words / 128 : 1;
    for (int w = threadIdx.x; w < words; w += blockDim.x) {
        const uint32_t c = __builtin_popcountll(b[w]);
        atomicAdd(&cnt[min(w / per, 127)], c);
        atomicAdd(&total, c);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint64_t bits = 0;
        for (int k = 0; k < 64; ++k) {
            const int bucket = threadIdx.x * 64 + k;
            if ((uint64_t)cnt[bucket] * 128u > (uint64_t)total)
                bits |= 1ull << k;
        }
        sketch[(size_t)row * 2 + threadIdx.x] = bits;
    }
    if (threadIdx.x == 0)
        roww[row] = total;
}

// Features of pair (s -> d) on island isl: sketch(s) | sketch(d) as +-1 bf16.
__global__ void ga_features_kernel(const uint64_t* __restrict__ sketch, int rows,
                                   const uint16_t* __restrict__ pairs,
                                   const int32_t* __restrict__ pair_isl, int n,
                                   __bf16* __restrict__ x) {
    const int i = blockIdx.x;
    const int f = threadIdx.x;   // 0..255
    if (i >= n) {
        x[(size_t)i * 256 + f] = (__bf16)0.f;
        return;
    }
    const int isl = pair_isl[i];
    const int s = pairs[i] & 0xff, d = pairs[i] >> 8;
    const int row = f < 128 ? s : d;
    const int bit = f & 127;
    const uint64_t w = sketch[((size_t)isl * rows + row) * 2 + (bit >> 6)];
    x[(size_t)i * 256 + f] = (__bf16)(((w >> (bit & 63)) & 1u) ? 1.f : -1.f);
}

// Exact label of pair (s -> d): (wt(r_s ^ r_d) - wt(r_d)) / scale.
__global__ void ga_labels_kernel(const uint64_t* __restrict__ base, int rows, int words,
                                 const uint16_t* __restrict__ pairs,
                                 const int32_t* __restrict__ pair_isl, int n, float scale,
                                 float* __restrict__ t) {
    const int i = blockIdx.x;
    if (i >= n)
        return;
    const int isl = pair_isl[i];
    const int s = pairs[i] & 0xff, d = pairs[i] >> 8;
    const uint64_t* rs = base + ((size_t)isl * rows + s) * words;
    const uint64_t* rd = base + ((size_t)isl * rows + d) * words;
    int delta = 0;
    for (int w = threadIdx.x; w < words; w += blockDim.x)
        delta += __builtin_popcountll(rs[w] ^ rd[w]) - __builtin_popcountll(rd[w]);
    delta = (int)wave_sum_u32((uint32_t)delta);
    __shared__ int red[4];
    if (threadIdx.x % kWave == 0)
        red[threadIdx.x / kWave] = delta;
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int k = 0; k < (int)(blockDim.x + kWave - 1) / kWave; ++k)
            sum += red[k];
        t[i] = (float)sum / scale;
    }
}

__global__ void ga_random_pairs_kernel(uint64_t* __restrict__ rng, int rows, int per_island,
                                       int islands, uint16_t* __restrict__ pairs,
                                       int32_t* __restrict__ pair_isl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per_island * islands)
        return;
    uint64_t s = rng[i];
    const int a = (int)(xs_next(s) % (uint64_t)rows);
    int b = (int)(xs_next(s) % (uint64_t)(rows - 1));
    if (b >= a)
        ++b;
    rng[i] = s;
    pairs[i] = (uint16_t)(a | (b << 8));
    pair_isl[i] = i / per_island;
}

// Keeps the n_keep lowest-scored of per_island scored pairs per island
// (bitonic sort of (score, pair) in LDS; per_island a power of two <= 2048).
__global__ void ga_select_kernel(const float* __restrict__ score,
                                 const uint16_t* __restrict__ pairs, int per_island, int n_keep,
                                 uint16_t* __restrict__ cand) {
    __shared__ unsigned long long key[2048];
    const int isl = blockIdx.x;
    for (int i = threadIdx.x; i < per_island; i += blockDim.x) {
        float sc = score[(size_t)isl * per_island + i];
        uint32_t u = __float_as_uint(sc);
        u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);   // order-preserving
        key[i] = ((unsigned long long)u << 32) | pairs[(size_t)isl * per_island + i];
    }
    __syncthreads();
    for (int k = 2; k <= per_island; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < per_island; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long a = key[i], b = key[ixj];
                    if ((a > b) == ((i & k) == 0)) {
                        key[i] = b;
                        key[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < n_keep; i += blockDim.x)
        cand[(size_t)isl * n_keep + i] = (uint16_t)(key[i] & 0xffffu);
}

// Local improvement on the island bases (genetic_algorithm_operators.cpp:
// 131-192: replace row i by r_i ^ r_j for the j minimising
// wt(r_i) + wt(r_j) - 2 popcount(r_i & r_j) when that beats wt(r_i)).  The
// reference sweeps rows in order; here one round takes the rows whose index
// bit `bit` equals `phase` as targets and the other rows as sources, so all
// chosen moves are independent and apply in parallel; rounds over every index
// bit and both phases cover every pair.  One workgroup per (island, row):
// target rows score each source row by AND-popcount over the row words.
// Scores of one round as tiles of 32 targets x 32 sources per workgroup
// (grid: 16 tile pairs x islands; the k-th target of a round is row k with
// `phase` inserted at bit `bit`, the k-th source likewise with the other
// phase).  Word chunks of both row sets are staged in LDS word-major and
// each thread forms 2 x 2 overlaps pop(r_i & r_j); a pair with
// cw = wt(r_i) + wt(r_j) - 2 pop(r_i & r_j) below wt(r_i) offers the key
// cw << 32 | j, and atomicMin per target keeps the lightest, lowest-index
// source over all tiles (keys[] stays ~0 when none improves; the apply
// kernel consumes and resets it).
constexpr int kLiTile = 32;
constexpr int kLiChunk = 64;   // u64 words per LDS stage

__global__ __launch_bounds__(256) void ga_li_tile_kernel(const uint64_t* __restrict__ base,
                                                         int rows, int words,
                                                         const uint32_t* __restrict__ roww,
                                                         int bit, int phase,
                                                         unsigned long long* __restrict__ keys) {
    const int isl = blockIdx.y;
    const int tt = blockIdx.x / 4, st = blockIdx.x % 4;
    const int tid = threadIdx.x;
    __shared__ uint64_t A[kLiChunk][kLiTile + 1];
    __shared__ uint64_t B[kLiChunk][kLiTile + 1];
    __shared__ int ra[kLiTile], rb[kLiTile];
    __shared__ uint32_t wa[kLiTile], wb[kLiTile];
    const uint32_t lowm = (1u << bit) - 1u;
    if (tid < 2 * kLiTile) {
        const int t = tid % kLiTile;
        const bool tgt = tid < kLiTile;
        const int k = (tgt ? tt : st) * kLiTile + t;
        const int r = (int)((((uint32_t)k >> bit) << (bit + 1)) |
                            ((uint32_t)(tgt ? phase : 1 - phase) << bit) | ((uint32_t)k & lowm));
        const bool ok = r < rows;
        (tgt ? ra : rb)[t] = ok ? r : -1;
        (tgt ? wa : wb)[t] = ok ? roww[(size_t)isl * rows + r] : 0u;
    }
    __syncthreads();
    if (ra[0] < 0 || rb[0] < 0)
        return;   // rows grow with k: an empty first slot means an empty tile
    const uint64_t* b = base + (size_t)isl * rows * words;
    const int tx = tid % 16, ty = tid / 16;
    uint32_t acc[2][2] = {};
    for (int w0 = 0; w0 < words; w0 += kLiChunk) {
        const int cw = min(kLiChunk, words - w0);
        {
            // Every load of the thread in flight at once: clamped rows and
            // words, masked when stored.
            constexpr int kPer = kLiTile * kLiChunk / 256;
            uint64_t xa[kPer], xb[kPer];
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int q = tid + 256 * k, r = q / kLiChunk, w = min(q % kLiChunk, cw - 1);
                xa[k] = b[(size_t)max(ra[r], 0) * words + w0 + w];
                xb[k] = b[(size_t)max(rb[r], 0) * words + w0 + w];
            }
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int q = tid + 256 * k, r = q / kLiChunk, w = q % kLiChunk;
                A[w][r] = (w < cw && ra[r] >= 0) ? xa[k] : 0ull;
                B[w][r] = (w < cw && rb[r] >= 0) ? xb[k] : 0ull;
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int w = 0; w < cw; ++w) {
            const uint64_t a0 = A[w][ty], a1 = A[w][ty + 16];
            const uint64_t b0 = B[w][tx], b1 = B[w][tx + 16];
            const uint64_t x00 = a0 & b0, x01 = a0 & b1, x10 = a1 & b0, x11 = a1 & b1;
            acc[0][0] = bcnt_acc((uint32_t)(x00 >> 32), bcnt_acc((uint32_t)x00, acc[0][0]));
            acc[0][1] = bcnt_acc((uint32_t)(x01 >> 32), bcnt_acc((uint32_t)x01, acc[0][1]));
            acc[1][0] = bcnt_acc((uint32_t)(x10 >> 32), bcnt_acc((uint32_t)x10, acc[1][0]));
            acc[1][1] = bcnt_acc((uint32_t)(x11 >> 32), bcnt_acc((uint32_t)x11, acc[1][1]));
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int a = ty + 16 * i;
        unsigned long long key = ~0ull;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = tx + 16 * j;
            const uint32_t wi = wa[a], wj = wb[c];
            if (ra[a] >= 0 && rb[c] >= 0 && wi > 0u) {
                const uint32_t cw = wi + wj - 2u * acc[i][j];
                if (cw < wi)
                    key = min(key, ((unsigned long long)cw << 32) | (unsigned)rb[c]);
            }
        }
        // The 16 threads of one target row are 16 consecutive lanes.
#pragma unroll
        for (int off = 1; off < 16; off <<= 1)
            key = min(key, (unsigned long long)__shfl_xor(key, off, kWave));
        if (tx == 0 && key != ~0ull)
            atomicMin(&keys[(size_t)isl * rows + ra[a]], key);
    }
}

// Applies a round's moves: r_i ^= r_j for each target i whose key names a
// source j, with the new weight from the key; clears the key and counts the
// moves in changed[island] (one atomic per workgroup: thousands of moves on
// one address serialise).  One wavefront per target of the round (grid:
// targets / 8 x islands; the k-th target is row k with `phase` inserted at
// bit `bit`).
constexpr int kLiApplyWaves = 8;
__global__ __launch_bounds__(kLiApplyWaves * kWave) void ga_li_apply_kernel(uint64_t* __restrict__ base, int rows,
                                                          int words,
                                                          unsigned long long* __restrict__ keys,
                                                          uint32_t* __restrict__ roww, int bit,
                                                          int phase,
                                                          unsigned int* __restrict__ changed) {
    const int isl = blockIdx.y, lane = threadIdx.x % kWave;
    const uint32_t k = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const uint32_t lowm = (1u << bit) - 1u;
    const int i = (int)(((k >> bit) << (bit + 1)) | ((uint32_t)phase << bit) | (k & lowm));
    __shared__ uint32_t s_moves;
    if (threadIdx.x == 0)
        s_moves = 0;
    __syncthreads();
    const size_t id = (size_t)isl * rows + i;
    const unsigned long long key = i < rows ? keys[id] : ~0ull;
    if (key != ~0ull) {
        const int j = (int)(key & 0xffffffffu);
        uint64_t* b = base + (size_t)isl * rows * words;
        for (int w = lane; w < words; w += kWave)
            b[(size_t)i * words + w] ^= b[(size_t)j * words + w];
        if (lane == 0) {   // the wavefront's key load came before this store
            roww[id] = (uint32_t)(key >> 32);
            keys[id] = ~0ull;
            atomicAdd(&s_moves, 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_moves)
        atomicAdd(&changed[isl], s_moves);
}

// Window local improvement: for each target row i of a round (index bit `bit`
// equal to `phase`) and a window of W rows from the other half, the exact
// minimum over the 2^W subsets S of wt(r_i ^ XOR_S) (the coset minimum of the
// FWHT engine: signed column-pattern histogram, packed-int16 blocks of 2^14
// masks, fwht_pk16.hpp), applied when it beats wt(r_i).  Targets never serve
// as sources within a round, so all moves are independent row operations.

__device__ inline uint64_t ga_mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// Candidate sources per target: C distinct rows of the other half (all of
// them, in index order, when the half has at most C rows; -1 pads), one
// thread per (island, row).
__global__ void ga_win_pick_kernel(int islands, int rows, int bit, int phase, int C,
                                   uint64_t seed, int16_t* __restrict__ cand) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= islands * rows)
        return;
    const int i = id % rows;
    int16_t* out = cand + (size_t)id * C;
    if (((i >> bit) & 1) != phase)
        return;
    int nsrc = 0;
    for (int j = 0; j < rows; ++j)
        nsrc += ((j >> bit) & 1) != phase;
    if (nsrc <= C) {
        int c = 0;
        for (int j = 0; j < rows; ++j)
            if (((j >> bit) & 1) != phase)
                out[c++] = (int16_t)j;
        for (; c < C; ++c)
            out[c] = -1;
        return;
    }
    uint32_t used[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t st = ga_mix64(seed ^ ((uint64_t)id << 1 | 1ull));
    for (int c = 0; c < C;) {
        st = ga_mix64(st + 0x9E3779B97F4A7C15ull);
        const int j = (int)((st >> 32) % (uint64_t)rows);
        if (((j >> bit) & 1) == phase || ((used[j >> 5] >> (j & 31)) & 1u))
            continue;
        used[j >> 5] |= 1u << (j & 31);
        out[c++] = (int16_t)j;
    }
}

// Pairs (candidate -> target) for the proposal model, one per (island, row,
// candidate); non-target rows and padding get the pair (0 -> 0).
__global__ void ga_win_pairs_kernel(int islands, int rows, int C, const int16_t* __restrict__ cand,
                                    int bit, int phase, uint16_t* __restrict__ pairs,
                                    int32_t* __restrict__ pair_isl) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)islands * rows * C)
        return;
    const size_t id = k / C;
    const int i = (int)(id % rows);
    const int s = cand[k];
    const bool ok = ((i >> bit) & 1) == phase && s >= 0;
    pairs[k] = ok ? (uint16_t)(s | (i << 8)) : (uint16_t)0;
    pair_isl[k] = (int32_t)(id / rows);
}

// Window = the W candidates with the lowest score (learned: predicted delta;
// score == nullptr: the first W candidates).  One thread per (island, row).
__global__ void ga_win_rank_kernel(int islands, int rows, int C, int W, int bit, int phase,
                                   const int16_t* __restrict__ cand,
                                   const float* __restrict__ score,
                                   int16_t* __restrict__ win) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= islands * rows)
        return;
    const int i = id % rows;
    if (((i >> bit) & 1) != phase)
        return;
    const int16_t* c = cand + (size_t)id * C;
    int16_t* w = win + (size_t)id * W;
    if (!score) {
        for (int k = 0; k < W; ++k)
            w[k] = c[k];
        return;
    }
    const float* sc = score + (size_t)id * C;
    uint64_t taken[2] = {0, 0};   // C <= 128
    for (int k = 0; k < W; ++k) {
        int best = -1;
        float bv = 0.f;
        for (int q = 0; q < C; ++q) {
            if (c[q] < 0 || ((taken[q >> 6] >> (q & 63)) & 1u))
                continue;
            if (best < 0 || sc[q] < bv) {
                best = q;
                bv = sc[q];
            }
        }
        if (best < 0) {
            w[k] = -1;
            continue;
        }
        taken[best >> 6] |= 1ull << (best & 63);
        w[k] = c[best];
    }
}

// One workgroup per (island, row); target rows only.  Columns: thread t holds
// the CP = cols / 256 columns [t CP, (t + 1) CP) (cols <= 8192 and a multiple
// of 2048, so CP is 8, 16, 24 or 32)
// as W-bit window patterns with the target bit at bit 31.  Blocks mh of 2^14
// masks (2^(W-14) of them) are visited in order; the first block holding the
// largest H keeps the lowest mask.  move[id] = the subset mask (0: none).
__global__ __launch_bounds__(256, 2) void ga_win_kernel(const uint64_t* __restrict__ base, int rows,
                                                     int words, int W, int bit, int phase,
                                                     const int16_t* __restrict__ win,
                                                     const uint32_t* __restrict__ roww,
                                                     uint32_t* __restrict__ move,
                                                     uint32_t* __restrict__ neww) {
    __shared__ __align__(16) uint32_t s[pk_lds_words() + 4];
    __shared__ int32_t wbest[4];
    __shared__ uint32_t wcount[4];
    __shared__ int32_t wins[32];   // win_rows <= 24
    const int isl = blockIdx.x / rows;
    const int i = blockIdx.x % rows;
    if (((i >> bit) & 1) != phase)
        return;
    const size_t id = (size_t)isl * rows + i;
    const uint64_t* b = base + (size_t)isl * rows * words;
    if (threadIdx.x < W)
        wins[threadIdx.x] = win[id * W + threadIdx.x];
    __syncthreads();
    bool complete = true;
    for (int k = 0; k < W; ++k)
        complete &= wins[k] >= 0;
    if (!complete) {   // fewer than W sources: no move
        if (threadIdx.x == 0)
            move[id] = 0;
        return;
    }
    const int CP = words * 64 / 256;
    const int c0 = (int)threadIdx.x * CP;
    const uint32_t cmask = CP >= 32 ? 0xffffffffu : ((1u << CP) - 1u);
    auto chunk = [&](int r) -> uint32_t {
        const uint64_t* row = b + (size_t)r * words;
        const int w0 = c0 >> 6, off = c0 & 63;
        uint64_t v = row[w0] >> off;
        if (off + CP > 64)   // the thread's columns straddle two words (CP = 24)
            v |= row[w0 + 1] << (64 - off);
        return (uint32_t)v & cmask;
    };
    uint32_t pat[32];
    {
        const uint32_t xt = chunk(i);
#pragma unroll
        for (int k = 0; k < 32; ++k)
            pat[k] = ((xt >> k) & 1u) << 31;
    }
    for (int r = 0; r < W; ++r) {
        const uint32_t x = chunk(wins[r]);
#pragma unroll
        for (int k = 0; k < 32; ++k)
            pat[k] |= ((x >> k) & 1u) << r;
    }
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k)
        cnt += pat[k] != 0u;
    cnt = wave_sum_u32(cnt);
    if (threadIdx.x % kWave == 0)
        wcount[threadIdx.x / kWave] = cnt;
    __syncthreads();
    const int32_t total = (int32_t)(wcount[0] + wcount[1] + wcount[2] + wcount[3]);
    const uint32_t hi_mask = (1u << (W - kPkL)) - 1u;
    auto build = [&](uint32_t mh) {
        int4* s4 = reinterpret_cast<int4*>(s);
        const int bias = (int)kPkBias;
        for (int q = threadIdx.x; q < (pk_lds_words() + 3) / 4; q += blockDim.x)
            s4[q] = make_int4(bias, bias, bias, bias);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const uint32_t p = pat[k];
            if (p != 0u) {
                const uint32_t pl = p & ((1u << kPkL) - 1u);
                const uint32_t one = (pl >> kPkW) ? 0x10000u : 1u;
                const uint32_t neg = ((p >> 31) ^ __builtin_popcount((p >> kPkL) & hi_mask & mh)) & 1u;
                atomicAdd(&s[pk_pad((int)(pl & ((1u << kPkW) - 1u)))], neg ? 0u - one : one);
            }
        }
    };
    // Pass 1: the largest H of each block; the first block reaching the
    // overall maximum holds the lowest mask reaching it.
    __shared__ int32_t s_best_h;
    __shared__ uint32_t s_best_mh;
    int32_t best_h = INT32_MIN;   // thread 0's running values
    uint32_t best_mh = 0;
    for (uint32_t mh = 0; mh <= hi_mask; ++mh) {
        build(mh);
        s16x2 m = {INT16_MIN, INT16_MIN};
        pk_block_fwht_visit_paired(s, [&](int, const uint32_t (&v)[8]) {
#pragma unroll
            for (int e = 0; e < 8; ++e)
                m = pk_max(m, pk_as(v[e]));
        });
        const int32_t bm = wave_max_i32(max(pk_lo(m), pk_hi(m)));
        if (threadIdx.x % kWave == 0)
            wbest[threadIdx.x / kWave] = bm;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int32_t h = max(max(wbest[0], wbest[1]), max(wbest[2], wbest[3]));
            if (h > best_h) {
                best_h = h;
                best_mh = mh;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        s_best_h = best_h;
        s_best_mh = best_mh;
    }
    __syncthreads();
    best_h = s_best_h;
    const uint32_t mh = s_best_mh;
    // Pass 2: that block again, keyed H * 2^16 | (65535 - 4 m') per entry
    // (thread t's entries: ml = 2 t + m' with m' = j + 512 h + 1024 e, + 2^13
    // for the high half), so one max keeps the largest H and the lowest ml.
    build(mh);
    int32_t best = INT32_MIN;
    pk_block_fwht_visit_paired(s, [&](int jh, const uint32_t (&v)[8]) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t m_lo = (uint32_t)((jh & 1) + 512 * (jh >> 1) + 1024 * e);
            best = max(best, (int32_t)((v[e] << 16) | (65535u - 4u * m_lo)));
            best = max(best, (int32_t)((v[e] & 0xffff0000u) |
                                       (65535u - 4u * (m_lo + (1u << kPkW)))));
        }
    });
    {
        const int32_t h = best >> 16;
        const uint32_t ml = ((65535u - ((uint32_t)best & 0xffffu)) >> 2) + 2u * threadIdx.x;
        best = (int32_t)(((uint32_t)h << 16) | (65535u - 4u * ml));
    }
    best = wave_max_i32(best);
    if (threadIdx.x % kWave == 0)
        wbest[threadIdx.x / kWave] = best;
    __syncthreads();
    uint32_t best_mask = 0;
    if (threadIdx.x == 0) {
        const int32_t k = max(max(wbest[0], wbest[1]), max(wbest[2], wbest[3]));
        best_mask = (mh << kPkL) | ((65535u - ((uint32_t)k & 0xffffu)) >> 2);
    }
    if (threadIdx.x == 0) {
        const uint32_t w = (uint32_t)(total - best_h) >> 1;
        const bool better = w < roww[id];
        move[id] = better ? best_mask : 0u;
        neww[id] = w;
    }
}

__global__ void ga_win_apply_kernel(uint64_t* __restrict__ base, int rows, int words, int W,
                                    int bit, int phase, const int16_t* __restrict__ win,
                                    const uint32_t* __restrict__ move,
                                    const uint32_t* __restrict__ neww,
                                    uint32_t* __restrict__ roww,
                                    unsigned int* __restrict__ changed) {
    const int isl = blockIdx.x / rows;
    const int i = blockIdx.x % rows;
    if (((i >> bit) & 1) != phase)
        return;
    const size_t id = (size_t)isl * rows + i;
    const uint32_t m = move[id];
    if (m == 0u)
        return;
    uint64_t* b = base + (size_t)isl * rows * words;
    const int16_t* w = win + id * W;
    for (int x = threadIdx.x; x < words; x += blockDim.x) {
        uint64_t acc = b[(size_t)i * words + x];
        for (int k = 0; k < W; ++k)
            if ((m >> k) & 1u)
                acc ^= b[(size_t)w[k] * words + x];
        b[(size_t)i * words + x] = acc;
    }
    if (threadIdx.x == 0) {
        roww[id] = neww[id];
        atomicAdd(changed, 1u);
    }
}

// best_w[isl] = total weight of the island base (after local improvement).
// ------------------------------------------- information-set improvement --
// Random information set per island (Prange's step of information-set
// decoding): Gauss-Jordan elimination of the island base B on columns drawn
// at random, pivoting on the lowest not-yet-pivoted row that has the column,
// turns B into rows R = T B that each hold one pivot column alone.  Any
// sparse vector of the row space that meets the chosen columns once is one
// of these rows.  The rows of R (coefficient masks T) and the rows of B (unit
// masks) then form the pool of GlobalGreedyApplier
// (src/apply/global_greedy_applier.cpp:31-105): stably sorted by weight,
// each entry kept when its mask is independent of the kept ones, until a
// full basis is chosen.  The result is the lightest basis the pool holds, so
// the island's weight never grows.
//
// Elimination in two kernels.  The draws (column c = xorshift state mod
// cols) do not depend on the matrix, and row r's current bit in column c is
// parity(T_r & a(c)) with a(c) column c of B, so the pivoting runs on the
// 256-bit coefficient masks alone (ga_isd_pivot_kernel: one workgroup per
// island and trial gathers the columns of the next kIsdDraws draws as
// 256-bit vectors, then one wavefront runs the draws with every mask in
// registers).  R = T B is then
// one GF(2) product (ga_isd_rows_kernel), with the row weights.  Both apply
// exactly the row operations of elimination on the full rows.
constexpr int kIsdMaskWords = 4;   // rows <= 256
constexpr int kIsdDraws = 512;     // draws whose columns are gathered at once
// Trials of one round run side by side from the same base (grid y) and are
// selected from together.  A trial's eliminated rows depend only on the row
// space and its draws (they are the rows reduced on its pivot columns), and
// the lightest basis of a union is the lightest basis of one part's
// lightest basis and the rest (matroid greedy), so a round of trials gives a
// basis of the same weight as the trials one after another.
constexpr int kIsdBatch = 6;       // trials per round
constexpr int kIsdKeyBits = 11;    // pool index bits of a selection key
constexpr int kIsdPool = 1 << kIsdKeyBits;   // (kIsdBatch + 1) * 256 entries fit
static_assert((kIsdBatch + 1) * 256 <= kIsdPool, "selection pool too small");

// Seed of trial `round` (counted from 1): the same stream the trials have
// taken one after another.
__host__ __device__ inline uint64_t isd_seed(uint64_t seed, uint64_t round) {
    return seed ^ (0xA0761D6478BD642Full * round);
}

// The draw chain runs on one wavefront with every mask in registers: lane
// l holds T_r for the rows r = 64 s + l (s < 4).  The drawn columns are
// reduced against T 64 at a time by all four wavefronts and kept current
// through the draws by one rank-1 update per pivot, and the pivot row's mask
// is broadcast with readlane, so a draw takes no barrier, no LDS round trip
// and no parity work of its own (all 256 threads gather the drawn columns
// first).
// The base's columns as 256-bit vectors, colmaj[isl][c][q] bit i = row 64 q
// + i's bit in column c (one round of trials reads a drawn column as one
// 32-byte vector instead of a word from each row).  Grid: words x islands;
// thread r holds row r's word, each wavefront ballots its 64 rows per bit.
__global__ __launch_bounds__(256) void ga_isd_transpose_kernel(const uint64_t* __restrict__ base,
                                                               int rows, int words,
                                                               uint64_t* __restrict__ colmaj) {
    const int w = blockIdx.x, isl = blockIdx.y, tid = threadIdx.x;
    const int wave = tid / kWave, lane = tid % kWave;
    const uint64_t x =
        tid < rows ? base[((size_t)isl * rows + tid) * words + w] : 0ull;
    uint64_t* out = colmaj + ((size_t)isl * words * 64 + (size_t)w * 64) * kIsdMaskWords;
    uint64_t mine = 0;   // lane b keeps bit b's ballot: one store per lane
#pragma unroll 8
    for (int b = 0; b < 64; ++b) {
        const uint64_t bal = __ballot((x >> b) & 1ull);
        if (lane == b)
            mine = bal;
    }
    out[(size_t)lane * kIsdMaskWords + wave] = mine;
}

__device__ inline uint64_t readlane_u64(uint64_t x, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

__global__ __launch_bounds__(256) void ga_isd_pivot_kernel(const uint64_t* __restrict__ colmaj,
                                                           int rows, int words, int cols,
                                                           uint64_t seed0, uint64_t round0,
                                                           uint64_t* __restrict__ tmask_all) {
    const int isl = blockIdx.x, tid = threadIdx.x;
    const uint64_t seed = isd_seed(seed0, round0 + blockIdx.y);
    uint64_t* tmask = tmask_all + (size_t)blockIdx.y * gridDim.x * rows * kIsdMaskWords;
    // The wavefront index as a uniform value: the draw loop under
    // `wave == 0` then keeps its pivot search and counters in scalar registers.
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave), lane = tid % kWave;
    const uint64_t* A = colmaj + (size_t)isl * words * 64 * kIsdMaskWords;   // columns
    __shared__ uint64_t colv[kIsdDraws][kIsdMaskWords];   // a(c) of the chunk's draws
    __shared__ int cidx[kIsdDraws];
    __shared__ uint64_t state[kIsdDraws];
    __shared__ uint64_t tl[kIsdMaskWords * 64][kIsdMaskWords];   // T at a sub-batch start
    __shared__ uint64_t vl[kWave][kIsdMaskWords];   // the sub-batch's reduced columns
    __shared__ int s_npiv;
    // Wave 0's registers: t[s] = T of row 64 s + lane; unpiv = rows not yet
    // pivoted (bit l of word s: row 64 s + l; the same on every lane).
    uint64_t t[kIsdMaskWords][kIsdMaskWords] = {};
    uint64_t unpiv[kIsdMaskWords] = {};
#pragma unroll
    for (int sl = 0; sl < kIsdMaskWords; ++sl) {
        if (64 * sl + lane < rows)
            t[sl][sl] = 1ull << lane;
        const int n = min(64, max(0, rows - 64 * sl));
        unpiv[sl] = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
    }
    uint64_t rng = seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(isl + 1));
    if (rng == 0)
        rng = 1;
    int npiv = 0;
    const int max_draws = 8 * cols + 64;
    for (int d0 = 0; d0 < max_draws && npiv < rows; d0 += kIsdDraws) {
        const int nd = min(kIsdDraws, max_draws - d0);
        __syncthreads();   // the previous chunk's reads of cidx / colv are done
        // Thread 0 steps the generator; the (64-bit, software) modulo runs
        // on all threads.
        if (tid == 0)
            for (int d = 0; d < nd; ++d) {
                rng ^= rng << 13;
                rng ^= rng >> 7;
                rng ^= rng << 17;
                state[d] = rng;
            }
        __syncthreads();
        for (int d = tid; d < nd; d += blockDim.x)
            cidx[d] = (int)(state[d] % (uint64_t)cols);
        __syncthreads();
        // The drawn columns: one 32-byte vector each, all of a thread's
        // loads in flight at once (clamped indices, stores masked after).
        {
            constexpr int kPer = kIsdDraws * kIsdMaskWords / 256;
            const int last = nd * kIsdMaskWords - 1;
            uint64_t g[kPer];
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int i = min(tid + 256 * k, last);
                g[k] = A[(size_t)cidx[i / kIsdMaskWords] * kIsdMaskWords + i % kIsdMaskWords];
            }
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int i = tid + 256 * k;
                if (i <= last)
                    colv[i / kIsdMaskWords][i % kIsdMaskWords] = g[k];
            }
        }
        __syncthreads();
        // Sub-batches of 64 draws.  All four wavefronts reduce the
        // sub-batch's columns against T as it stands (lane l of wavefront w:
        // word w of v_l = T a(c_l), the parities of rows 64 w + i), then
        // wavefront 0 runs the draws in order: draw j pivots on the lowest
        // unpivoted row r of v_j, rows i with v_j bit i (i != r) take T_r,
        // and every later column follows: v_l bit i ^= v_j bit i & v_l bit r
        // (i != r), so v_l stays T a(c_l) for the current T.
        for (int e0 = 0; e0 < nd && npiv < rows; e0 += kWave) {
            if (wave == 0) {
#pragma unroll
                for (int sl = 0; sl < kIsdMaskWords; ++sl)
#pragma unroll
                    for (int q = 0; q < kIsdMaskWords; ++q)
                        tl[64 * sl + lane][q] = t[sl][q];
            }
            __syncthreads();
            {
                const int dd = min(e0 + lane, nd - 1);
                uint64_t a[kIsdMaskWords];
#pragma unroll
                for (int q = 0; q < kIsdMaskWords; ++q)
                    a[q] = colv[dd][q];
                uint32_t half[2] = {0u, 0u};
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int i = 0; i < 32; ++i) {
                        const uint64_t* tr = tl[64 * wave + 32 * h + i];
                        uint32_t p = 0;
#pragma unroll
                        for (int q = 0; q < kIsdMaskWords; ++q)
                            p += (uint32_t)__builtin_popcountll(tr[q] & a[q]);
                        half[h] |= (p & 1u) << i;
                    }
                vl[lane][wave] = ((uint64_t)half[1] << 32) | half[0];
            }
            __syncthreads();
            if (wave == 0) {
                uint64_t vr[kIsdMaskWords];   // lane l: v of draw e0 + l
#pragma unroll
                for (int q = 0; q < kIsdMaskWords; ++q)
                    vr[q] = vl[lane][q];
                const int ne = min(kWave, nd - e0);
                // Draws of the sub-batch not yet taken.  A draw whose column
                // has no bit in an unpivoted row never gains one (unpiv only
                // shrinks, and the column changes only when it holds the new
                // pivot's bit), so the loop jumps from one pivoting draw to
                // the next: the draws that would be skipped in order.
                uint64_t rem = ne >= kWave ? ~0ull : ((1ull << ne) - 1ull);
                while (npiv < rows) {
                    uint64_t any = 0;
#pragma unroll
                    for (int q = 0; q < kIsdMaskWords; ++q)
                        any |= vr[q] & unpiv[q];
                    const uint64_t live = (uint64_t)__ballot(any != 0ull) & rem;
                    if (live == 0ull)
                        break;
                    const int j = __builtin_ctzll(live);
                    rem &= j == kWave - 1 ? 0ull : ~((2ull << j) - 1ull);
                    // vj[s] bit l: row 64 s + l has the column's bit.
                    uint64_t vj[kIsdMaskWords];
#pragma unroll
                    for (int q = 0; q < kIsdMaskWords; ++q)
                        vj[q] = readlane_u64(vr[q], j);
                    int r = -1;   // lowest row not yet pivoted that has the bit
#pragma unroll
                    for (int sl = 0; sl < kIsdMaskWords; ++sl)
                        if (r < 0 && (vj[sl] & unpiv[sl]) != 0ull)
                            r = 64 * sl + __builtin_ctzll(vj[sl] & unpiv[sl]);
                    ++npiv;
                    const int rs = r / 64, rl = r % 64;
                    uint64_t tr[kIsdMaskWords];
#pragma unroll
                    for (int q = 0; q < kIsdMaskWords; ++q) {
                        uint64_t src = 0;
#pragma unroll
                        for (int sl = 0; sl < kIsdMaskWords; ++sl)
                            if (sl == rs)
                                src = t[sl][q];
                        tr[q] = readlane_u64(src, rl);
                    }
                    uint64_t vrs = 0;   // this lane's column: row r's bit
#pragma unroll
                    for (int sl = 0; sl < kIsdMaskWords; ++sl) {
                        if (sl == rs) {
                            unpiv[sl] &= ~(1ull << rl);
                            vrs = vr[sl];
                        }
                        // Every other row with the bit takes the pivot row.
                        if (((vj[sl] >> lane) & 1ull) && !(sl == rs && lane == rl)) {
#pragma unroll
                            for (int q = 0; q < kIsdMaskWords; ++q)
                                t[sl][q] ^= tr[q];
                        }
                    }
                    if ((vrs >> rl) & 1ull) {
#pragma unroll
                        for (int q = 0; q < kIsdMaskWords; ++q)
                            vr[q] ^= q == rs ? vj[q] & ~(1ull << rl) : vj[q];
                    }
                }
                if (lane == 0)
                    s_npiv = npiv;
            }
            __syncthreads();
            npiv = s_npiv;
        }
    }
    if (wave == 0) {
#pragma unroll
        for (int sl = 0; sl < kIsdMaskWords; ++sl) {
            const int r = 64 * sl + lane;
            if (r < rows) {
                uint64_t* T = tmask + ((size_t)isl * rows + r) * kIsdMaskWords;
#pragma unroll
                for (int q = 0; q < kIsdMaskWords; ++q)
                    T[q] = t[sl][q];
            }
        }
    }
}

// R = T B for one island and a slice of kIsdSlice words (grid: slices x
// islands x trials), by the method of four Russians: the base rows in
// groups of four, each group's 16 subset XORs of the slice built once in LDS
// (entries padded to 80 bytes, so the 16 entries of a group fall on
// different banks), then thread r XORs, for each group, the entry its mask
// nibble names (64 table reads instead of a masked XOR per base row),
// stores its slice of R and adds the slice's popcount to rw (zeroed before).
constexpr int kIsdSlice = 8;
constexpr int kIsdTabStride = kIsdSlice + 2;   // u64 per table entry

__global__ __launch_bounds__(256) void ga_isd_rows_kernel(const uint64_t* __restrict__ base,
                                                          int rows, int words,
                                                          const uint64_t* __restrict__ tmask_all,
                                                          uint64_t* __restrict__ rbuf_all,
                                                          uint32_t* __restrict__ rw_all) {
    const int slice = blockIdx.x, isl = blockIdx.y, tid = threadIdx.x;
    const size_t per = (size_t)rows * words;
    const size_t n_isl = gridDim.y;   // trial blockIdx.z's buffers follow the earlier trials'
    const uint64_t* tmask = tmask_all + (size_t)blockIdx.z * n_isl * rows * kIsdMaskWords;
    uint64_t* rbuf = rbuf_all + (size_t)blockIdx.z * n_isl * per;
    uint32_t* rw = rw_all + (size_t)blockIdx.z * n_isl * rows;
    const uint64_t* B = base + (size_t)isl * per;
    __shared__ __align__(16) uint64_t tab[64 * 16 * kIsdTabStride];
    const int w0 = slice * kIsdSlice;
    // Thread p builds word k of group g's 16 entries (p = 8 g + k, two
    // (g, k) pairs per thread): loads at clamped rows, masked after.
#pragma unroll
    for (int h = 0; h < 64 * kIsdSlice / 256; ++h) {
        const int p = tid + 256 * h, g = p / kIsdSlice, k = p % kIsdSlice;
        uint64_t x[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * g + r;
            const uint64_t v = B[(size_t)min(row, rows - 1) * words + w0 + k];
            x[r] = row < rows ? v : 0ull;
        }
        uint64_t e[16];
        e[0] = 0ull;
#pragma unroll
        for (int m = 1; m < 16; ++m)
            e[m] = e[m & (m - 1)] ^ x[__builtin_ctz(m)];
        uint64_t* T = tab + (size_t)g * 16 * kIsdTabStride + k;
#pragma unroll
        for (int m = 0; m < 16; ++m)
            T[m * kIsdTabStride] = e[m];
    }
    uint64_t t[kIsdMaskWords] = {};
    if (tid < rows) {
        const uint64_t* T = tmask + ((size_t)isl * rows + tid) * kIsdMaskWords;
#pragma unroll
        for (int q = 0; q < kIsdMaskWords; ++q)
            t[q] = T[q];
    }
    __syncthreads();
    uint64_t acc[kIsdSlice] = {};
#pragma unroll
    for (int q = 0; q < kIsdMaskWords; ++q) {   // groups 16 q .. 16 q + 15
        const uint64_t tq = t[q];
#pragma unroll 4
        for (int gg = 0; gg < 16; ++gg) {
            const int nib = (int)((tq >> (4 * gg)) & 15ull);
            const uint64_t* en = tab + (size_t)((16 * q + gg) * 16 + nib) * kIsdTabStride;
#pragma unroll
            for (int k = 0; k < kIsdSlice; ++k)
                acc[k] ^= en[k];
        }
    }
    if (tid < rows) {
        uint64_t* R = rbuf + (size_t)isl * per + (size_t)tid * words + w0;
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < kIsdSlice; ++k) {
            R[k] = acc[k];
            c += (uint32_t)__builtin_popcountll(acc[k]);
        }
        atomicAdd(&rw[(size_t)isl * rows + tid], c);
    }
}

// GlobalGreedyApplier over the pool {R rows of each trial (masks T)} + {B
// rows (unit masks)}, pool index = trial * rows + row, then rows + B row:
// bitonic sort of (weight, pool index) keys, then the entries in order
// against a fully reduced basis held in one wavefront's registers (reduce
// by the pivots set in the mask, pivot = highest remaining bit, clear that
// bit from the other basis rows; entries reduced 64 at a time by all
// wavefronts).  chosen[k] = pool index of the k-th kept entry.
__global__ __launch_bounds__(256) void ga_isd_select_kernel(int rows, int n_trials,
                                                            const uint64_t* __restrict__ tmask,
                                                            const uint32_t* __restrict__ rw,
                                                            const uint32_t* __restrict__ roww,
                                                            int16_t* __restrict__ chosen,
                                                            unsigned int* __restrict__ n_isd) {
    const int isl = blockIdx.x, tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave), lane = tid % kWave;
    const size_t n_isl = gridDim.x;
    __shared__ unsigned long long key[kIsdPool];
    __shared__ uint64_t tm[kIsdBatch * 256][kIsdMaskWords];
    const int n_r = n_trials * rows;   // pool entries from the trials
    const int pool = n_r + rows;
    int np2 = 512;
    while (np2 < pool)
        np2 <<= 1;
    for (int k = tid; k < np2; k += blockDim.x) {
        unsigned long long kk = ~0ull;
        if (k < pool) {
            const uint32_t w = k < n_r ? rw[((size_t)(k / rows) * n_isl + isl) * rows + k % rows]
                                       : roww[(size_t)isl * rows + (k - n_r)];
            kk = ((unsigned long long)w << kIsdKeyBits) | (unsigned)k;
        }
        key[k] = kk;
    }
    for (int k = tid; k < n_r; k += blockDim.x) {
#pragma unroll
        for (int q = 0; q < kIsdMaskWords; ++q)
            tm[k][q] = tmask[(((size_t)(k / rows) * n_isl + isl) * rows + k % rows) *
                                 kIsdMaskWords + q];
    }
    __syncthreads();
    // Bitonic sort of np2 keys (ascending).
    for (int size = 2; size <= np2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < np2 / 2; i += blockDim.x) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const unsigned long long a = key[lo], b = key[hi];
                if ((a > b) == up) {
                    key[lo] = b;
                    key[hi] = a;
                }
            }
            __syncthreads();
        }
    // Selection with the basis held by columns: basis row r is the row whose
    // pivot is bit r, and wavefront 0's lane l holds, for its bit positions
    // j = 64 s + l (s < 4), the 256-bit column col[s] (bit r = row r's bit
    // j).  The basis stays fully reduced (no row holds another row's pivot
    // bit), so an entry's mask m reduces to v = m ^ x with x_j =
    // parity(col_j & m & used).  Entries go in batches of 64: all four
    // wavefronts reduce the batch against the basis as it stands (the
    // columns staged in LDS; wavefront w forms word w of every entry's v),
    // then wavefront 0 takes the entries in order.  An entry with v = 0 is
    // dependent; otherwise v (pivot pv = its highest bit) is XORed into the
    // rows holding bit pv (column pv, broadcast from its lane) and installed
    // as row pv, and every later entry of the batch follows with
    // v_l ^= v_l bit pv ? v : 0, which is its reduction against the new
    // basis.
    __shared__ uint64_t cl[kIsdMaskWords * 64][kIsdMaskWords];   // basis columns
    __shared__ uint64_t s_used[kIsdMaskWords];
    __shared__ uint64_t vb[kWave][kIsdMaskWords];                // the batch's v
    __shared__ int s_count;
    uint64_t col[kIsdMaskWords][kIsdMaskWords] = {};   // [slot s][word of r] (wavefront 0)
    uint64_t used[kIsdMaskWords] = {};                  // pivot rows taken (uniform)
    int count = 0, kept_isd = 0;
    for (int e0 = 0; e0 < pool && count < rows; e0 += kWave) {
        if (wave == 0) {
#pragma unroll
            for (int sl = 0; sl < kIsdMaskWords; ++sl)
#pragma unroll
                for (int q = 0; q < kIsdMaskWords; ++q)
                    cl[64 * sl + lane][q] = col[sl][q];
            if (lane < kIsdMaskWords) {
                uint64_t u = 0;
#pragma unroll
                for (int q = 0; q < kIsdMaskWords; ++q)
                    if (q == lane)
                        u = used[q];
                s_used[lane] = u;
            }
        }
        __syncthreads();
        // Entry e0 + l of the sorted pool: pool index and mask.
        const int il = (int)(key[min(e0 + lane, pool - 1)] & (kIsdPool - 1u));
        {
            uint64_t mu[kIsdMaskWords], mw = 0;
#pragma unroll
            for (int q = 0; q < kIsdMaskWords; ++q) {
                const uint64_t m = il < n_r ? tm[il][q]
                                            : ((il - n_r) / 64 == q ? 1ull << ((il - n_r) % 64) : 0ull);
                mu[q] = m & s_used[q];
                if (q == wave)
                    mw = m;
            }
            uint32_t half[2] = {0u, 0u};
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int i = 0; i < 32; ++i) {
                    const uint64_t* cj = cl[64 * wave + 32 * h + i];
                    uint64_t a = 0;
#pragma unroll
                    for (int q = 0; q < kIsdMaskWords; ++q)
                        a ^= cj[q] & mu[q];
                    half[h] |= ((uint32_t)__builtin_popcountll(a) & 1u) << i;
                }
            vb[lane][wave] = mw ^ (((uint64_t)half[1] << 32) | half[0]);
        }
        __syncthreads();
        if (wave == 0) {
            uint64_t vr[kIsdMaskWords];   // lane l: v of entry e0 + l
#pragma unroll
            for (int q = 0; q < kIsdMaskWords; ++q)
                vr[q] = vb[lane][q];
            const int ne = min(kWave, pool - e0);
            // Entries of the batch not yet taken.  A dependent entry (v = 0)
            // stays dependent (the update adds v only where v_l has the new
            // pivot's bit), so the loop jumps from one kept entry to the next.
            uint64_t rem = ne >= kWave ? ~0ull : ((1ull << ne) - 1ull);
            while (count < rows) {
                const uint64_t live =
                    (uint64_t)__ballot((vr[0] | vr[1] | vr[2] | vr[3]) != 0ull) & rem;
                if (live == 0ull)
                    break;
                const int j = __builtin_ctzll(live);
                rem &= j == kWave - 1 ? 0ull : ~((2ull << j) - 1ull);
                uint64_t v[kIsdMaskWords];
#pragma unroll
                for (int q = 0; q < kIsdMaskWords; ++q)
                    v[q] = readlane_u64(vr[q], j);
                int pv = -1;   // highest set bit of v (the same on every lane)
#pragma unroll
                for (int q = kIsdMaskWords - 1; q >= 0; --q)
                    if (pv < 0 && v[q])
                        pv = q * 64 + 63 - __builtin_clzll(v[q]);
                const int idx = __builtin_amdgcn_readlane(il, j);
                const int ps = pv / 64, pl = pv % 64;
                // Column pv: the rows that hold bit pv.
                uint64_t cp[kIsdMaskWords];
#pragma unroll
                for (int q = 0; q < kIsdMaskWords; ++q) {
                    uint64_t src = 0;
#pragma unroll
                    for (int sl = 0; sl < kIsdMaskWords; ++sl)
                        if (sl == ps)
                            src = col[sl][q];
                    cp[q] = readlane_u64(src, pl);
                }
                // Rows holding bit pv take v (clearing their bit pv); row pv = v.
#pragma unroll
                for (int sl = 0; sl < kIsdMaskWords; ++sl)
                    if ((v[sl] >> lane) & 1ull) {
#pragma unroll
                        for (int q = 0; q < kIsdMaskWords; ++q) {
                            col[sl][q] ^= cp[q];
                            if (q == ps)
                                col[sl][q] |= 1ull << pl;
                        }
                    }
                uint64_t vps = 0;   // this lane's entry: bit pv
#pragma unroll
                for (int q = 0; q < kIsdMaskWords; ++q)
                    if (q == ps) {
                        used[q] |= 1ull << pl;
                        vps = vr[q];
                    }
                if ((vps >> pl) & 1ull) {
#pragma unroll
                    for (int q = 0; q < kIsdMaskWords; ++q)
                        vr[q] ^= v[q];
                }
                if (lane == 0)
                    chosen[(size_t)isl * rows + count] = (int16_t)idx;
                ++count;
                kept_isd += idx < n_r ? 1 : 0;
            }
            if (lane == 0)
                s_count = count;
        }
        __syncthreads();
        count = s_count;
    }
    if (wave == 0 && lane == 0 && kept_isd)
        atomicAdd(n_isd, (unsigned)kept_isd);
}

// New base: row k = row c % rows of trial c / rows's R (c = chosen[k] <
// n_trials * rows) or B row c - n_trials * rows, into `out`.
__global__ void ga_isd_apply_kernel(const uint64_t* __restrict__ base,
                                    const uint64_t* __restrict__ rbuf, int rows, int words,
                                    int n_trials, const int16_t* __restrict__ chosen,
                                    uint64_t* __restrict__ out) {
    const int isl = blockIdx.y, k = blockIdx.x;
    const int c = chosen[(size_t)isl * rows + k];
    const int n_r = n_trials * rows;
    const size_t per = (size_t)rows * words;
    const uint64_t* src = c < n_r ? rbuf + ((size_t)(c / rows) * gridDim.y + isl) * per +
                                        (size_t)(c % rows) * words
                                  : base + isl * per + (size_t)(c - n_r) * words;
    uint64_t* dst = out + isl * per + (size_t)k * words;
    for (int w = threadIdx.x; w < words; w += blockDim.x)
        dst[w] = src[w];
}

__global__ void ga_base_weight_kernel(const uint32_t* __restrict__ roww, int rows,
                                      unsigned int* __restrict__ best_w) {
    const int isl = blockIdx.x;
    uint32_t s = 0;
    for (int r = threadIdx.x; r < rows; r += blockDim.x)
        s += roww[(size_t)isl * rows + r];
    s = wave_sum_u32(s);
    __shared__ uint32_t red[4];
    if (threadIdx.x % kWave == 0)
        red[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < (int)(blockDim.x + kWave - 1) / kWave; ++k)
            t += red[k];
        best_w[isl] = t;
    }
}

// --------------------------------------------------------------- host ----
struct Ga {
    bmm_ga_params P{};
    int W32 = 0, words = 0;
    DevBuf base, base_tmp, ops[2], len[2], fit, rng, best_w, best_ops, best_len, cand, flag,
        w_snap, roww, sketch, evals, n_ops, base_total, eval_tier;
    // proposal model data
    DevBuf pairs, pair_isl, feats, labels, prng;
    DevBuf li_key, li_changed;
    DevBuf win_cand, win, win_move, win_w;
    DevBuf isd_r, isd_t, isd_w, isd_chosen, isd_count, isd_cols;
    DevBuf isl_state, mig;
    uint64_t win_round = 0;
    uint64_t win_moves = 0;
    uint64_t isd_round = 0;
    bool started = false;   // the first run has taken its initial step
    uint64_t isd_rows = 0;
    Mlp* mlp = nullptr;
    int cur = 0;
    long generation = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t gen_exec[2] = {nullptr, nullptr};
    hipGraph_t gen_graph[2] = {nullptr, nullptr};
    double train_loss = 0.0;
    uint64_t li_moves = 0;
    ~Ga() {
        for (int i = 0; i < 2; ++i) {
            if (gen_exec[i])
                (void)hipGraphExecDestroy(gen_exec[i]);
            if (gen_graph[i])
                (void)hipGraphDestroy(gen_graph[i]);
        }
        if (mlp)
            mlp_destroy(mlp);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace {

GaDev dev_params(const Ga& G) {
    GaDev d;
    d.islands = G.P.islands;
    d.pop = G.P.population;
    d.rows = G.P.rows;
    d.max_ops = G.P.max_ops;
    d.tournament = G.P.tournament;
    d.elite = G.P.elite;
    d.n_cand = G.P.learned ? G.P.n_candidates : 0;
    d.learned = G.P.learned;
    d.p_append = G.P.p_append;
    d.p_replace = G.P.p_replace;
    d.p_delete = G.P.p_delete;
    d.mutations = G.P.mutations;
    d.p_learned = G.P.p_learned;
    d.num_parents = G.P.num_parents;
    d.num_offspring = G.P.num_offspring;
    d.mutation_rate = G.P.mutation_rate;
    d.catastrophe_threshold = G.P.catastrophe_threshold;
    d.catastrophe_survival = G.P.catastrophe_survival;
    d.restart_after = G.P.restart_after;
    d.restart_p = G.P.restart_p;
    return d;
}

size_t breed_lds_bytes(const bmm_ga_params& P) {
    const size_t pop = (size_t)P.population;
    return pop * 4 + ((pop + 1) & ~(size_t)1) * 2 + pop * (size_t)P.num_parents * 3 * 2;
}

uint64_t xs_host(uint64_t& s) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}

void enqueue_generation(Ga& G, int parity, hipStream_t s) {
    const int I = G.P.islands, Pp = G.P.population;
    unsigned int* fit = static_cast<unsigned int*>(G.fit.ptr);
    const size_t n_ind = (size_t)I * Pp;
    hipLaunchKernelGGL(ga_fit_init_kernel, dim3((unsigned)((n_ind + 255) / 256)), dim3(256), 0, s,
                       fit, static_cast<const unsigned int*>(G.base_total.ptr), Pp, n_ind);
    const int slices = G.W32 / 64;
    const int n_tasks = I * Pp * slices;
    const size_t ops_words = (size_t)((G.P.max_ops + 3) & ~3);
    uint8_t* tier = static_cast<uint8_t*>(G.eval_tier.ptr);
    auto lds = [&](int slots) { return ((size_t)slots * 64 + ops_words) * 4; };
    const int s0 = std::min(G.P.rows, kGaEvalSlots0);
    hipLaunchKernelGGL(ga_eval_kernel, dim3((unsigned)n_tasks), dim3(64), lds(s0), s,
                       static_cast<const uint32_t*>(G.base.ptr),
                       static_cast<const uint16_t*>(G.ops[parity].ptr),
                       static_cast<const int32_t*>(G.len[parity].ptr), Pp, G.P.rows, G.W32,
                       G.P.max_ops, fit, s0, tier);
    BMM_HIP(hipGetLastError());
    // Later tiers on fixed grids, as many wavefronts per CU as their LDS
    // allows; a tier no larger than the previous one is not needed.
    const int cus = default_persistent_wgs() / 8;
    int prev = s0, level = 1;
    for (int k = 0; k < kGaEvalTiers && prev < G.P.rows; ++k) {
        const int slots = std::min(G.P.rows, kGaEvalSlots[k]);
        if (slots <= prev)
            continue;
        hipLaunchKernelGGL(ga_eval_tier_kernel,
                           dim3((unsigned)std::min(n_tasks, cus * kGaEvalPerCu[k])), dim3(64),
                           lds(slots), s, static_cast<const uint32_t*>(G.base.ptr),
                           static_cast<const uint16_t*>(G.ops[parity].ptr),
                           static_cast<const int32_t*>(G.len[parity].ptr), Pp, G.P.rows, G.W32,
                           G.P.max_ops, n_tasks, slots, level, fit, tier);
        BMM_HIP(hipGetLastError());
        prev = slots;
        ++level;
    }
    hipLaunchKernelGGL(ga_breed_kernel, dim3(I), dim3(std::max(64, (Pp + 63) / 64 * 64)),
                       breed_lds_bytes(G.P), s, dev_params(G), fit,
                       static_cast<const uint16_t*>(G.ops[parity].ptr),
                       static_cast<const int32_t*>(G.len[parity].ptr),
                       static_cast<uint16_t*>(G.ops[1 - parity].ptr),
                       static_cast<int32_t*>(G.len[1 - parity].ptr),
                       static_cast<uint64_t*>(G.rng.ptr), static_cast<unsigned int*>(G.best_w.ptr),
                       static_cast<uint16_t*>(G.best_ops.ptr),
                       static_cast<int32_t*>(G.best_len.ptr),
                       static_cast<const uint16_t*>(G.cand.ptr),
                       static_cast<unsigned long long*>(G.evals.ptr),
                       static_cast<unsigned long long*>(G.n_ops.ptr),
                       static_cast<GaIslandState*>(G.isl_state.ptr));
    BMM_HIP(hipGetLastError());
}

void run_generation(Ga& G) {
    const int parity = G.cur;
    if (G.P.use_graph) {
        if (!G.gen_exec[parity]) {
            hipStream_t cap;
            BMM_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            BMM_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
            enqueue_generation(G, parity, cap);
            BMM_HIP(hipStreamEndCapture(cap, &G.gen_graph[parity]));
            BMM_HIP(hipGraphInstantiate(&G.gen_exec[parity], G.gen_graph[parity], nullptr,
                                        nullptr, 0));
            (void)hipStreamDestroy(cap);
        }
        BMM_HIP(hipGraphLaunch(G.gen_exec[parity], G.stream));
    } else {
        enqueue_generation(G, parity, G.stream);
    }
    G.cur = 1 - parity;
    ++G.generation;
}

void refresh_sketches(Ga& G) {
    const int I = G.P.islands;
    hipLaunchKernelGGL(ga_sketch_kernel, dim3((unsigned)(I * G.P.rows)), dim3(128), 0, G.stream,
                       static_cast<const uint64_t*>(G.base.ptr), G.P.rows, G.words,
                       static_cast<uint32_t*>(G.roww.ptr), static_cast<uint64_t*>(G.sketch.ptr));
    BMM_HIP(hipGetLastError());
}

// Per island the base's total weight (from the row weights of the last
// refresh), the starting fitness of every evaluation.
void update_base_total(Ga& G) {
    hipLaunchKernelGGL(ga_base_weight_kernel, dim3(G.P.islands), dim3(256), 0, G.stream,
                       static_cast<const uint32_t*>(G.roww.ptr), G.P.rows,
                       static_cast<unsigned int*>(G.base_total.ptr));
    BMM_HIP(hipGetLastError());
}

// Samples n pairs per island into G.pairs / G.pair_isl.
void sample_pairs(Ga& G, int per_island) {
    const int n = per_island * G.P.islands;
    hipLaunchKernelGGL(ga_random_pairs_kernel, dim3((n + 255) / 256), dim3(256), 0, G.stream,
                       static_cast<uint64_t*>(G.prng.ptr), G.P.rows, per_island, G.P.islands,
                       static_cast<uint16_t*>(G.pairs.ptr), static_cast<int32_t*>(G.pair_isl.ptr));
    BMM_HIP(hipGetLastError());
}

void make_features(Ga& G, int n_padded, int n) {
    hipLaunchKernelGGL(ga_features_kernel, dim3((unsigned)n_padded), dim3(256), 0, G.stream,
                       static_cast<const uint64_t*>(G.sketch.ptr), G.P.rows,
                       static_cast<const uint16_t*>(G.pairs.ptr),
                       static_cast<const int32_t*>(G.pair_isl.ptr), n,
                       static_cast<__bf16*>(G.feats.ptr));
    BMM_HIP(hipGetLastError());
}

constexpr int kScorePerIsland = 1024;
constexpr int kTrainPerIsland = 64;
constexpr float kLabelScale = 64.f;

// Learned proposal: train on exact deltas of fresh random pairs, then rank a
// fresh pool per island and keep the lowest predicted deltas.
void update_proposals(Ga& G) {
    const int I = G.P.islands;
    refresh_sketches(G);
    update_base_total(G);
    if (!G.P.learned)
        return;
    // training batch
    const int nt = kTrainPerIsland * I;
    const int ntp = (nt + 31) / 32 * 32;
    sample_pairs(G, kTrainPerIsland);
    make_features(G, ntp, nt);
    hipLaunchKernelGGL(ga_labels_kernel, dim3((unsigned)nt), dim3(128), 0, G.stream,
                       static_cast<const uint64_t*>(G.base.ptr), G.P.rows, G.words,
                       static_cast<const uint16_t*>(G.pairs.ptr),
                       static_cast<const int32_t*>(G.pair_isl.ptr), nt, kLabelScale,
                       static_cast<float*>(G.labels.ptr));
    BMM_HIP(hipGetLastError());
    float* loss = nullptr;
    for (int it = 0; it < G.P.train_steps; ++it) {
        loss = mlp_backward_dev(G.mlp, static_cast<const __bf16*>(G.feats.ptr),
                                static_cast<const float*>(G.labels.ptr), ntp, nt, G.stream);
        mlp_adam_dev(G.mlp, G.P.learning_rate, G.stream);
    }
    if (loss) {
        float l = 0.f;
        BMM_HIP(hipMemcpyAsync(&l, loss, 4, hipMemcpyDeviceToHost, G.stream));
        BMM_HIP(hipStreamSynchronize(G.stream));
        G.train_loss = l;
    }
    // score a pool per island
    const int ns = kScorePerIsland * I;
    sample_pairs(G, kScorePerIsland);
    float* y = mlp_score_pairs_dev(G.mlp, static_cast<const uint64_t*>(G.sketch.ptr), G.P.rows,
                                   static_cast<const uint16_t*>(G.pairs.ptr),
                                   static_cast<const int32_t*>(G.pair_isl.ptr), ns, G.stream);
    hipLaunchKernelGGL(ga_select_kernel, dim3(I), dim3(256), 0, G.stream, y,
                       static_cast<const uint16_t*>(G.pairs.ptr), kScorePerIsland,
                       G.P.n_candidates, static_cast<uint16_t*>(G.cand.ptr));
    BMM_HIP(hipGetLastError());
}

// Rounds of parallel pairwise improvement on every island base; returns the
// number of applied moves.  Needs up-to-date row weights (ga_sketch_kernel).
unsigned int local_improvement(Ga& G) {
    const int I = G.P.islands, R = G.P.rows;
    int bits = 0;
    while ((1 << bits) < R)
        ++bits;
    unsigned int* changed = static_cast<unsigned int*>(G.li_changed.ptr);
    fill_u32(changed, (size_t)I, 0u, G.stream);
    for (int sweep = 0; sweep < G.P.li_sweeps; ++sweep)
        for (int bit = 0; bit < bits; ++bit)
            for (int phase = 0; phase < 2; ++phase) {
                // Up to 128 targets and 128 sources: 4 x 4 tiles of 32.
                hipLaunchKernelGGL(ga_li_tile_kernel, dim3(16, (unsigned)I), dim3(256), 0,
                                   G.stream, static_cast<const uint64_t*>(G.base.ptr), R, G.words,
                                   static_cast<const uint32_t*>(G.roww.ptr), bit, phase,
                                   static_cast<unsigned long long*>(G.li_key.ptr));
                int targets = 0;   // rows whose index bit `bit` equals `phase`
                for (int i = 0; i < R; ++i)
                    targets += ((i >> bit) & 1) == phase;
                hipLaunchKernelGGL(ga_li_apply_kernel,
                                   dim3((unsigned)((targets + kLiApplyWaves - 1) / kLiApplyWaves),
                                        (unsigned)I),
                                   dim3(kLiApplyWaves * kWave),
                                   0, G.stream, static_cast<uint64_t*>(G.base.ptr), R, G.words,
                                   static_cast<unsigned long long*>(G.li_key.ptr),
                                   static_cast<uint32_t*>(G.roww.ptr), bit, phase, changed);
            }
    hipLaunchKernelGGL(ga_base_weight_kernel, dim3(I), dim3(256), 0, G.stream,
                       static_cast<const uint32_t*>(G.roww.ptr), R,
                       static_cast<unsigned int*>(G.best_w.ptr));
    BMM_HIP(hipGetLastError());
    std::vector<unsigned int> per(I);
    BMM_HIP(hipMemcpyAsync(per.data(), changed, (size_t)I * 4, hipMemcpyDeviceToHost, G.stream));
    BMM_HIP(hipStreamSynchronize(G.stream));
    unsigned int n = 0;
    for (unsigned int c : per)
        n += c;
    return n;
}

// Rounds of window local improvement on every island base (windows chosen at
// random, or the learned proposal's top-W of C random candidates per
// target); returns the number of applied moves.
unsigned int window_improvement(Ga& G) {
    const int I = G.P.islands, R = G.P.rows, W = G.P.win_rows;
    const int C = G.P.learned ? G.P.win_candidates : W;
    int bits = 0;
    while ((1 << bits) < R)
        ++bits;
    unsigned int* changed = static_cast<unsigned int*>(G.li_changed.ptr);
    fill_u32(changed, 1, 0u, G.stream);
    int16_t* cand = static_cast<int16_t*>(G.win_cand.ptr);
    int16_t* win = static_cast<int16_t*>(G.win.ptr);
    const int n_rows = I * R;
    for (int round = 0; round < G.P.win_rounds; ++round) {
        const uint64_t r = G.win_round++;
        const int bit = (int)(r % (uint64_t)bits);
        const int phase = (int)((r / (uint64_t)bits) & 1u);
        refresh_sketches(G);
        hipLaunchKernelGGL(ga_win_pick_kernel, dim3((n_rows + 255) / 256), dim3(256), 0,
                           G.stream, I, R, bit, phase, C,
                           (uint64_t)G.P.seed ^ (0xD1B54A32D192ED03ull * (r + 1)), cand);
        const float* score = nullptr;
        if (G.P.learned) {
            const int n = n_rows * C;
            hipLaunchKernelGGL(ga_win_pairs_kernel, dim3((n + 255) / 256), dim3(256), 0,
                               G.stream, I, R, C, cand, bit, phase,
                               static_cast<uint16_t*>(G.pairs.ptr),
                               static_cast<int32_t*>(G.pair_isl.ptr));
            score = mlp_score_pairs_dev(G.mlp, static_cast<const uint64_t*>(G.sketch.ptr), R,
                                        static_cast<const uint16_t*>(G.pairs.ptr),
                                        static_cast<const int32_t*>(G.pair_isl.ptr), n,
                                        G.stream);
        }
        hipLaunchKernelGGL(ga_win_rank_kernel, dim3((n_rows + 255) / 256), dim3(256), 0,
                           G.stream, I, R, C, W, bit, phase, cand, score, win);
        hipLaunchKernelGGL(ga_win_kernel, dim3((unsigned)n_rows), dim3(256), 0, G.stream,
                           static_cast<const uint64_t*>(G.base.ptr), R, G.words, W, bit, phase,
                           win, static_cast<const uint32_t*>(G.roww.ptr),
                           static_cast<uint32_t*>(G.win_move.ptr),
                           static_cast<uint32_t*>(G.win_w.ptr));
        hipLaunchKernelGGL(ga_win_apply_kernel, dim3((unsigned)n_rows), dim3(128), 0, G.stream,
                           static_cast<uint64_t*>(G.base.ptr), R, G.words, W, bit, phase, win,
                           static_cast<const uint32_t*>(G.win_move.ptr),
                           static_cast<const uint32_t*>(G.win_w.ptr),
                           static_cast<uint32_t*>(G.roww.ptr), changed);
        BMM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(ga_base_weight_kernel, dim3(I), dim3(256), 0, G.stream,
                       static_cast<const uint32_t*>(G.roww.ptr), R,
                       static_cast<unsigned int*>(G.best_w.ptr));
    BMM_HIP(hipGetLastError());
    unsigned int n = 0;
    BMM_HIP(hipMemcpyAsync(&n, changed, 4, hipMemcpyDeviceToHost, G.stream));
    BMM_HIP(hipStreamSynchronize(G.stream));
    return n;
}

// Information-set trials on every island base (ga_isd_*_kernel); returns the
// rows taken from the eliminated bases.
unsigned int isd_improvement(Ga& G) {
    const int I = G.P.islands, R = G.P.rows;
    const size_t per = (size_t)R * G.words;
    unsigned int* cnt = static_cast<unsigned int*>(G.isd_count.ptr);
    fill_u32(cnt, 1, 0u, G.stream);
    refresh_sketches(G);
    // Rounds of up to kIsdBatch trials side by side (same base, one
    // selection over all their rows and the island's).
    for (int t0 = 0; t0 < G.P.isd_trials; t0 += kIsdBatch) {
        const int nb = std::min(kIsdBatch, G.P.isd_trials - t0);
        const uint64_t round0 = G.isd_round + 1;
        G.isd_round += (uint64_t)nb;
        hipLaunchKernelGGL(ga_isd_transpose_kernel, dim3(G.words, I), dim3(256), 0, G.stream,
                           static_cast<const uint64_t*>(G.base.ptr), R, G.words,
                           static_cast<uint64_t*>(G.isd_cols.ptr));
        hipLaunchKernelGGL(ga_isd_pivot_kernel, dim3(I, nb), dim3(256), 0, G.stream,
                           static_cast<const uint64_t*>(G.isd_cols.ptr), R, G.words,
                           G.words * 64, (uint64_t)G.P.seed, round0,
                           static_cast<uint64_t*>(G.isd_t.ptr));
        fill_u32(static_cast<uint32_t*>(G.isd_w.ptr), (size_t)nb * I * R, 0u, G.stream);
        hipLaunchKernelGGL(ga_isd_rows_kernel, dim3(G.words / kIsdSlice, I, nb), dim3(256), 0,
                           G.stream, static_cast<const uint64_t*>(G.base.ptr), R, G.words,
                           static_cast<const uint64_t*>(G.isd_t.ptr),
                           static_cast<uint64_t*>(G.isd_r.ptr), static_cast<uint32_t*>(G.isd_w.ptr));
        hipLaunchKernelGGL(ga_isd_select_kernel, dim3(I), dim3(256), 0, G.stream, R, nb,
                           static_cast<const uint64_t*>(G.isd_t.ptr),
                           static_cast<const uint32_t*>(G.isd_w.ptr),
                           static_cast<const uint32_t*>(G.roww.ptr),
                           static_cast<int16_t*>(G.isd_chosen.ptr), cnt);
        hipLaunchKernelGGL(ga_isd_apply_kernel, dim3(R, I), dim3(128), 0, G.stream,
                           static_cast<const uint64_t*>(G.base.ptr),
                           static_cast<const uint64_t*>(G.isd_r.ptr), R, G.words, nb,
                           static_cast<const int16_t*>(G.isd_chosen.ptr),
                           static_cast<uint64_t*>(G.base_tmp.ptr));
        BMM_HIP(hipGetLastError());
        BMM_HIP(hipMemcpyAsync(G.base.ptr, G.base_tmp.ptr, (size_t)I * per * 8,
                               hipMemcpyDeviceToDevice, G.stream));
        refresh_sketches(G);
    }
    hipLaunchKernelGGL(ga_base_weight_kernel, dim3(I), dim3(256), 0, G.stream,
                       static_cast<const uint32_t*>(G.roww.ptr), R,
                       static_cast<unsigned int*>(G.best_w.ptr));
    BMM_HIP(hipGetLastError());
    unsigned int n = 0;
    BMM_HIP(hipMemcpyAsync(&n, cnt, 4, hipMemcpyDeviceToHost, G.stream));
    BMM_HIP(hipStreamSynchronize(G.stream));
    return n;
}

void epoch(Ga& G) {
    const int I = G.P.islands;
    const int parity = G.cur;   // population about to be evaluated
    hipLaunchKernelGGL(ga_rebase_kernel, dim3((G.words + 63) / 64, I), dim3(64), 0, G.stream,
                       static_cast<uint64_t*>(G.base.ptr), G.P.rows, G.words,
                       static_cast<const uint16_t*>(G.best_ops.ptr),
                       static_cast<const int32_t*>(G.best_len.ptr), G.P.max_ops);
    hipLaunchKernelGGL(ga_after_rebase_kernel, dim3((I + 63) / 64), dim3(64), 0, G.stream,
                       static_cast<int32_t*>(G.best_len.ptr),
                       static_cast<int32_t*>(G.len[parity].ptr), G.P.population, I);
    const size_t per = (size_t)G.P.rows * G.words;
    hipLaunchKernelGGL(ga_migrate_decide_kernel, dim3((I + 63) / 64), dim3(64), 0, G.stream,
                       static_cast<const unsigned int*>(G.best_w.ptr), I,
                       static_cast<const int32_t*>(G.mig.ptr), static_cast<int32_t*>(G.flag.ptr),
                       static_cast<unsigned int*>(G.w_snap.ptr));
    hipLaunchKernelGGL(ga_migrate_copy_kernel, dim3(64, I), dim3(256), 0, G.stream,
                       static_cast<const uint64_t*>(G.base.ptr),
                       static_cast<uint64_t*>(G.base_tmp.ptr), per,
                       static_cast<const int32_t*>(G.flag.ptr), I, 1);
    hipLaunchKernelGGL(ga_migrate_copy_kernel, dim3(64, I), dim3(256), 0, G.stream,
                       static_cast<const uint64_t*>(G.base_tmp.ptr),
                       static_cast<uint64_t*>(G.base.ptr), per,
                       static_cast<const int32_t*>(G.flag.ptr), I, 0);
    hipLaunchKernelGGL(ga_migrate_weights_kernel, dim3((I + 63) / 64), dim3(64), 0, G.stream,
                       static_cast<unsigned int*>(G.best_w.ptr),
                       static_cast<const unsigned int*>(G.w_snap.ptr),
                       static_cast<const int32_t*>(G.flag.ptr), I);
    BMM_HIP(hipGetLastError());
    if (G.P.li_sweeps > 0) {
        refresh_sketches(G);
        G.li_moves += local_improvement(G);
    }
    if (G.P.win_rows > 0 && G.P.win_rounds > 0)
        G.win_moves += window_improvement(G);
    if (G.P.isd_trials > 0)
        G.isd_rows += isd_improvement(G);
    update_proposals(G);
}

// With init_rounds set, before the first generation the island bases go
// through the epoch's improvement steps (pairwise, window and
// information-set rounds; with none of them configured, nothing changes)
// and the proposals are refreshed on the new bases: the first epoch then
// evolves op lists on the
// improved bases instead of long ones on the input (on the planted
// 256 x 8192 instance, 1.01 x planted in a third of the time).
void initial_improvement(Ga& G) {
    bool ran = false;
    if (G.P.li_sweeps > 0) {
        refresh_sketches(G);
        G.li_moves += local_improvement(G);
        ran = true;
    }
    if (G.P.win_rows > 0 && G.P.win_rounds > 0) {
        G.win_moves += window_improvement(G);
        ran = true;
    }
    if (G.P.isd_trials > 0) {
        G.isd_rows += isd_improvement(G);
        ran = true;
    }
    if (ran)
        update_proposals(G);
}

} // namespace

Ga* ga_create(const uint64_t* matrix, size_t rows, size_t cols, const bmm_ga_params* params) {
    auto G = std::make_unique<Ga>();
    G->P = *params;
    bmm_ga_params& P = G->P;
    if (rows < 2 || rows > 256)
        fail(BMM_ERR_INVALID, "ga: rows must be in [2, 256]");
    if (cols == 0 || cols % 2048 != 0)
        fail(BMM_ERR_INVALID, "ga: cols must be a positive multiple of 2048");
    if (P.islands < 1 || P.population < 1 || P.population > 1024 || P.max_ops < 1 ||
        P.max_ops > 4096 || P.tournament < 1 || P.elite < 0 || P.elite > P.population ||
        P.migrate_every < 1 || P.mutations < 0)
        fail(BMM_ERR_INVALID, "ga: parameter out of range");
    if (P.num_parents < 1 || P.num_parents > kMaxParents || P.num_offspring < 1 ||
        P.num_offspring > P.population)
        fail(BMM_ERR_INVALID, "ga: num_parents must be in [1, 4] and num_offspring in "
                              "[1, population]");
    if (!(P.mutation_rate >= 0.f && P.mutation_rate <= 1000.f) || P.catastrophe_threshold < 0 ||
        !(P.catastrophe_survival >= 0.f && P.catastrophe_survival <= 1.f) ||
        P.restart_after < 0 || !(P.restart_p >= 0.f && P.restart_p <= 1.f))
        fail(BMM_ERR_INVALID, "ga: mutation_rate must be in [0, 1000], catastrophe_survival and "
                              "restart_p in [0, 1], catastrophe_threshold and restart_after >= 0");
    if (P.learned && (P.n_candidates < 1 || P.n_candidates > kScorePerIsland))
        fail(BMM_ERR_INVALID, "ga: n_candidates must be in [1, 1024] for the learned proposal");
    if (P.win_rows != 0) {
        if (P.win_rows < kPkL || P.win_rows > 24)
            fail(BMM_ERR_INVALID, "ga: win_rows must be 0 or in [14, 24]");
        if (P.win_rounds < 0)
            fail(BMM_ERR_INVALID, "ga: win_rounds must be >= 0");
        if (P.learned && (P.win_candidates < P.win_rows || P.win_candidates > 128))
            fail(BMM_ERR_INVALID, "ga: win_candidates must be in [win_rows, 128]");
        if (cols > 8192)
            P.win_rows = 0;   // one thread holds cols / 256 <= 32 column patterns
    }
    if (P.isd_trials < 0 || P.isd_trials > 64)
        fail(BMM_ERR_INVALID, "ga: isd_trials must be in [0, 64]");
    if (P.init_rounds < 0 || P.init_rounds > 1)
        fail(BMM_ERR_INVALID, "ga: init_rounds must be 0 or 1");
    P.rows = (int)rows;
    G->words = (int)(cols / 64);
    G->W32 = (int)(cols / 32);
    const int I = P.islands, Pp = P.population;
    const size_t per = rows * (size_t)G->words;
    BMM_HIP(hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking));
    uint64_t* base = G->base.get<uint64_t>(per * I);
    G->base_tmp.get<uint64_t>(per * I);
    for (int i = 0; i < I; ++i)
        BMM_HIP(hipMemcpy(base + per * i, matrix, per * 8, hipMemcpyHostToDevice));
    for (int b = 0; b < 2; ++b) {
        G->ops[b].get<uint16_t>((size_t)I * Pp * P.max_ops);
        G->len[b].get<int32_t>((size_t)I * Pp);
        BMM_HIP(hipMemset(G->len[b].ptr, 0, (size_t)I * Pp * 4));
    }
    G->fit.get<unsigned int>((size_t)I * Pp);
    // One tier flag per fitness task (written by the first tier).
    G->eval_tier.get<uint8_t>((size_t)I * Pp * (size_t)(G->W32 / 64));
    G->best_ops.get<uint16_t>((size_t)I * P.max_ops);
    int32_t* bl = G->best_len.get<int32_t>(I);
    BMM_HIP(hipMemset(bl, 0, (size_t)I * 4));
    G->flag.get<int32_t>(I);
    G->w_snap.get<unsigned int>(I);
    G->roww.get<uint32_t>((size_t)I * rows);
    G->base_total.get<unsigned int>((size_t)I);
    G->sketch.get<uint64_t>((size_t)I * rows * 2);
    G->cand.get<uint16_t>((size_t)I * std::max(1, P.n_candidates));
    unsigned long long* ev = G->evals.get<unsigned long long>(1);
    BMM_HIP(hipMemset(ev, 0, 8));
    BMM_HIP(hipMemset(G->n_ops.get<unsigned long long>(1), 0, 8));
    // base weight as the initial best
    uint64_t w0 = 0;
    for (size_t i = 0; i < per; ++i)
        w0 += (uint64_t)__builtin_popcountll(matrix[i]);
    if (w0 >= 0xffffffffull)
        fail(BMM_ERR_OVERFLOW, "ga: matrix weight exceeds 32 bits");
    std::vector<unsigned int> bw(I, (unsigned int)w0);
    BMM_HIP(hipMemcpy(G->best_w.get<unsigned int>(I), bw.data(), (size_t)I * 4,
                      hipMemcpyHostToDevice));
    // RNG streams: one per individual, one per proposal slot.
    const int n_pairs = std::max(kScorePerIsland, kTrainPerIsland) * I;
    std::vector<uint64_t> seeds((size_t)I * Pp + n_pairs);
    uint64_t s = P.seed == 0 ? 0x9E3779B97F4A7C15ull : P.seed;
    for (auto& x : seeds)
        x = xs_host(s) | 1ull;
    BMM_HIP(hipMemcpy(G->rng.get<uint64_t>((size_t)I * Pp), seeds.data(), (size_t)I * Pp * 8,
                      hipMemcpyHostToDevice));
    BMM_HIP(hipMemcpy(G->prng.get<uint64_t>(n_pairs), seeds.data() + (size_t)I * Pp,
                      (size_t)n_pairs * 8, hipMemcpyHostToDevice));
    std::vector<GaIslandState> st0(I, GaIslandState{0u, 0u, P.mutation_rate, 0u});
    BMM_HIP(hipMemcpy(G->isl_state.get<GaIslandState>(I), st0.data(), (size_t)I * sizeof(GaIslandState),
                      hipMemcpyHostToDevice));
    std::vector<int32_t> mig0(I, 3);
    BMM_HIP(hipMemcpy(G->mig.get<int32_t>(I), mig0.data(), (size_t)I * 4, hipMemcpyHostToDevice));
    // Pairwise-improvement keys, ~0 between rounds (the apply kernel resets them).
    fill_u64(reinterpret_cast<uint64_t*>(G->li_key.get<unsigned long long>((size_t)I * rows)),
             (size_t)I * rows, ~0ull, G->stream);
    G->li_changed.get<unsigned int>((size_t)P.islands);   // per-island move counts
    const size_t n_win = P.win_rows && P.learned ? (size_t)I * rows * P.win_candidates : 0;
    const size_t n_pair_buf = std::max((size_t)n_pairs, n_win);
    G->pairs.get<uint16_t>(n_pair_buf + 32);
    G->pair_isl.get<int32_t>(n_pair_buf + 32);
    G->feats.get<__bf16>((size_t)(kTrainPerIsland * I + 32) * 256);   // training batch
    G->labels.get<float>(n_pairs + 32);
    if (P.win_rows) {
        const int C = P.learned ? P.win_candidates : P.win_rows;
        G->win_cand.get<int16_t>((size_t)I * rows * C);
        G->win.get<int16_t>((size_t)I * rows * P.win_rows);
        G->win_move.get<uint32_t>((size_t)I * rows);
        G->win_w.get<uint32_t>((size_t)I * rows);
    }
    if (P.isd_trials) {
        const size_t nb = (size_t)std::min(kIsdBatch, P.isd_trials);
        G->isd_r.get<uint64_t>(nb * I * rows * G->words);
        G->isd_t.get<uint64_t>(nb * I * rows * kIsdMaskWords);
        G->isd_w.get<uint32_t>(nb * I * rows);
        G->isd_chosen.get<int16_t>((size_t)I * rows);
        G->isd_count.get<uint32_t>(1);
        G->isd_cols.get<uint64_t>((size_t)I * G->words * 64 * kIsdMaskWords);
    }
    if (P.learned)
        G->mlp = mlp_create(P.seed ^ 0xA5A5A5A5ull);
    update_proposals(*G);
    BMM_HIP(hipStreamSynchronize(G->stream));
    return G.release();
}

void ga_destroy(Ga* G) { delete G; }

void ga_run(Ga* G, size_t max_generations, uint64_t target, double time_budget_s,
            bmm_ga_status* st) {
    const auto t0 = std::chrono::steady_clock::now();
    const int I = G->P.islands;
    std::vector<unsigned int> bw(I);
    size_t done = 0;
    bool reached = false;
    if (!G->started) {
        if (G->P.init_rounds)
            initial_improvement(*G);
        G->started = true;
    }
    while (done < max_generations) {
        run_generation(*G);
        ++done;
        if (G->generation % G->P.migrate_every == 0)
            epoch(*G);
        if (done % (size_t)G->P.check_every == 0 || done == max_generations) {
            BMM_HIP(hipMemcpyAsync(bw.data(), G->best_w.ptr, (size_t)I * 4,
                                   hipMemcpyDeviceToHost, G->stream));
            BMM_HIP(hipStreamSynchronize(G->stream));
            const unsigned int b = *std::min_element(bw.begin(), bw.end());
            const double el =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (b <= target) {
                reached = true;
                break;
            }
            if (time_budget_s > 0 && el >= time_budget_s)
                break;
        }
    }
    BMM_HIP(hipMemcpyAsync(bw.data(), G->best_w.ptr, (size_t)I * 4, hipMemcpyDeviceToHost,
                           G->stream));
    unsigned long long ev = 0, nops = 0;
    BMM_HIP(hipMemcpyAsync(&ev, G->evals.ptr, 8, hipMemcpyDeviceToHost, G->stream));
    BMM_HIP(hipMemcpyAsync(&nops, G->n_ops.ptr, 8, hipMemcpyDeviceToHost, G->stream));
    BMM_HIP(hipStreamSynchronize(G->stream));
    st->generations = (size_t)G->generation;
    st->ran = done;
    st->evaluations = ev;
    st->best_weight = *std::min_element(bw.begin(), bw.end());
    st->best_island = (int)(std::min_element(bw.begin(), bw.end()) - bw.begin());
    st->reached_target = reached ? 1 : 0;
    st->elapsed_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st->train_loss = G->train_loss;
    st->li_moves = G->li_moves;
    st->win_moves = G->win_moves;
    st->isd_rows = G->isd_rows;
    st->replayed_ops = nops;
}

// Best matrix: the best island's base with its best sequence replayed.
// Island `island`'s best individual (its base with its best op sequence
// replayed); island < 0: the island with the lightest best.
void ga_island_best(Ga* G, int island, uint64_t* out, uint64_t* weight) {
    const int I = G->P.islands;
    if (island >= I)
        fail(BMM_ERR_RANGE, "ga_island_best: island out of range");
    std::vector<unsigned int> bw(I);
    BMM_HIP(hipMemcpy(bw.data(), G->best_w.ptr, (size_t)I * 4, hipMemcpyDeviceToHost));
    const int isl =
        island >= 0 ? island : (int)(std::min_element(bw.begin(), bw.end()) - bw.begin());
    const size_t per = (size_t)G->P.rows * G->words;
    std::vector<uint64_t> m(per);
    BMM_HIP(hipMemcpy(m.data(), static_cast<uint64_t*>(G->base.ptr) + per * isl, per * 8,
                      hipMemcpyDeviceToHost));
    std::vector<int32_t> bl(I);
    BMM_HIP(hipMemcpy(bl.data(), G->best_len.ptr, (size_t)I * 4, hipMemcpyDeviceToHost));
    std::vector<uint16_t> ops((size_t)G->P.max_ops);
    BMM_HIP(hipMemcpy(ops.data(), static_cast<uint16_t*>(G->best_ops.ptr) + (size_t)isl * G->P.max_ops,
                      ops.size() * 2, hipMemcpyDeviceToHost));
    // Replay on the GPU through the same replay path as the appliers.
    std::vector<uint16_t> pairs;
    for (int k = 0; k < bl[isl]; ++k) {
        pairs.push_back((uint16_t)(ops[k] & 0xff));
        pairs.push_back((uint16_t)(ops[k] >> 8));
    }
    uint64_t w = 0;
    replay_host(m.data(), (size_t)G->P.rows, (size_t)G->words, pairs.data(), pairs.size() / 2, &w);
    std::memcpy(out, m.data(), per * 8);
    *weight = w;
}

void ga_best(Ga* G, uint64_t* out, uint64_t* weight) { ga_island_best(G, -1, out, weight); }

// Island `island`'s base matrix and best op sequence: replaying the ops on
// the base gives ga_island_best's matrix.
void ga_best_ops(Ga* G, int island, uint64_t* base, uint16_t* ops, size_t* n_ops) {
    if (island < 0 || island >= G->P.islands)
        fail(BMM_ERR_RANGE, "ga_best_ops: island out of range");
    BMM_HIP(hipStreamSynchronize(G->stream));
    const size_t per = (size_t)G->P.rows * G->words;
    BMM_HIP(hipMemcpy(base, static_cast<uint64_t*>(G->base.ptr) + per * island, per * 8,
                      hipMemcpyDeviceToHost));
    int32_t n = 0;
    BMM_HIP(hipMemcpy(&n, static_cast<int32_t*>(G->best_len.ptr) + island, 4,
                      hipMemcpyDeviceToHost));
    if (n > 0)
        BMM_HIP(hipMemcpy(ops, static_cast<uint16_t*>(G->best_ops.ptr) + (size_t)island * G->P.max_ops,
                          (size_t)n * 2, hipMemcpyDeviceToHost));
    *n_ops = (size_t)n;
}

// Populations and counters of one island (see bmm_hip_ga_population).
void ga_population(Ga* G, int island, uint16_t* ops_next, int32_t* len_next, uint16_t* ops_prev,
                   int32_t* len_prev, uint32_t* fit_prev, uint64_t* rng, bmm_ga_island_state* st) {
    if (island < 0 || island >= G->P.islands)
        fail(BMM_ERR_RANGE, "ga_population: island out of range");
    BMM_HIP(hipStreamSynchronize(G->stream));
    const size_t Pp = (size_t)G->P.population, off = (size_t)island * Pp;
    const size_t mo = (size_t)G->P.max_ops;
    const int nx = G->cur, pv = 1 - G->cur;
    if (ops_next)
        BMM_HIP(hipMemcpy(ops_next, static_cast<uint16_t*>(G->ops[nx].ptr) + off * mo, Pp * mo * 2,
                          hipMemcpyDeviceToHost));
    if (len_next)
        BMM_HIP(hipMemcpy(len_next, static_cast<int32_t*>(G->len[nx].ptr) + off, Pp * 4,
                          hipMemcpyDeviceToHost));
    if (ops_prev)
        BMM_HIP(hipMemcpy(ops_prev, static_cast<uint16_t*>(G->ops[pv].ptr) + off * mo, Pp * mo * 2,
                          hipMemcpyDeviceToHost));
    if (len_prev)
        BMM_HIP(hipMemcpy(len_prev, static_cast<int32_t*>(G->len[pv].ptr) + off, Pp * 4,
                          hipMemcpyDeviceToHost));
    if (fit_prev)
        BMM_HIP(hipMemcpy(fit_prev, static_cast<unsigned int*>(G->fit.ptr) + off, Pp * 4,
                          hipMemcpyDeviceToHost));
    if (rng)
        BMM_HIP(hipMemcpy(rng, static_cast<uint64_t*>(G->rng.ptr) + off, Pp * 8,
                          hipMemcpyDeviceToHost));
    if (st) {
        GaIslandState s{};
        BMM_HIP(hipMemcpy(&s, static_cast<GaIslandState*>(G->isl_state.ptr) + island, sizeof(s),
                          hipMemcpyDeviceToHost));
        BMM_HIP(hipMemcpy(&st->best_weight, static_cast<unsigned int*>(G->best_w.ptr) + island, 4,
                          hipMemcpyDeviceToHost));
        BMM_HIP(hipMemcpy(&st->best_len, static_cast<int32_t*>(G->best_len.ptr) + island, 4,
                          hipMemcpyDeviceToHost));
        BMM_HIP(hipMemcpy(&st->migrate, static_cast<int32_t*>(G->mig.ptr) + island, 4,
                          hipMemcpyDeviceToHost));
        st->no_improvement = s.no_improvement;
        st->catastrophe = s.catastrophe;
        st->mutation_rate = s.rate;
    }
}

// The island's ranked candidate pairs (learned proposal; empty otherwise).
void ga_proposals(Ga* G, int island, uint16_t* out, size_t cap, size_t* n) {
    if (island < 0 || island >= G->P.islands)
        fail(BMM_ERR_RANGE, "ga_proposals: island out of range");
    const size_t k = G->P.learned ? (size_t)G->P.n_candidates : 0;
    if (k > cap)
        fail(BMM_ERR_INVALID, "ga_proposals: output capacity below n_candidates");
    BMM_HIP(hipStreamSynchronize(G->stream));
    if (k)
        BMM_HIP(hipMemcpy(out, static_cast<uint16_t*>(G->cand.ptr) + (size_t)island * k, k * 2,
                          hipMemcpyDeviceToHost));
    *n = k;
}

// Replaces island `island`'s base by `matrix` (rows x cols/64 words); the
// island's best becomes that base (weight = its popcount, empty sequence).
void ga_set_island_base(Ga* G, int island, const uint64_t* matrix) {
    if (island < 0 || island >= G->P.islands)
        fail(BMM_ERR_RANGE, "ga_set_island_base: island out of range");
    const size_t per = (size_t)G->P.rows * G->words;
    uint64_t w = 0;
    for (size_t i = 0; i < per; ++i)
        w += (uint64_t)__builtin_popcountll(matrix[i]);
    if (w >= 0xffffffffull)
        fail(BMM_ERR_OVERFLOW, "ga: matrix weight exceeds 32 bits");
    BMM_HIP(hipStreamSynchronize(G->stream));
    BMM_HIP(hipMemcpy(static_cast<uint64_t*>(G->base.ptr) + per * island, matrix, per * 8,
                      hipMemcpyHostToDevice));
    const unsigned int w32 = (unsigned int)w;
    const int32_t zero = 0;
    BMM_HIP(hipMemcpy(static_cast<unsigned int*>(G->best_w.ptr) + island, &w32, 4,
                      hipMemcpyHostToDevice));
    BMM_HIP(hipMemcpy(static_cast<int32_t*>(G->best_len.ptr) + island, &zero, 4,
                      hipMemcpyHostToDevice));
    refresh_sketches(*G);
    update_base_total(*G);
    BMM_HIP(hipStreamSynchronize(G->stream));
}

// Per-island seeds (individual streams of island i restarted from seeds[i],
// as the reference seeds each island's GA, bmmpy/ga.py:373) and migration
// flags (bit 0 export, bit 1 import).  Either array may be null.
void ga_configure_islands(Ga* G, const uint64_t* seeds, const int32_t* migrate, size_t n) {
    const int I = G->P.islands;
    if (n != (size_t)I)
        fail(BMM_ERR_INVALID, "ga_configure_islands: one entry per island expected");
    BMM_HIP(hipStreamSynchronize(G->stream));
    const size_t Pp = (size_t)G->P.population;
    if (seeds) {
        std::vector<uint64_t> st(Pp * I);
        for (int i = 0; i < I; ++i) {
            uint64_t s = seeds[i] == 0 ? 0x9E3779B97F4A7C15ull : seeds[i];
            for (size_t k = 0; k < Pp; ++k)
                st[i * Pp + k] = xs_host(s) | 1ull;
        }
        BMM_HIP(hipMemcpy(G->rng.ptr, st.data(), st.size() * 8, hipMemcpyHostToDevice));
    }
    if (migrate) {
        for (int i = 0; i < I; ++i)
            if (migrate[i] < 0 || migrate[i] > 3)
                fail(BMM_ERR_INVALID, "ga_configure_islands: migrate flags must be in [0, 3]");
        BMM_HIP(hipMemcpy(G->mig.ptr, migrate, (size_t)I * 4, hipMemcpyHostToDevice));
    }
}

} // namespace bmmhip
