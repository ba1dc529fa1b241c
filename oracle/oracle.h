/*
 * Test infrastructure only: a plain-C restatement of the reference's CPU
 * algorithms for the row-XOR weight-minimization path.  It is the checker the
 * tests compare the HIP path against, and the timed CPU baseline in bench.py.
 * No product code links or calls it.
 *
 * Conventions shared by every function:
 *   - A matrix is `n_rows` rows of `stride` uint64 words; bit c of a row is
 *     bit (c % 64) of word (c / 64) (LSB-first), words_per_row = ceil(cols/64).
 *   - A window is an ordered list of `t` row indices `idx`; local row r is
 *     mask bit r.
 *   - Results are ordered ascending by (weight, mask); mask 0 is never
 *     reported by the top-k searches.
 *
 * Pinning status per function is listed in oracle.c's header.
 */
#ifndef BMMPY_ORACLE_H
#define BMMPY_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* XorShift64 (include/bmmpy/core/detail/xorshift64.hpp:10-40). */
void or_xs_seed(uint64_t *state, uint64_t seed);
uint64_t or_xs_next(uint64_t *state);
uint64_t or_xs_index(uint64_t *state, uint64_t upper);
double or_xs_unit(uint64_t *state);

/* Synthetic inputs (fixed by the workload definitions, not by the reference). */
void or_fill_bernoulli(uint64_t *state, uint64_t *rows, int n_rows, int stride, int cols,
                       double p);
int or_rank(const uint64_t *rows, int n_rows, int stride, int cols);
void or_mul(const uint64_t *a, int a_rows, int a_stride, int a_cols, const uint64_t *b,
            int b_stride, int b_cols, uint64_t *out);
int or_random_invertible(uint64_t *state, uint64_t *m, int n, int stride);
int or_gen_planted(uint64_t seed, int n_rows, int cols, double p, uint64_t *sparse,
                   uint64_t *mixed, uint64_t *transform);
long long or_gen_batch(uint64_t seed, int n_mat, int *rows_out, int *cols_out, double *dens_out,
                       uint64_t *words, size_t words_cap, size_t *word_offsets);

/* FWHT, natural order, unnormalised, wrapping (src/math/fwht.cpp:28-57). */
int or_fwht_i32(int32_t *data, size_t n);
int or_fwht_i16(int16_t *data, size_t n);

/* Column-pattern histogram (src/search/fwht_search.cpp:53-91). */
int or_histogram(const uint64_t *base, int stride, int cols, const int *idx, int t,
                 int32_t *buckets, int64_t *total);

/* FwhtSearch top-k (src/search/fwht_search.cpp:34-132). */
long long or_fwht_topk(const uint64_t *base, int stride, int cols, const int *idx, int t,
                       size_t k, uint64_t *out_mask, uint32_t *out_w);

/* build_compact_split_window (src/search/split_window_prep.cpp:42-137). */
long long or_split_window(const uint64_t *base, int stride, int cols, const int *idx, int t,
                          int low_bits, int32_t *total, uint64_t *q, uint64_t *r, int32_t *mult,
                          size_t cap);

/* MitmFwhtSearch top-k (src/search/mitm_fwht_search.cpp:267-322). */
long long or_mitm_topk(const uint64_t *base, int stride, int cols, const int *idx, int t,
                       size_t k, uint64_t *out_mask, uint32_t *out_w);

/* BruteforceSearch top-k (src/search/bruteforce_search.cpp:113-330). */
long long or_bruteforce_topk(const uint64_t *base, int stride, int cols, const int *idx, int t,
                             size_t k, int chunk_bits, uint64_t *out_mask, uint32_t *out_w);

/* Per-target coset minimum: for each target row i of an n_rows matrix, the
 * minimum over subsets S of the other rows of wt(row_i ^ XOR_S) and the lowest
 * index of S (bit j = j-th other row in ascending order). */
int or_coset_min(const uint64_t *base, int stride, int cols, int n_rows, uint32_t *out_w,
                 uint64_t *out_arg);

/* Gray-code coset sweep: target row `target` against block rows
 * [0, block_rows) \ {target}; reports min weight, first Gray step s reaching it
 * and the lowest block mask reaching it. */
int or_gray_min(const uint64_t *base, int stride, int words, int block_rows, int target,
                uint32_t *out_w, uint64_t *out_first_s, uint64_t *out_low_mask);

/* Row-XOR replay (RowWindow::row_xor, src/core/bit_matrix_rows.cpp:12-20):
 * ops[2i] = source, ops[2i+1] = destination; returns the total weight. */
long long or_replay(uint64_t *rows, int n_rows, int stride, int cols, const uint16_t *ops,
                    size_t n_ops);

/* GreedyApplier::apply (src/apply/greedy_applier.cpp:13-109). */
int or_greedy_apply(uint64_t *rows, int n_rows, int stride, int cols, const uint64_t *masks,
                    const uint32_t *weights, size_t n_cand, uint64_t min_gain, int stochastic,
                    uint64_t seed, uint64_t *out_applied, uint64_t *out_improvement);

/* ExactBasisSolver::solve (src/solver/exact_basis_solver.cpp:29-157): masks
 * gray(i), i = 1 .. 2^t - 1, counting-sorted by weight (stable), inserted
 * greedily into a GF(2) pivot basis until the rank is reached.  Writes the
 * chosen masks / weights (selection order, up to cap) and the total; returns
 * the rank (t <= 24). */
int or_exact_basis(const uint64_t *base, int stride, int cols, const int *idx, int t,
                   uint32_t *out_masks, uint32_t *out_w, size_t cap, uint64_t *out_total);

/* The first min(k, 2^t - 1) masks of that order (the GPU candidate stream). */
long long or_weight_order(const uint64_t *base, int stride, int cols, const int *idx, int t,
                          size_t k, uint32_t *out_masks, uint32_t *out_w);

/* calc_scores_and_order (src/math/fwht.cpp:59-100); cnt and off hold n + 1
 * entries, order m - 1.  Returns 0, -1 for n < 0, -2 (int16 form) for a
 * score above 32767. */
int or_scores_order_i32(const int32_t *h, size_t m, int32_t n, int32_t *s_by_mask,
                        int32_t *order, int32_t *cnt, int32_t *off);
int or_scores_order_i16(const int16_t *h, size_t m, int32_t n, int16_t *s_by_mask,
                        int32_t *order, int32_t *cnt, int32_t *off);

/* GlobalGreedyApplier::apply (src/apply/global_greedy_applier.cpp:15-105)
 * on t rows (stride words) with n_cand candidates of mask_words words each.
 * Rows replaced in place when applied.  Returns 0, -1 (mask bit >= t) or
 * -4 (no full transform). */
/* fixed_weight_masks_u32 / _u64 (src/math/comb.cpp:13-82) for a mask width
 * of 32 or 64: writes up to `capacity` masks (n bits, k set, increasing) and
 * returns their count; OR_EINVAL for n above the width when k > 0. */
long long or_fixed_weight_masks(uint32_t n, uint32_t k, int width, uint64_t *out,
                                size_t capacity);

int or_global_greedy(uint64_t *rows, int t, int stride, int cols, const uint64_t *masks,
                     int mask_words, const uint32_t *weights, size_t n_cand,
                     int require_improvement, uint64_t *out_applied,
                     uint64_t *out_improvement);

#ifdef __cplusplus
}
#endif

#endif
