/*
 * bmmpy_hip.h - C ABI of the MI355X (gfx950) row-XOR weight-minimization
 * library `libbmmpy_hip.so`.
 *
 * Plain pointers and sizes only.  Every entry point returns a BMM_* status; on
 * failure `bmm_hip_last_error()` (thread-local) holds the message.  Status
 * codes map onto the reference's exception types:
 *   BMM_ERR_INVALID  -> std::invalid_argument / ValueError
 *   BMM_ERR_OVERFLOW -> std::overflow_error / OverflowError
 *   BMM_ERR_RANGE    -> std::out_of_range / IndexError
 *   BMM_ERR_DEVICE   -> std::runtime_error / RuntimeError
 * Validation happens before any device query, as in the reference.
 *
 * Row data: `rows` rows of `words_per_row` uint64 words, contiguous and
 * unpadded, bit c of a row at word c/64 bit c%64 (the reference BitMatrix
 * layout minus its stride padding).  "host" entry points take host pointers
 * and are synchronous (the reference device seams are); "_dev" entry points
 * take device pointers, enqueue on `stream` (a hipStream_t, NULL = default)
 * and return without synchronising.
 *
 * Result records are ordered ascending by (weight, mask); mask 0 is never
 * reported.
 */
#ifndef BMMPY_HIP_H
#define BMMPY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMM_OK 0
#define BMM_ERR_INVALID 1
#define BMM_ERR_OVERFLOW 2
#define BMM_ERR_RANGE 3
#define BMM_ERR_DEVICE 4

/* Result record; same fields as the reference DeviceTopKEntry
 * (include/bmmpy/search/cuda_launch_topk.cuh:11-14), padded to 16 bytes. */
typedef struct bmm_topk_entry {
    uint64_t mask;
    uint32_t weight;
    uint32_t pad;
} bmm_topk_entry;

typedef struct bmm_device_info {
    int device_count;
    int device;
    int compute_units;
    int wave_size;
    int lds_per_block;
    int clock_khz;
    char arch[64];
    char name[128];
} bmm_device_info;

/* Thread-local message of the last failing call ("" after a success). */
const char *bmm_hip_last_error(void);

/* Library version string. */
const char *bmm_hip_version(void);

/* HIP-event timing of the dominant kernel launched by each entry point
 * (recorded on the launch stream).  timing_end synchronises on the recorded
 * events and returns the summed kernel time and the number of launches. */
int bmm_hip_timing_begin(void);
int bmm_hip_timing_end(double *total_ms, size_t *launches);

/* Device discovery for RuntimeFeatures (src/stub.cpp:17-43). device_count is 0
 * (and BMM_OK returned) when no device is visible. */
int bmm_hip_device_info(bmm_device_info *out);

/* ---- src/math: fwht_inplace (include/bmmpy/math/fwht.hpp:8-9) ----------- */
/* In-place unnormalised Walsh-Hadamard transform, natural order, wrapping
 * arithmetic; n must be a power of two (n <= 1 is a no-op). Host pointers. */
int bmm_hip_fwht_i32(int32_t *data, size_t n);
int bmm_hip_fwht_i16(int16_t *data, size_t n);
/* Same on device memory, enqueued on `stream`.  An int32 array of 2^28
 * elements on a 16-byte boundary takes two passes over HBM (other sizes and
 * alignments take three or fewer passes of another form; the result is the
 * same). */
int bmm_hip_fwht_i32_dev(int32_t *d_data, size_t n, void *stream);
int bmm_hip_fwht_i16_dev(int16_t *d_data, size_t n, void *stream);

/* ---- FwhtSearch::search inner loops (src/search/fwht_search.cpp:53-131) -
 * Histogram of active-column patterns, FWHT, exact top-k over masks
 * 1..2^rows-1.  Writes min(k, 2^rows-1) entries to `out` (capacity at least
 * that many) and the count to *n_out.  rows <= 48.  k <= BMM_TOPK_MAX runs the
 * on-chip top-k; any larger k (the reference takes any max_candidates,
 * src/search/fwht_search.cpp:95-131) ranks every mask through the column
 * histogram's transform in HBM and calc_scores_and_order's stable order,
 * for windows of at most 30 rows.  The same holds for bmm_hip_mitm_topk and
 * bmm_hip_bruteforce_topk. */
#define BMM_TOPK_MAX 4096
int bmm_hip_fwht_topk(const uint64_t *row_words, size_t rows, size_t words_per_row,
                      size_t cols, size_t k, bmm_topk_entry *out, size_t *n_out);

/* ---- run_cuda_mitm_fwht_search (include/bmmpy/search/cuda_mitm_fwht_launch.hpp:
 * 15-16) over a CompactSplitWindow (include/bmmpy/search/split_window_prep.hpp:
 * 10-20): `unique` entries of (q, r, multiplicity), pattern = q << low_bits |
 * r, total = sum of multiplicities.  Exact top-k over masks of
 * low_bits + high_bits rows. */
int bmm_hip_mitm_topk(const uint64_t *q, const uint64_t *r, const int32_t *mult, size_t unique,
                      size_t low_bits, size_t high_bits, int32_t total, size_t k,
                      bmm_topk_entry *out, size_t *n_out);

/* ---- run_cuda_bruteforce_search (include/bmmpy/search/cuda_bruteforce_launch.hpp:
 * 22-23): Gray-code sweep over all masks of `rows` rows, popcount over
 * words_per_row words without tail masking (src/search/bruteforce_search.cpp:
 * 157-169).  chunk_bits = 0 selects automatically. */
int bmm_hip_bruteforce_topk(const uint64_t *row_words, size_t rows, size_t words_per_row,
                            size_t chunk_bits, size_t k, bmm_topk_entry *out, size_t *n_out);

/* ---- per-target coset minimum (no single reference function: the reference
 * FwhtSearch ranking restricted to masks containing the target's bit,
 * src/search/fwht_search.cpp:59-126).  For every row i of an n_rows x cols
 * matrix: out_weight[i] = min over subsets S of the other rows of
 * wt(row_i ^ XOR_S), out_index[i] = the lowest such S as a bit set over the
 * other rows in ascending order.  n_rows in [2, 31]. */
int bmm_hip_coset_min(const uint64_t *row_words, size_t n_rows, size_t words_per_row,
                      size_t cols, uint32_t *out_weight, uint64_t *out_index);

/* Batched form: n_mat matrices, matrix m has rows[m] rows and cols[m] columns,
 * its words at row_words + word_offset[m] (rows[m] x ceil(cols[m]/64) words).
 * Outputs are indexed by out_offset[m] + target.  One persistent launch fed by
 * a device work queue. */
int bmm_hip_coset_min_batch(const uint64_t *row_words, const int64_t *word_offset,
                            const int32_t *rows, const int32_t *cols, size_t n_mat,
                            const int64_t *out_offset, uint32_t *out_weight, uint64_t *out_index);

/* Device-resident form: d_row_words on the device (offsets/sizes on the
 * host), outputs d_out_weight / d_out_index on the device, enqueued on
 * `stream`.  persistent_wgs = 0 selects the launch width automatically. */
int bmm_hip_coset_min_batch_dev(const uint64_t *d_row_words, const int64_t *word_offset,
                                const int32_t *rows, const int32_t *cols, size_t n_mat,
                                const int64_t *out_offset, uint32_t *d_out_weight,
                                uint64_t *d_out_index, size_t persistent_wgs, void *stream);

/* Reusable plan for repeated batched runs on device-resident data: the
 * descriptors and work queue are built once (optionally autotuning the launch
 * shape: workgroups, queue chunk, block size, timed on this batch), and
 * run() enqueues the whole sequence, as one captured hipGraph when use_graph
 * is set.  Small batches (few blocks, at most 2048 columns per matrix) run
 * as one kernel, which run() launches directly either way: a graph of one
 * node adds only the graph launch's own cost.  Pointers given at creation
 * must stay valid until destroy. */
typedef struct bmm_batch_info {
    size_t n_matrices;
    size_t n_slots;
    size_t total_blocks;
    int persistent_wgs;
    int chunk;
    int max_block_bits;
    double subset_evals;        /* sum over matrices of rows * 2^(rows-1) */
    double transform_elements;  /* sum over matrices of 2^rows */
} bmm_batch_info;
int bmm_hip_batch_plan_create(const uint64_t *d_row_words, const int64_t *word_offset,
                              const int32_t *rows, const int32_t *cols, size_t n_mat,
                              const int64_t *out_offset, uint32_t *d_out_weight,
                              uint64_t *d_out_index, size_t persistent_wgs, int autotune,
                              void **plan);
int bmm_hip_batch_plan_run(void *plan, int use_graph, void *stream);
int bmm_hip_batch_plan_info(const void *plan, bmm_batch_info *out);
/* Fixes the launch shape instead of the autotuned one: max_block_bits = 12
 * or 13 (int32 blocks of up to 2^bits masks) or 14 (packed int16 blocks,
 * every matrix at most 32767 columns); persistent_wgs / chunk = 0 pick
 * defaults.  Drops a captured graph (recaptured on the next graph run). */
int bmm_hip_batch_plan_set_shape(void *plan, int max_block_bits, size_t persistent_wgs,
                                 size_t chunk);
int bmm_hip_batch_plan_destroy(void *plan);

/* ---- single-target coset minimum over a wide window (planted MITM
 * workload; the split-transform enumeration of src/search/mitm_fwht_search.cpp:
 * 164-265 with the target bit forced).  For each i < n_targets: window =
 * window_rows[i*width .. i*width+width) (rows other than targets[i]),
 * out_weight[i] = min over subsets S of the window of wt(row_t ^ XOR_S),
 * out_mask[i] = the lowest such S (bit j = window row j).  width <= 47. */
int bmm_hip_window_coset_min(const uint64_t *row_words, size_t n_rows, size_t words_per_row,
                             size_t cols, const int32_t *window_rows, size_t width,
                             const int32_t *targets, size_t n_targets, uint32_t *out_weight,
                             uint64_t *out_mask);

/* ---- pairwise row overlaps (the synergy table of SASelector,
 * src/search/sa_selector_pairwise.cpp:21-52): out[i * n_rows + j] =
 * popcount(row_i & row_j) for i != j, 0 on the diagonal; rows of
 * words_per_row words with zero bits beyond the columns. */
int bmm_hip_pair_overlaps(const uint64_t *row_words, size_t n_rows, size_t words_per_row,
                          uint32_t *out);

/* ---- GF(2) row combination: out row i = XOR of the input rows selected by
 * mask i (mask_words words per mask, bit j of word q = input row 64 q + j).
 * Replaces BitMatrix::mul (src/core/bit_matrix_algebra.cpp:27-50), as used
 * by GlobalGreedyApplier::apply to apply its transform
 * (src/apply/global_greedy_applier.cpp:93-96).  *out_weight = total popcount
 * of the n_out output rows.  A mask bit at or beyond n_rows is
 * BMM_ERR_RANGE. */
int bmm_hip_row_combine(const uint64_t *row_words, size_t n_rows, size_t words_per_row,
                        const uint64_t *masks, size_t mask_words, size_t n_out,
                        uint64_t *out_words, uint64_t *out_weight);

/* ---- calc_scores_and_order (include/bmmpy/math/fwht.hpp:11-25,
 * src/math/fwht.cpp:59-126): s_by_mask[u] = clamp((n - h[u]) >> 1, 0, n)
 * for u >= 1, s_by_mask[0] = 0; cnt[v] = number of masks u in 1 .. m-1 with
 * score v; order[0 .. m-2] = those masks sorted by score, ascending u within
 * a score; off[v] = end of bucket v in order (inclusive prefix of cnt).
 * cnt and off hold n + 1 entries.  n < 0 is BMM_ERR_INVALID; m == 0 zeroes
 * cnt and off only.  Limits: n <= 2^28, m <= 2^31; the int16 form fails with
 * BMM_ERR_OVERFLOW when a score exceeds 32767 (the reference indexes out of
 * its offsets there). */
int bmm_hip_scores_order_i32(const int32_t *h, size_t m, int32_t n, int32_t *s_by_mask,
                             int32_t *order, int32_t *cnt, int32_t *off);
int bmm_hip_scores_order_i16(const int16_t *h, size_t m, int32_t n, int16_t *s_by_mask,
                             int32_t *order, int32_t *cnt, int32_t *off);

/* ---- fixed_weight_masks_u32 / _u64 (include/bmmpy/math/comb.hpp:7-12,
 * src/math/comb.cpp:13-82): every n-bit mask with exactly k set bits, in
 * increasing order.  *count = the number of masks (1 for k = 0, whose one
 * mask is 0 whatever n; 0 for k > n; else C(n, k)); with out == NULL only
 * *count is set, otherwise out must hold capacity >= *count masks.  n above
 * the mask width (32 / 64) is BMM_ERR_INVALID, checked after k = 0 as in the
 * reference. */
int bmm_hip_fixed_weight_masks_u32(uint32_t n, uint32_t k, uint32_t *out, size_t capacity,
                                   size_t *count);
int bmm_hip_fixed_weight_masks_u64(uint32_t n, uint32_t k, uint64_t *out, size_t capacity,
                                   size_t *count);

/* ---- candidate stream of ExactBasisSolver (src/solver/exact_basis_solver.cpp:
 * 29-124): the first min(k, 2^t - 1) nonzero masks of a t-row window
 * (t <= 32) in the order (weight, Gray index), i.e. the reference's Gray
 * enumeration counting-sorted by weight; mask bit r = row r.
 * *n_out = number written. */
int bmm_hip_weight_order(const uint64_t *row_words, size_t t, size_t words_per_row,
                         size_t cols, size_t k, uint32_t *out_mask, uint32_t *out_weight,
                         size_t *n_out);

/* ---- Gray-code coset sweep (config: pivot block of rows [0, block_rows)
 * minus the target).  For each target t in targets[]: block = rows
 * [0, block_rows) \ {t} in ascending order; step s visits block subset
 * gray(s) = s ^ (s >> 1) and evaluates wt(row_t ^ XOR(subset)) over
 * words_per_row words.  out_weight = minimum, out_first = first step s that
 * reaches it, out_mask = lowest block subset that reaches it. */
int bmm_hip_gray_coset_min(const uint64_t *row_words, size_t n_rows, size_t words_per_row,
                           size_t block_rows, const int32_t *targets, size_t n_targets,
                           uint32_t *out_weight, uint64_t *out_first, uint64_t *out_mask);
/* Device-resident form: d_rows on the device, results written as packed keys
 * d_first_key[i] = weight << 32 | first step, d_mask_key[i] = weight << 32 |
 * lowest mask.  chunk_bits = 0 selects automatically. */
int bmm_hip_gray_coset_min_dev(const uint64_t *d_rows, size_t n_rows, size_t words_per_row,
                               size_t block_rows, const int32_t *targets, size_t n_targets,
                               size_t chunk_bits, uint64_t *d_first_key, uint64_t *d_mask_key,
                               void *stream);

/* Allocation-free form of bmm_hip_gray_coset_min_dev (no reference
 * counterpart): for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results; nothing is allocated or freed: the task tables
 * (and rows zero-padded to a supported width) live in a 1 MiB device array
 * each GPU holds for the life of the process, built on the device from
 * by-value kernel arguments.  Limits: at most 64 targets, at most 61,440
 * workgroups (sum over targets of 2^(block bits - chunk_bits - 8)), and for
 * words_per_row other than 1, 2, 4, 8, 16, 32 at most 98,304 padded words;
 * beyond them BMM_ERR_INVALID.  Calls that overlap on one GPU share the array
 * and race: serialise them (one stream, or events). */
int bmm_hip_gray_coset_min_noalloc(const uint64_t *d_rows, size_t n_rows, size_t words_per_row,
                                   size_t block_rows, const int32_t *targets, size_t n_targets,
                                   size_t chunk_bits, uint64_t *d_first_key,
                                   uint64_t *d_mask_key, void *stream);

/* ---- src/apply: replay of a row-XOR sequence (RowWindow::row_xor,
 * src/core/bit_matrix_rows.cpp:12-20, as driven by GreedyApplier::apply,
 * src/apply/greedy_applier.cpp:92-95).  ops holds n_ops (source, target)
 * pairs of row indices; target ^= source, a no-op when they are equal.  Rows
 * are updated in place (host memory); *out_weight = total popcount after. */
int bmm_hip_replay(uint64_t *rows, size_t n_rows, size_t words_per_row, const uint16_t *ops,
                   size_t n_ops, uint64_t *out_weight);

/* ---- src/ga on the GPU: island model over row-XOR sequences ---------------
 * Individual = up to max_ops ops (src, dst) applied to the island's base
 * matrix; fitness = exact total weight after replay; elitism + tournament +
 * one-point crossover + mutation (append / replace / delete an op); every
 * migrate_every generations each island folds its best sequence into its base,
 * adopts its ring predecessor's base when that one is lighter, and runs
 * li_sweeps sweeps of pairwise local improvement on it (the reference GA's
 * local_improvement rule, applied in parallel rounds).  learned = 1
 * draws new ops from pairs ranked by the proposal MLP (trained online on exact
 * pair deltas); learned = 0 draws uniform random pairs.  rows <= 256, cols a
 * multiple of 2048. */
typedef struct bmm_ga_params {
    int islands;
    int population;
    int max_ops;
    int tournament;
    int elite;
    int migrate_every;
    int mutations;      /* mutation draws per child */
    int learned;
    int n_candidates;   /* ranked pairs kept per island (learned) */
    int train_steps;    /* MLP steps per migration epoch (learned) */
    int check_every;    /* generations between host checks of the best weight */
    int use_graph;      /* replay each generation as a captured hipGraph */
    int li_sweeps;      /* pairwise local-improvement sweeps per epoch (0 = off) */
    int rows;           /* filled in by create */
    float p_append;
    float p_replace;
    float p_delete;
    float learning_rate;
    float p_learned;    /* share of new ops drawn from the ranked pairs */
    uint64_t seed;
    int win_rows;       /* window local improvement: rows per window (0 = off, 14..24;
                           off for matrices wider than 8192 columns) */
    int win_rounds;     /* window rounds per epoch */
    int win_candidates; /* learned: candidates scored per target, best win_rows kept */
    int isd_trials;     /* information-set trials per island per epoch (0 = off, <= 64):
                           Gauss-Jordan elimination on random columns, then the lightest
                           basis from those rows and the island's (GlobalGreedyApplier);
                           up to 6 trials run side by side from one base and are selected
                           from together */
    /* The reference GA's operator settings (GeneticAlgorithmConfig,
     * include/bmmpy/ga/genetic_algorithm.hpp:17-29; src/ga/genetic_algorithm_*.cpp). */
    int num_parents;          /* tournament winners per crossover, 1..4 (1: copy) */
    int num_offspring;        /* children sharing one parent set and its cut points */
    float mutation_rate;      /* > 0: each child gets max(1, floor(rows * rate)) row XORs
                                 r_i ^= r_j (i, j uniform, i == j skipped), the rate adapted
                                 every generation (genetic_algorithm_core.cpp:33-41);
                                 0: `mutations` append / replace / delete draws */
    int catastrophe_threshold;  /* 0 = off; else after this many stale generations all but
                                   the best floor(population * catastrophe_survival) are
                                   replaced by random individuals (:15-37) */
    float catastrophe_survival;
    int restart_after;        /* 0 = off; after this many stale generations a child is a
                                 random individual with probability restart_p
                                 (genetic_algorithm_lifecycle.cpp:118-142: 15 and 0.2) */
    float restart_p;
    int init_rounds;          /* 1: before the first generation the island bases go through the
                                 epoch's improvement steps (li_sweeps, win_rounds, isd_trials);
                                 0: the first generation starts from the input */
} bmm_ga_params;

typedef struct bmm_ga_status {
    size_t generations;   /* total so far */
    size_t ran;           /* in this call */
    uint64_t evaluations; /* exact fitness evaluations so far */
    uint64_t best_weight;
    int best_island;
    int reached_target;
    double elapsed_s;
    double train_loss;    /* last MLP training loss (learned) */
    uint64_t li_moves;    /* local-improvement moves applied so far */
    uint64_t win_moves;   /* window-improvement moves applied so far */
    uint64_t isd_rows;    /* rows taken from information-set trials so far */
    uint64_t replayed_ops; /* row XOR ops replayed by fitness evaluation so far */
} bmm_ga_status;

int bmm_hip_ga_default_params(bmm_ga_params *out);
int bmm_hip_ga_create(const uint64_t *matrix, size_t rows, size_t cols,
                      const bmm_ga_params *params, void **ga);
/* Runs until max_generations, best weight <= target_weight, or time budget
 * (seconds, 0 = none), checking every check_every generations. */
int bmm_hip_ga_run(void *ga, size_t max_generations, uint64_t target_weight,
                   double time_budget_s, bmm_ga_status *out);
/* Best matrix found (rows x cols/64 words, host) and its weight. */
int bmm_hip_ga_best(void *ga, uint64_t *out_matrix, uint64_t *out_weight);
/* Island `island`'s best matrix and weight (island < 0: the lightest
 * island's); the per-island snapshot of IslandModel
 * (include/bmmpy/ga/types.hpp:45-53).  An island out of range is
 * BMM_ERR_RANGE. */
int bmm_hip_ga_island_best(void *ga, int island, uint64_t *out_matrix, uint64_t *out_weight);
int bmm_hip_ga_destroy(void *ga);

/* Per-island settings of an island model (IslandSpec / GeneticAlgorithmConfig
 * per island, bmmpy/ga.py:360-380): seeds[i] restarts island i's individual
 * random streams (XorShift64, seed 0 -> 0x9E3779B97F4A7C15); migrate[i] bit 0
 * = island i exports its best to its ring successor, bit 1 = it imports from
 * its predecessor (a zero export_count / import_count or interval clears the
 * bit, src/ga/island.cpp:149-228).  n = islands; either array may be NULL. */
int bmm_hip_ga_configure_islands(void *ga, const uint64_t *seeds, const int32_t *migrate,
                                 size_t n);

/* Island `island`'s base matrix (rows x cols/64 words) and best op sequence
 * (*n_ops ops of src | dst << 8, at most max_ops): replaying the ops on the
 * base (RowWindow::row_xor, src/core/bit_matrix_rows.cpp:12-20) gives
 * bmm_hip_ga_island_best's matrix. */
int bmm_hip_ga_best_ops(void *ga, int island, uint64_t *out_base, uint16_t *out_ops,
                        size_t *n_ops);

/* The learned proposal's ranked candidate pairs of island `island`
 * (src | dst << 8, *n_pairs = n_candidates; 0 without the learned proposal):
 * mutations draw from them with probability p_learned, and a crossover's
 * later parents start their segment at the best of 4 cut draws by how many
 * of their next 8 ops are among them. */
int bmm_hip_ga_proposals(void *ga, int island, uint16_t *out_pairs, size_t capacity,
                         size_t *n_pairs);

/* Replaces island `island`'s base matrix; its best becomes that base. */
int bmm_hip_ga_set_island_base(void *ga, int island, const uint64_t *matrix);

/* Counters of one island (the reference GA's _best_score, _no_improvement,
 * _catastrophe and adapted mutation rate, include/bmmpy/ga/genetic_algorithm.hpp). */
typedef struct bmm_ga_island_state {
    uint32_t best_weight;
    int32_t best_len;
    uint32_t no_improvement;
    uint32_t catastrophe;
    float mutation_rate;
    int32_t migrate;
} bmm_ga_island_state;

/* Island `island`'s populations: "next" = the population the next generation
 * evaluates (population x max_ops ops, lengths), "prev" = the population the
 * last generation evaluated with its exact fitness, rng = the individual
 * streams the next breeding step draws from, and the island counters.  Any
 * output may be NULL.  An island out of range is BMM_ERR_RANGE. */
int bmm_hip_ga_population(void *ga, int island, uint16_t *ops_next, int32_t *len_next,
                          uint16_t *ops_prev, int32_t *len_prev, uint32_t *fit_prev,
                          uint64_t *rng, bmm_ga_island_state *state);

/* ---- proposal MLP (256 -> 256 -> 256 -> 1, ReLU, bf16 products with fp32
 * accumulation on MFMA, fp32 master weights, Adam).  Parameters flat in the
 * order W1[256][256] (in x out), b1[256], W2[256][256], b2[256], w3[256],
 * b3[1].  Features are host floats, batch x 256; loss = mean squared error. */
size_t bmm_hip_mlp_num_params(void);
int bmm_hip_mlp_create(uint64_t seed, void **mlp);
int bmm_hip_mlp_destroy(void *mlp);
int bmm_hip_mlp_set_params(void *mlp, const float *params);
int bmm_hip_mlp_get_params(void *mlp, float *params);
int bmm_hip_mlp_forward(void *mlp, const float *x, size_t batch, float *y);
/* Fused scoring of row pairs: sketch = 2 u64 words per row (128 bits),
 * pairs[i] = src | dst << 8; y[i] = model output for the features
 * sketch(src) | sketch(dst) as +-1 (the same value mlp_forward gives). */
int bmm_hip_mlp_score_pairs(void *mlp, const uint64_t *sketch, size_t rows,
                            const uint16_t *pairs, size_t n, float *y);
int bmm_hip_mlp_gradients(void *mlp, const float *x, const float *target, size_t batch,
                          float *grads, float *loss);
int bmm_hip_mlp_train_step(void *mlp, const float *x, const float *target, size_t batch,
                           float learning_rate, float *loss);

/* ---- seeded synthetic workloads (no reference counterpart: they define the
 * benchmark inputs).  XorShift64 stream (include/bmmpy/core/detail/
 * xorshift64.hpp:10-40); Bernoulli(p) = next_unit_double() < p, row-major. */
int bmm_gen_bernoulli(uint64_t seed, size_t n_rows, size_t cols, double p, uint64_t *out);
/* sparse = Bernoulli(p) n_rows x cols; transform = random invertible
 * n_rows x n_rows (bit = top bit of splitmix64(next output), rejected until
 * full rank: XorShift64 is GF(2)-linear, so linearly extracted bits would cap
 * the rank near 64); mixed = transform * sparse.  Row strides ceil(cols/64)
 * and ceil(n_rows/64) words. */
int bmm_gen_planted(uint64_t seed, size_t n_rows, size_t cols, double p, uint64_t *sparse,
                    uint64_t *mixed, uint64_t *transform);
/* Per matrix: rows = 8 + index(21), cols = 64 (1 + index(32)),
 * density = 0.1 + 0.4 unit(), then the bits.  words may be NULL to size. */
int bmm_gen_batch(uint64_t seed, size_t n_mat, int32_t *rows, int32_t *cols, double *density,
                  uint64_t *words, size_t words_cap, int64_t *word_offset, size_t *n_words);

#ifdef __cplusplus
}
#endif

#endif
