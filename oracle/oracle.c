/*
 * Test infrastructure only (see oracle.h): CPU restatement of the reference's
 * row-XOR weight-minimization algorithms, used as the checker for the HIP path
 * and as the CPU baseline.  The product library never links or calls this.
 *
 * Pinning (what fixes each function independently of this file):
 *   or_xs_*            reference XorShift64 recurrence; first outputs checked in
 *                      tests/test_oracle.py against the closed recurrence.
 *   or_fwht_i32/i16    reference KATs fwht_tests.cpp:58-70; involution
 *                      H(H(x)) = N x; Parseval; brute-force Walsh sum N<=2^10;
 *                      reference fwht_inplace via oracle/_ref.
 *   or_histogram       KAT search_cpu_tests.cpp:128-152 (through split window);
 *                      weight identity vs direct popcount for every mask (t<=12).
 *   or_fwht_topk       KATs search_cpu_tests.cpp:18-52; golden vectors from the
 *                      reference FwhtSearch (tests/golden); brute-force sort.
 *   or_split_window    KAT search_cpu_tests.cpp:128-152; reference golden.
 *   or_mitm_topk       equals or_fwht_topk on random windows; reference golden
 *                      (28x96 k=8/128 equivalence matrix, edge cases).
 *   or_bruteforce_topk reference golden (16x512/1024/4096, k=8/128); untruncated
 *                      tail popcount checked against a planted tail bit.
 *   or_coset_min       equals the first target-containing entry of the reference
 *                      FwhtSearch full ranking (20x128 golden); brute force t<=12.
 *   or_gray_min        min weight and lowest mask equal the reference FwhtSearch
 *                      ranking of [block..., target]; first Gray step equals the
 *                      minimum inverse-Gray index over all minimising masks.
 *   or_replay          reference RowWindow::row_xor replay golden; weight
 *                      identity wt(a^b) = wt(a)+wt(b)-2 pop(a&b).
 *   or_greedy_apply    reference GreedyApplier golden.
 *   or_scores_order_*  reference KATs fwht_tests.cpp:72-107 (i32 and i16); the
 *                      live reference calc_scores_and_order via oracle/_ref on
 *                      random spectra; order checked as a stable sort.
 *   or_global_greedy   reference KAT test_matrix_api.py:66-105 (applies, keeps
 *                      the rank; identity untouched) and the live reference
 *                      GlobalGreedyApplier on random windows and candidates.
 *   or_exact_basis,    reference ExactBasisSolver KAT (tests/python/
 *   or_weight_order    test_matrix_api.py: 3x3, rank 2, total 3) and the live
 *                      reference on random and rank-deficient windows;
 *                      the order against a sort of all masks by
 *                      (weight, inverse Gray index).
 *   or_mul, or_rank    the live reference BitMatrix::mul / rank on random
 *                      shapes (tests/test_oracle.py); identity and
 *                      associativity; rank of random invertible products.
 *   generators         not reference behaviour (workload definitions); the
 *                      product generator must match them bit for bit, and
 *                      rank/inverse properties are checked directly (incl. a
 *                      256 x 256 transform: full rank needs nonlinear bits).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define OR_EINVAL (-1)
#define OR_EOVERFLOW (-2)
#define OR_ENOMEM (-3)
#define OR_ERANK (-4)

/* ---------------------------------------------------------------- RNG --- */
/* include/bmmpy/core/detail/xorshift64.hpp:10-40 */
void or_xs_seed(uint64_t *state, uint64_t seed) {
    *state = seed == 0 ? 0x9E3779B97F4A7C15ull : seed;
}

uint64_t or_xs_next(uint64_t *state) {
    uint64_t x = *state;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *state = x;
    return x;
}

uint64_t or_xs_index(uint64_t *state, uint64_t upper) {
    if (upper == 0)
        return 0;
    return or_xs_next(state) % upper;
}

double or_xs_unit(uint64_t *state) {
    return (double)(or_xs_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* ---------------------------------------------------------- matrices --- */
static int words_for(int cols) { return (cols + 63) / 64; }

static int get_bit(const uint64_t *row, int c) { return (int)((row[c >> 6] >> (c & 63)) & 1u); }

void or_fill_bernoulli(uint64_t *state, uint64_t *rows, int n_rows, int stride, int cols,
                       double p) {
    for (int r = 0; r < n_rows; ++r) {
        uint64_t *row = rows + (size_t)r * stride;
        memset(row, 0, sizeof(uint64_t) * (size_t)stride);
        for (int c = 0; c < cols; ++c)
            if (or_xs_unit(state) < p)
                row[c >> 6] |= 1ull << (c & 63);
    }
}

/* rank over GF(2) by forward elimination (src/core/bit_matrix_algebra.cpp:76-99). */
int or_rank(const uint64_t *rows, int n_rows, int stride, int cols) {
    uint64_t *tmp = malloc(sizeof(uint64_t) * (size_t)n_rows * (size_t)stride + 8);
    if (!tmp)
        return OR_ENOMEM;
    memcpy(tmp, rows, sizeof(uint64_t) * (size_t)n_rows * (size_t)stride);
    int rank = 0;
    for (int c = 0; c < cols && rank < n_rows; ++c) {
        int piv = rank;
        while (piv < n_rows && !get_bit(tmp + (size_t)piv * stride, c))
            ++piv;
        if (piv == n_rows)
            continue;
        if (piv != rank)
            for (int w = 0; w < stride; ++w) {
                uint64_t a = tmp[(size_t)piv * stride + w];
                tmp[(size_t)piv * stride + w] = tmp[(size_t)rank * stride + w];
                tmp[(size_t)rank * stride + w] = a;
            }
        for (int r = rank + 1; r < n_rows; ++r)
            if (get_bit(tmp + (size_t)r * stride, c))
                for (int w = 0; w < stride; ++w)
                    tmp[(size_t)r * stride + w] ^= tmp[(size_t)rank * stride + w];
        ++rank;
    }
    free(tmp);
    return rank;
}

/* out = a * b over GF(2): row i of out = XOR of rows k of b with a[i][k] = 1
 * (src/core/bit_matrix_algebra.cpp:27-50). */
void or_mul(const uint64_t *a, int a_rows, int a_stride, int a_cols, const uint64_t *b,
            int b_stride, int b_cols, uint64_t *out) {
    (void)b_cols;
    for (int i = 0; i < a_rows; ++i) {
        uint64_t *o = out + (size_t)i * b_stride;
        memset(o, 0, sizeof(uint64_t) * (size_t)b_stride);
        for (int k = 0; k < a_cols; ++k)
            if (get_bit(a + (size_t)i * a_stride, k))
                for (int w = 0; w < b_stride; ++w)
                    o[w] ^= b[(size_t)k * b_stride + w];
    }
}

/* splitmix64 finalizer: a nonlinear (over GF(2)) bijection of 64-bit words. */
static uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

/* Bit (r, c) = top bit of mix64(next output), row-major; rejected until full
 * rank.  XorShift64 is GF(2)-linear, so bits taken linearly from its outputs
 * (as `unit() < 0.5` does) would confine the matrix to rank about 64. Returns
 * the number of draws. */
int or_random_invertible(uint64_t *state, uint64_t *m, int n, int stride) {
    int draws = 0;
    for (;;) {
        ++draws;
        for (int r = 0; r < n; ++r) {
            uint64_t *row = m + (size_t)r * stride;
            memset(row, 0, sizeof(uint64_t) * (size_t)stride);
            for (int c = 0; c < n; ++c)
                if (mix64(or_xs_next(state)) >> 63)
                    row[c >> 6] |= 1ull << (c & 63);
        }
        if (or_rank(m, n, stride, n) == n)
            return draws;
    }
}

/* Planted instance: sparse = Bernoulli(p) n_rows x cols, transform = random
 * invertible n_rows x n_rows (same stream, drawn after), mixed = transform *
 * sparse.  Strides: ceil(cols/64) words for sparse/mixed, ceil(n/64) for the
 * transform. */
int or_gen_planted(uint64_t seed, int n_rows, int cols, double p, uint64_t *sparse,
                   uint64_t *mixed, uint64_t *transform) {
    uint64_t st;
    or_xs_seed(&st, seed);
    const int w = words_for(cols);
    const int wn = words_for(n_rows);
    or_fill_bernoulli(&st, sparse, n_rows, w, cols, p);
    int draws = or_random_invertible(&st, transform, n_rows, wn);
    or_mul(transform, n_rows, wn, n_rows, sparse, w, cols, mixed);
    return draws;
}

/* Batched uneven workload: per matrix r = 8 + idx(21), cols = 64 (1 + idx(32)),
 * density = 0.1 + 0.4 unit(), then bits (row-major).  Matrix m occupies
 * words[word_offsets[m] ...] as r rows of cols/64 words. Returns total words or
 * a negative error when words_cap is too small. */
long long or_gen_batch(uint64_t seed, int n_mat, int *rows_out, int *cols_out, double *dens_out,
                       uint64_t *words, size_t words_cap, size_t *word_offsets) {
    uint64_t st;
    or_xs_seed(&st, seed);
    size_t off = 0;
    for (int m = 0; m < n_mat; ++m) {
        const int r = 8 + (int)or_xs_index(&st, 21);
        const int c = 64 * (1 + (int)or_xs_index(&st, 32));
        const double d = 0.1 + 0.4 * or_xs_unit(&st);
        const int w = c / 64;
        if (off + (size_t)r * (size_t)w > words_cap)
            return OR_ENOMEM;
        rows_out[m] = r;
        cols_out[m] = c;
        dens_out[m] = d;
        word_offsets[m] = off;
        or_fill_bernoulli(&st, words + off, r, w, c, d);
        off += (size_t)r * (size_t)w;
    }
    return (long long)off;
}

/* -------------------------------------------------------------- FWHT --- */
/* Radix-2 butterflies len = 1, 2, 4, ... with wrapping add/sub
 * (src/math/fwht.cpp:28-57, src/math/fwht_scalar.cpp:19-53). */
int or_fwht_i32(int32_t *data, size_t n) {
    if (n <= 1)
        return 0;
    if ((n & (n - 1)) != 0)
        return OR_EINVAL;
    for (size_t len = 1; len < n; len <<= 1)
        for (size_t base = 0; base < n; base += 2 * len)
            for (size_t i = 0; i < len; ++i) {
                uint32_t u = (uint32_t)data[base + i], v = (uint32_t)data[base + len + i];
                data[base + i] = (int32_t)(u + v);
                data[base + len + i] = (int32_t)(u - v);
            }
    return 0;
}

int or_fwht_i16(int16_t *data, size_t n) {
    if (n <= 1)
        return 0;
    if ((n & (n - 1)) != 0)
        return OR_EINVAL;
    for (size_t len = 1; len < n; len <<= 1)
        for (size_t base = 0; base < n; base += 2 * len)
            for (size_t i = 0; i < len; ++i) {
                uint16_t u = (uint16_t)data[base + i], v = (uint16_t)data[base + len + i];
                data[base + i] = (int16_t)(uint16_t)(u + v);
                data[base + len + i] = (int16_t)(uint16_t)(u - v);
            }
    return 0;
}

/* --------------------------------------------------------- histogram --- */
static uint64_t tail_mask(int cols) {
    int tb = cols % 64;
    return tb == 0 ? ~0ull : ((1ull << tb) - 1);
}

/* b[p] = number of active columns with row pattern p (bit r = local row r);
 * total = number of active columns (src/search/fwht_search.cpp:53-91). */
int or_histogram(const uint64_t *base, int stride, int cols, const int *idx, int t,
                 int32_t *buckets, int64_t *total) {
    if (t < 0 || t > 30)
        return OR_EINVAL;
    const int w = words_for(cols);
    memset(buckets, 0, sizeof(int32_t) << t);
    int64_t tot = 0;
    for (int wi = 0; wi < w; ++wi) {
        uint64_t active = 0;
        for (int r = 0; r < t; ++r)
            active |= base[(size_t)idx[r] * stride + wi];
        if (wi + 1 == w)
            active &= tail_mask(cols);
        while (active) {
            int bit = __builtin_ctzll(active);
            uint32_t pat = 0;
            for (int r = 0; r < t; ++r)
                pat |= (uint32_t)((base[(size_t)idx[r] * stride + wi] >> bit) & 1u) << r;
            if (buckets[pat] == 0x7fffffff)
                return OR_EOVERFLOW;
            ++buckets[pat];
            if (++tot > 0x7fffffffll)
                return OR_EOVERFLOW;
            active &= active - 1;
        }
    }
    *total = tot;
    return 0;
}

/* ------------------------------------------------------------- top-k --- */
typedef struct {
    uint64_t mask;
    uint32_t weight;
} or_entry;

static int entry_less(const or_entry *a, const or_entry *b) {
    return a->weight < b->weight || (a->weight == b->weight && a->mask < b->mask);
}

static int entry_cmp(const void *pa, const void *pb) {
    const or_entry *a = pa, *b = pb;
    if (entry_less(a, b))
        return -1;
    if (entry_less(b, a))
        return 1;
    return 0;
}

/* Max-heap of the k smallest entries seen. */
typedef struct {
    or_entry *e;
    size_t n, k;
} or_topk;

static int topk_init(or_topk *h, size_t k) {
    h->n = 0;
    h->k = k;
    h->e = malloc(sizeof(or_entry) * (k ? k : 1));
    return h->e ? 0 : OR_ENOMEM;
}

static void topk_push(or_topk *h, uint64_t mask, uint32_t w) {
    if (h->k == 0 || mask == 0)
        return;
    or_entry x = {mask, w};
    if (h->n < h->k) {
        size_t i = h->n++;
        while (i > 0) {
            size_t p = (i - 1) / 2;
            if (!entry_less(&h->e[p], &x))
                break;
            h->e[i] = h->e[p];
            i = p;
        }
        h->e[i] = x;
        return;
    }
    if (!entry_less(&x, &h->e[0]))
        return;
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        const or_entry *big = &x;
        if (l < h->n && entry_less(big, &h->e[l])) {
            m = l;
            big = &h->e[l];
        }
        if (r < h->n && entry_less(big, &h->e[r]))
            m = r;
        if (m == i)
            break;
        h->e[i] = h->e[m];
        i = m;
    }
    h->e[i] = x;
}

static long long topk_finish(or_topk *h, uint64_t *out_mask, uint32_t *out_w) {
    qsort(h->e, h->n, sizeof(or_entry), entry_cmp);
    for (size_t i = 0; i < h->n; ++i) {
        out_mask[i] = h->e[i].mask;
        out_w[i] = h->e[i].weight;
    }
    long long n = (long long)h->n;
    free(h->e);
    return n;
}

/* ------------------------------------------------------ FWHT search --- */
/* src/search/fwht_search.cpp:34-132 (t>max_rows / k==0 early-outs are the
 * caller's; here t==0 and empty width return 0). */
long long or_fwht_topk(const uint64_t *base, int stride, int cols, const int *idx, int t,
                       size_t k, uint64_t *out_mask, uint32_t *out_w) {
    if (t == 0 || k == 0 || words_for(cols) == 0)
        return 0;
    if (t > 30)
        return OR_EINVAL;
    const size_t n = (size_t)1 << t;
    int32_t *b = malloc(sizeof(int32_t) * n);
    if (!b)
        return OR_ENOMEM;
    int64_t total = 0;
    int rc = or_histogram(base, stride, cols, idx, t, b, &total);
    if (rc) {
        free(b);
        return rc;
    }
    or_fwht_i32(b, n);
    or_topk h;
    if (topk_init(&h, k < n - 1 ? k : n - 1)) {
        free(b);
        return OR_ENOMEM;
    }
    for (size_t m = 1; m < n; ++m) {
        int64_t w2 = total - (int64_t)b[m];
        if (w2 < 0 || (w2 & 1))
            continue;
        topk_push(&h, m, (uint32_t)(w2 / 2));
    }
    free(b);
    return topk_finish(&h, out_mask, out_w);
}

/* ------------------------------------------------------ split window --- */
static int u64_cmp(const void *pa, const void *pb) {
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    return a < b ? -1 : (a > b);
}

/* Unique active-column patterns sorted ascending, split at low_bits
 * (src/search/split_window_prep.cpp:42-137). Returns the unique count (entries
 * beyond cap are counted but not written). */
long long or_split_window(const uint64_t *base, int stride, int cols, const int *idx, int t,
                          int low_bits, int32_t *total, uint64_t *q, uint64_t *r, int32_t *mult,
                          size_t cap) {
    if (low_bits > t || t > 64)
        return OR_EINVAL;
    *total = 0;
    const int w = words_for(cols);
    if (t == 0 || w == 0)
        return 0;
    uint64_t *pats = malloc(sizeof(uint64_t) * (size_t)cols + 8);
    if (!pats)
        return OR_ENOMEM;
    size_t np = 0;
    for (int wi = 0; wi < w; ++wi) {
        uint64_t active = 0;
        for (int rr = 0; rr < t; ++rr)
            active |= base[(size_t)idx[rr] * stride + wi];
        if (wi + 1 == w)
            active &= tail_mask(cols);
        while (active) {
            int bit = __builtin_ctzll(active);
            uint64_t pat = 0;
            for (int rr = 0; rr < t; ++rr)
                pat |= ((base[(size_t)idx[rr] * stride + wi] >> bit) & 1ull) << rr;
            pats[np++] = pat;
            active &= active - 1;
        }
    }
    qsort(pats, np, sizeof(uint64_t), u64_cmp);
    const uint64_t lowm = low_bits == 0 ? 0 : (low_bits >= 64 ? ~0ull : ((1ull << low_bits) - 1));
    size_t u = 0;
    for (size_t i = 0; i < np;) {
        size_t j = i;
        while (j < np && pats[j] == pats[i])
            ++j;
        if (u < cap) {
            q[u] = low_bits >= 64 ? 0 : (pats[i] >> low_bits);
            r[u] = pats[i] & lowm;
            mult[u] = (int32_t)(j - i);
        }
        ++u;
        i = j;
    }
    *total = (int32_t)np;
    free(pats);
    return (long long)u;
}

/* ------------------------------------------------------------- MITM --- */
/* src/search/mitm_fwht_search.cpp:56-60 (split), :145-186 (buckets, Gray
 * walk), :233-265 (per-x_low FWHT and scan), :292 ({} on all-zero window). */
long long or_mitm_topk(const uint64_t *base, int stride, int cols, const int *idx, int t,
                       size_t k, uint64_t *out_mask, uint32_t *out_w) {
    if (k == 0 || t == 0)
        return 0;
    if (t > 64)
        return OR_EINVAL;
    int high = t <= 26 ? t / 2 : (t / 2 < 13 ? t / 2 : 13);
    int low = t - high;
    if (low > 30 || high >= 31)
        return OR_EINVAL;
    if (words_for(cols) == 0)
        return 0;
    size_t cap = (size_t)cols + 1;
    uint64_t *q = malloc(sizeof(uint64_t) * cap), *r = malloc(sizeof(uint64_t) * cap);
    int32_t *mult = malloc(sizeof(int32_t) * cap);
    int32_t total = 0;
    long long u = or_split_window(base, stride, cols, idx, t, low, &total, q, r, mult, cap);
    if (u <= 0) {
        free(q);
        free(r);
        free(mult);
        return u;
    }
    const size_t nh = (size_t)1 << high;
    int32_t *buckets = calloc(nh, sizeof(int32_t)), *buf = malloc(sizeof(int32_t) * nh);
    int32_t *sgn = malloc(sizeof(int32_t) * (size_t)u);
    for (long long i = 0; i < u; ++i) {
        buckets[q[i]] += mult[i];
        sgn[i] = mult[i];
    }
    or_topk h;
    topk_init(&h, k);
    uint32_t xl = 0;
    const uint64_t nl = 1ull << low;
    for (uint64_t step = 0; step < nl; ++step) {
        if (step) {
            int bit = __builtin_ctzll(step);
            xl ^= 1u << bit;
            for (long long i = 0; i < u; ++i)
                if ((r[i] >> bit) & 1u) {
                    buckets[q[i]] -= 2 * sgn[i];
                    sgn[i] = -sgn[i];
                }
        }
        memcpy(buf, buckets, sizeof(int32_t) * nh);
        or_fwht_i32(buf, nh);
        for (size_t xh = 0; xh < nh; ++xh) {
            int64_t w2 = (int64_t)total - buf[xh];
            if (w2 < 0 || (w2 & 1))
                continue;
            uint64_t mask = ((uint64_t)xh << low) | xl;
            topk_push(&h, mask, (uint32_t)(w2 / 2));
        }
    }
    free(q);
    free(r);
    free(mult);
    free(buckets);
    free(buf);
    free(sgn);
    return topk_finish(&h, out_mask, out_w);
}

/* ------------------------------------------------------- brute force --- */
/* src/search/bruteforce_search.cpp:113-123 (chunk), :137-169 (prefix XOR,
 * prefix-only state, Gray sweep), popcount over words_per_row words with no
 * tail mask (:153, :167). */
long long or_bruteforce_topk(const uint64_t *base, int stride, int cols, const int *idx, int t,
                             size_t k, int chunk_bits, uint64_t *out_mask, uint32_t *out_w) {
    if (k == 0 || t == 0)
        return 0;
    if (t > 64)
        return OR_EINVAL;
    const int w = words_for(cols);
    if (w == 0)
        return 0;
    int chunk = chunk_bits == 0 ? (t < 12 ? t : 12) : chunk_bits;
    if (chunk_bits != 0 && (chunk_bits > t || chunk_bits >= 64))
        return OR_EINVAL;
    const int high = t - chunk;
    if (high >= 40)
        return OR_EINVAL;
    uint64_t *cur = malloc(sizeof(uint64_t) * (size_t)w);
    or_topk h;
    topk_init(&h, k);
    const uint64_t np = 1ull << high;
    for (uint64_t prefix = 0; prefix < np; ++prefix) {
        memset(cur, 0, sizeof(uint64_t) * (size_t)w);
        for (int hb = 0; hb < high; ++hb)
            if ((prefix >> hb) & 1u)
                for (int wi = 0; wi < w; ++wi)
                    cur[wi] ^= base[(size_t)idx[chunk + hb] * stride + wi];
        uint64_t mask = prefix << chunk;
        if (mask) {
            uint32_t pc = 0;
            for (int wi = 0; wi < w; ++wi)
                pc += (uint32_t)__builtin_popcountll(cur[wi]);
            topk_push(&h, mask, pc);
        }
        const uint64_t ls = 1ull << chunk;
        for (uint64_t step = 1; step < ls; ++step) {
            int bit = __builtin_ctzll(step);
            const uint64_t *row = base + (size_t)idx[bit] * stride;
            uint32_t pc = 0;
            for (int wi = 0; wi < w; ++wi) {
                cur[wi] ^= row[wi];
                pc += (uint32_t)__builtin_popcountll(cur[wi]);
            }
            mask ^= 1ull << bit;
            topk_push(&h, mask, pc);
        }
    }
    free(cur);
    return topk_finish(&h, out_mask, out_w);
}

/* ------------------------------------------------------- coset sweep --- */
/* For each target i: signed histogram over the other rows (pattern bit j = j-th
 * other row in ascending order; sign -1 where the target bit is set), FWHT,
 * weight(m) = (total - H[m]) / 2 with total = active columns of the whole
 * matrix.  This is the reference full-window ranking restricted to masks that
 * contain bit i (src/search/fwht_search.cpp:59-126). */
int or_coset_min(const uint64_t *base, int stride, int cols, int n_rows, uint32_t *out_w,
                 uint64_t *out_arg) {
    if (n_rows < 1 || n_rows > 31)
        return OR_EINVAL;
    const int t = n_rows - 1;
    const size_t n = (size_t)1 << t;
    const int w = words_for(cols);
    int32_t *b = malloc(sizeof(int32_t) * n);
    if (!b)
        return OR_ENOMEM;
    for (int tgt = 0; tgt < n_rows; ++tgt) {
        memset(b, 0, sizeof(int32_t) * n);
        int64_t total = 0;
        for (int wi = 0; wi < w; ++wi) {
            uint64_t active = 0;
            for (int r = 0; r < n_rows; ++r)
                active |= base[(size_t)r * stride + wi];
            if (wi + 1 == w)
                active &= tail_mask(cols);
            while (active) {
                int bit = __builtin_ctzll(active);
                uint32_t pat = 0;
                int j = 0;
                for (int r = 0; r < n_rows; ++r) {
                    if (r == tgt)
                        continue;
                    pat |= (uint32_t)((base[(size_t)r * stride + wi] >> bit) & 1u) << j;
                    ++j;
                }
                int tb = (int)((base[(size_t)tgt * stride + wi] >> bit) & 1u);
                b[pat] += tb ? -1 : 1;
                ++total;
                active &= active - 1;
            }
        }
        or_fwht_i32(b, n);
        uint32_t best = 0xffffffffu;
        uint64_t arg = 0;
        for (size_t m = 0; m < n; ++m) {
            uint32_t wt = (uint32_t)((total - (int64_t)b[m]) / 2);
            if (wt < best) {
                best = wt;
                arg = m;
            }
        }
        out_w[tgt] = best;
        out_arg[tgt] = arg;
    }
    free(b);
    return 0;
}

/* Gray sweep of target ^ XOR(block subset): step s visits block state
 * gray(s) = s ^ (s >> 1); per step one row XOR and a popcount over `words`
 * words (the reference brute-force inner loop, src/search/bruteforce_search.cpp:
 * 157-169, with the target row fixed into the state). */
int or_gray_min(const uint64_t *base, int stride, int words, int block_rows, int target,
                uint32_t *out_w, uint64_t *out_first_s, uint64_t *out_low_mask) {
    int blk[64];
    int b = 0;
    for (int r = 0; r < block_rows; ++r)
        if (r != target)
            blk[b++] = r;
    if (b > 40)
        return OR_EINVAL;
    uint64_t *cur = malloc(sizeof(uint64_t) * (size_t)words);
    memcpy(cur, base + (size_t)target * stride, sizeof(uint64_t) * (size_t)words);
    uint32_t best = 0;
    for (int wi = 0; wi < words; ++wi)
        best += (uint32_t)__builtin_popcountll(cur[wi]);
    uint64_t first_s = 0, low_mask = 0, mask = 0;
    const uint64_t ns = 1ull << b;
    for (uint64_t s = 1; s < ns; ++s) {
        int bit = __builtin_ctzll(s);
        const uint64_t *row = base + (size_t)blk[bit] * stride;
        uint32_t pc = 0;
        for (int wi = 0; wi < words; ++wi) {
            cur[wi] ^= row[wi];
            pc += (uint32_t)__builtin_popcountll(cur[wi]);
        }
        mask ^= 1ull << bit;
        if (pc < best) {
            best = pc;
            first_s = s;
            low_mask = mask;
        } else if (pc == best && mask < low_mask) {
            low_mask = mask;
        }
    }
    free(cur);
    *out_w = best;
    *out_first_s = first_s;
    *out_low_mask = low_mask;
    return 0;
}

/* ------------------------------------------------------------ replay --- */
long long or_replay(uint64_t *rows, int n_rows, int stride, int cols, const uint16_t *ops,
                    size_t n_ops) {
    for (size_t i = 0; i < n_ops; ++i) {
        int src = ops[2 * i], dst = ops[2 * i + 1];
        if (src >= n_rows || dst >= n_rows)
            return OR_EINVAL;
        if (src == dst)
            continue;
        for (int w = 0; w < stride; ++w)
            rows[(size_t)dst * stride + w] ^= rows[(size_t)src * stride + w];
    }
    long long total = 0;
    const int w = words_for(cols);
    for (int r = 0; r < n_rows; ++r)
        for (int wi = 0; wi < w; ++wi)
            total += __builtin_popcountll(rows[(size_t)r * stride + wi]);
    return total;
}

/* ----------------------------------------------------- greedy apply --- */
typedef struct {
    uint64_t mask;
    uint32_t weight;
    size_t order;
} or_cand;

static int cand_weight_stable(const void *pa, const void *pb) {
    const or_cand *a = pa, *b = pb;
    if (a->weight != b->weight)
        return a->weight < b->weight ? -1 : 1;
    return a->order < b->order ? -1 : (a->order > b->order);
}

/* src/apply/greedy_applier.cpp:13-109: stable sort by weight (after an
 * optional Fisher-Yates shuffle), map each mask through the running
 * coefficient basis c[], pick the lowest-index row with the best gain, replay
 * its XORs, update c[]. */
int or_greedy_apply(uint64_t *rows, int n_rows, int stride, int cols, const uint64_t *masks,
                    const uint32_t *weights, size_t n_cand, uint64_t min_gain, int stochastic,
                    uint64_t seed, uint64_t *out_applied, uint64_t *out_improvement) {
    if (n_rows > 64)
        return OR_EINVAL;
    uint64_t prng = seed == 0 ? 0xDEADBull : seed;
    uint64_t c[64];
    uint64_t cw[64];
    const int w = words_for(cols);
    for (int i = 0; i < n_rows; ++i) {
        c[i] = 1ull << i;
        cw[i] = 0;
        for (int wi = 0; wi < w; ++wi)
            cw[i] += (uint64_t)__builtin_popcountll(rows[(size_t)i * stride + wi]);
    }
    or_cand *cs = malloc(sizeof(or_cand) * (n_cand ? n_cand : 1));
    size_t *perm = malloc(sizeof(size_t) * (n_cand ? n_cand : 1));
    for (size_t i = 0; i < n_cand; ++i)
        perm[i] = i;
    if (stochastic)
        for (size_t i = n_cand; i > 1; --i) {
            uint64_t x = prng;
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            prng = x;
            size_t j = (size_t)(x % i);
            size_t tmp = perm[i - 1];
            perm[i - 1] = perm[j];
            perm[j] = tmp;
        }
    for (size_t i = 0; i < n_cand; ++i) {
        cs[i].mask = masks[perm[i]];
        cs[i].weight = weights[perm[i]];
        cs[i].order = i;
    }
    qsort(cs, n_cand, sizeof(or_cand), cand_weight_stable);
    uint64_t applied = 0, improvement = 0;
    for (size_t ci = 0; ci < n_cand; ++ci) {
        const uint64_t m = cs[ci].mask;
        uint64_t a = 0;
        for (int i = 0; i < n_rows; ++i)
            if (__builtin_popcountll(m & c[i]) & 1)
                a |= 1ull << i;
        if (a == 0)
            continue;
        int best_rows[64];
        int nbest = 0;
        uint64_t best_gain = 0;
        for (int r = 0; r < n_rows; ++r) {
            if (!((a >> r) & 1u))
                continue;
            uint64_t old = cw[r], cand = cs[ci].weight;
            if (old <= cand)
                continue;
            uint64_t gain = old - cand;
            if (gain < min_gain)
                continue;
            if (gain > best_gain) {
                best_gain = gain;
                best_rows[0] = r;
                nbest = 1;
            } else if (gain == best_gain) {
                best_rows[nbest++] = r;
            }
        }
        if (nbest == 0)
            continue;
        int chosen = best_rows[0];
        if (stochastic) {
            uint64_t x = prng;
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            prng = x;
            chosen = best_rows[x % (uint64_t)nbest];
        }
        for (int i = 0; i < n_rows; ++i)
            if (i != chosen && ((a >> i) & 1u))
                for (int wi = 0; wi < stride; ++wi)
                    rows[(size_t)chosen * stride + wi] ^= rows[(size_t)i * stride + wi];
        cw[chosen] = cs[ci].weight;
        for (int j = 0; j < n_rows; ++j)
            if (j != chosen && ((a >> j) & 1u))
                c[j] ^= c[chosen];
        ++applied;
        improvement += best_gain;
    }
    free(cs);
    free(perm);
    *out_applied = applied;
    *out_improvement = improvement;
    return 0;
}

/* ----------------------------------------------------- exact basis --- */
/* src/solver/exact_basis_solver.cpp:29-54: Gray enumeration of all nonzero
 * masks, weight = popcount of the running XOR (whole words; padding bits are
 * zero).  Fills w[i - 1] = weight(gray(i)). */
static int eb_weights(const uint64_t *base, int stride, int cols, const int *idx, int t,
                      uint32_t *w) {
    const int words = words_for(cols);
    uint64_t *cur = calloc((size_t)(words > 0 ? words : 1), sizeof(uint64_t));
    if (!cur)
        return OR_ENOMEM;
    const uint64_t states = ((uint64_t)1 << t) - 1;
    uint32_t prev = 0;
    for (uint64_t i = 1; i <= states; ++i) {
        const uint32_t g = (uint32_t)(i ^ (i >> 1));
        const int r = __builtin_ctz(g ^ prev);
        const uint64_t *row = base + (size_t)idx[r] * stride;
        uint32_t pc = 0;
        for (int k = 0; k < words; ++k) {
            cur[k] ^= row[k];
            pc += (uint32_t)__builtin_popcountll(cur[k]);
        }
        w[i - 1] = pc;
        prev = g;
    }
    free(cur);
    return 0;
}

/* :105-124: stable counting sort of the Gray stream by weight. */
static long long eb_order(const uint32_t *w, uint64_t states, int cols, uint32_t *out_idx) {
    size_t *off = calloc((size_t)cols + 2, sizeof(size_t));
    if (!off)
        return OR_ENOMEM;
    for (uint64_t i = 0; i < states; ++i)
        ++off[w[i] + 1];
    for (int q = 0; q <= cols; ++q)
        off[q + 1] += off[q];
    for (uint64_t i = 0; i < states; ++i)
        out_idx[off[w[i]]++] = (uint32_t)(i + 1);
    free(off);
    return (long long)states;
}

long long or_weight_order(const uint64_t *base, int stride, int cols, const int *idx, int t,
                          size_t k, uint32_t *out_masks, uint32_t *out_w) {
    if (t < 1 || t > 24 || cols < 0)
        return OR_EINVAL;
    const uint64_t states = ((uint64_t)1 << t) - 1;
    uint32_t *w = malloc(sizeof(uint32_t) * states);
    uint32_t *order = malloc(sizeof(uint32_t) * states);
    if (!w || !order) {
        free(w);
        free(order);
        return OR_ENOMEM;
    }
    int rc = eb_weights(base, stride, cols, idx, t, w);
    if (rc == 0 && eb_order(w, states, cols, order) < 0)
        rc = OR_ENOMEM;
    long long n = 0;
    if (rc == 0) {
        n = (long long)(k < states ? k : states);
        for (long long q = 0; q < n; ++q) {
            const uint32_t i = order[q];
            out_masks[q] = i ^ (i >> 1);
            out_w[q] = w[i - 1];
        }
    }
    free(w);
    free(order);
    return rc ? rc : n;
}

/* :58-157: greedy pivot-basis selection over the ordered stream. */
int or_exact_basis(const uint64_t *base, int stride, int cols, const int *idx, int t,
                   uint32_t *out_masks, uint32_t *out_w, size_t cap, uint64_t *out_total) {
    if (t < 1 || t > 24 || cols < 0)
        return OR_EINVAL;
    const int words = words_for(cols) > 0 ? words_for(cols) : 1;
    const uint64_t states = ((uint64_t)1 << t) - 1;
    uint32_t *w = malloc(sizeof(uint32_t) * states);
    uint32_t *order = malloc(sizeof(uint32_t) * states);
    uint64_t *piv = calloc((size_t)t * words, sizeof(uint64_t)); /* reduced rows by pivot */
    int *piv_bit = malloc(sizeof(int) * (size_t)t);
    uint64_t *v = malloc(sizeof(uint64_t) * words);
    uint64_t *sub = malloc(sizeof(uint64_t) * (size_t)t * words);
    int rc = (!w || !order || !piv || !piv_bit || !v || !sub) ? OR_ENOMEM : 0;
    if (rc == 0)
        rc = eb_weights(base, stride, cols, idx, t, w);
    if (rc == 0 && eb_order(w, states, cols, order) < 0)
        rc = OR_ENOMEM;
    int rank = 0, n_piv = 0;
    uint64_t total = 0;
    if (rc == 0) {
        for (int r = 0; r < t; ++r)
            memcpy(sub + (size_t)r * words, base + (size_t)idx[r] * stride, sizeof(uint64_t) * words);
        rank = or_rank(sub, t, words, cols);
        for (uint64_t q = 0; q < states && n_piv < rank; ++q) {
            const uint32_t i = order[q];
            const uint32_t mask = i ^ (i >> 1);
            memset(v, 0, sizeof(uint64_t) * words);
            for (int r = 0; r < t; ++r)
                if ((mask >> r) & 1u)
                    for (int k2 = 0; k2 < words; ++k2)
                        v[k2] ^= sub[(size_t)r * words + k2];
            /* reduce by existing pivots (highest set bit of each) */
            for (int p = 0; p < n_piv; ++p)
                if ((v[piv_bit[p] >> 6] >> (piv_bit[p] & 63)) & 1u)
                    for (int k2 = 0; k2 < words; ++k2)
                        v[k2] ^= piv[(size_t)p * words + k2];
            int hb = -1;
            for (int k2 = words - 1; k2 >= 0 && hb < 0; --k2)
                if (v[k2])
                    hb = k2 * 64 + 63 - __builtin_clzll(v[k2]);
            if (hb < 0)
                continue;
            /* keep pivots fully reduced against the new one */
            for (int p = 0; p < n_piv; ++p)
                if ((piv[(size_t)p * words + (hb >> 6)] >> (hb & 63)) & 1u)
                    for (int k2 = 0; k2 < words; ++k2)
                        piv[(size_t)p * words + k2] ^= v[k2];
            memcpy(piv + (size_t)n_piv * words, v, sizeof(uint64_t) * words);
            piv_bit[n_piv] = hb;
            if ((size_t)n_piv < cap) {
                out_masks[n_piv] = mask;
                out_w[n_piv] = w[i - 1];
            }
            total += w[i - 1];
            ++n_piv;
        }
    }
    free(w);
    free(order);
    free(piv);
    free(piv_bit);
    free(v);
    free(sub);
    if (rc)
        return rc;
    *out_total = total;
    return n_piv;
}

/* ------------------------------------------- scores and weight order --- */
/* calc_scores_and_order (src/math/fwht.cpp:59-100): score of mask u >= 1 =
 * clamp((n - h[u]) >> 1, 0, n) (floor shift), s[0] = 0; cnt = histogram of
 * the scores of masks 1 .. m-1; order = counting sort of those masks by
 * score, stable; off[v] ends as the end of bucket v.  The difference is
 * taken in 64 bits (the reference's 32-bit one agrees wherever it does not
 * overflow).  The int16 form returns OR_EOVERFLOW for a score above 32767,
 * where the reference would index outside its offsets. */
static int scores_order_core(const int32_t *h32, const int16_t *h16, size_t m, int32_t n,
                             int32_t *s32, int16_t *s16, int32_t *order, int32_t *cnt,
                             int32_t *off) {
    if (n < 0)
        return OR_EINVAL;
    for (int64_t v = 0; v <= n; ++v)
        cnt[v] = off[v] = 0;
    if (m == 0)
        return 0;
    if (s32)
        s32[0] = 0;
    else
        s16[0] = 0;
    for (size_t u = 1; u < m; ++u) {
        const int64_t hu = h32 ? h32[u] : h16[u];
        int64_t su = ((int64_t)n - hu) >> 1;
        su = su < 0 ? 0 : (su > n ? n : su);
        if (s32) {
            s32[u] = (int32_t)su;
        } else {
            if (su > 32767)
                return OR_EOVERFLOW;
            s16[u] = (int16_t)su;
        }
        cnt[su] += 1;
    }
    for (int64_t v = 0; v < n; ++v)
        off[v + 1] = off[v] + cnt[v];
    for (size_t u = 1; u < m; ++u) {
        const int32_t su = s32 ? s32[u] : s16[u];
        order[off[su]++] = (int32_t)u;
    }
    return 0;
}

int or_scores_order_i32(const int32_t *h, size_t m, int32_t n, int32_t *s_by_mask,
                        int32_t *order, int32_t *cnt, int32_t *off) {
    return scores_order_core(h, NULL, m, n, s_by_mask, NULL, order, cnt, off);
}

int or_scores_order_i16(const int16_t *h, size_t m, int32_t n, int16_t *s_by_mask,
                        int32_t *order, int32_t *cnt, int32_t *off) {
    return scores_order_core(NULL, h, m, n, NULL, s_by_mask, order, cnt, off);
}

/* ------------------------------------------------ fixed-weight masks --- */
/* fixed_weight_masks_u32 / _u64 (src/math/comb.cpp:13-82): k = 0 gives the
 * single mask 0 before any width check; n above the width is an error; k > n
 * gives none; else the n-bit masks with k set bits in increasing order,
 * stepped by Gosper's successor m' = (((r ^ m) >> 2) / (m & -m)) | r with
 * r = m + (m & -m) until m reaches 2^n (r wraps to 0 after the last mask
 * when n = 64).  Pinned by the reference's known answers (n 5 k 3, n 6 k 2),
 * by brute force over all masks below 2^n for n <= 16, and by the count
 * C(n, k). */
long long or_fixed_weight_masks(uint32_t n, uint32_t k, int width, uint64_t *out,
                                size_t capacity) {
    size_t count = 0;
    if (k == 0) {
        if (capacity)
            out[0] = 0;
        return 1;
    }
    if (n > (uint32_t)width)
        return OR_EINVAL;
    if (k > n)
        return 0;
    uint64_t m = (k == 64) ? ~(uint64_t)0 : (((uint64_t)1 << k) - 1);
    for (;;) {
        if (n < 64 && (m >> n) != 0)
            break;
        if (count < capacity)
            out[count] = m;
        ++count;
        const uint64_t low = m & (~m + 1);
        const uint64_t r = m + low;
        if (r == 0)
            break;
        m = (((r ^ m) >> 2) / low) | r;
    }
    return (long long)count;
}

/* ------------------------------------------------ global greedy apply --- */
typedef struct {
    const uint64_t *mask;
    uint32_t weight;
    size_t order;
} or_pool;

static int pool_weight_stable(const void *pa, const void *pb) {
    const or_pool *a = pa, *b = pb;
    if (a->weight != b->weight)
        return a->weight < b->weight ? -1 : 1;
    return a->order < b->order ? -1 : (a->order > b->order);
}

static int highest_bit_words(const uint64_t *v, int words) {
    for (int w = words - 1; w >= 0; --w)
        if (v[w])
            return w * 64 + 63 - __builtin_clzll(v[w]);
    return -1;
}

/* GlobalGreedyApplier::apply (src/apply/global_greedy_applier.cpp:15-105):
 * the pool is the candidates (mask words over t bits) followed by the unit
 * rows weighted by their popcounts, stably sorted by weight; masks enter a
 * pivot basis (highest set bit, include/bmmpy/core/detail/xor_basis.hpp:
 * 123-160) in that order until t are chosen; the transform whose row i is
 * chosen mask i replaces the window when the total weight does not grow
 * (strictly drops when require_improvement).  Returns 0, OR_EINVAL for a
 * mask bit at or beyond t, OR_ERANK when no full transform exists. */
int or_global_greedy(uint64_t *rows, int t, int stride, int cols, const uint64_t *masks,
                     int mask_words, const uint32_t *weights, size_t n_cand,
                     int require_improvement, uint64_t *out_applied,
                     uint64_t *out_improvement) {
    *out_applied = 0;
    *out_improvement = 0;
    if (t == 0)
        return 0;
    const int mw = (t + 63) / 64, w = words_for(cols);
    const size_t n_pool = n_cand + (size_t)t;
    uint64_t *pm = calloc(n_pool * (size_t)mw, sizeof(uint64_t));
    or_pool *pool = malloc(sizeof(or_pool) * n_pool);
    uint64_t *piv = calloc((size_t)t * (size_t)mw, sizeof(uint64_t));
    char *used = calloc((size_t)t, 1);
    uint64_t *chosen = calloc((size_t)t * (size_t)mw, sizeof(uint64_t));
    uint64_t *after = calloc((size_t)t * (size_t)stride, sizeof(uint64_t));
    uint64_t *v = calloc((size_t)mw, sizeof(uint64_t));
    int rc = 0;
    for (size_t c = 0; c < n_cand && rc == 0; ++c)
        for (int q = 0; q < mask_words; ++q) {
            const uint64_t word = masks[c * (size_t)mask_words + q];
            const int lo = q * 64;
            const uint64_t allowed =
                lo >= t ? 0ull : (t - lo >= 64 ? ~0ull : (1ull << (t - lo)) - 1ull);
            if (word & ~allowed) {
                rc = OR_EINVAL;
                break;
            }
            if (q < mw)
                pm[c * (size_t)mw + q] = word;
        }
    if (rc == 0) {
        uint64_t old_weight = 0;
        for (size_t c = 0; c < n_cand; ++c)
            pool[c] = (or_pool){pm + c * (size_t)mw, weights[c], c};
        for (int r = 0; r < t; ++r) {
            uint64_t pc = 0;
            for (int wi = 0; wi < w; ++wi)
                pc += (uint64_t)__builtin_popcountll(rows[(size_t)r * stride + wi]);
            old_weight += pc;
            uint64_t *u = pm + (n_cand + (size_t)r) * (size_t)mw;
            u[r / 64] = 1ull << (r % 64);
            pool[n_cand + (size_t)r] = (or_pool){u, (uint32_t)pc, n_cand + (size_t)r};
        }
        qsort(pool, n_pool, sizeof(or_pool), pool_weight_stable);
        int n_chosen = 0;
        for (size_t p = 0; p < n_pool && n_chosen < t; ++p) {
            memcpy(v, pool[p].mask, sizeof(uint64_t) * (size_t)mw);
            int hb;
            while ((hb = highest_bit_words(v, mw)) >= 0 && used[hb])
                for (int q = 0; q < mw; ++q)
                    v[q] ^= piv[(size_t)hb * mw + q];
            if (hb < 0)
                continue;
            used[hb] = 1;
            memcpy(piv + (size_t)hb * mw, v, sizeof(uint64_t) * (size_t)mw);
            memcpy(chosen + (size_t)n_chosen * mw, pool[p].mask, sizeof(uint64_t) * (size_t)mw);
            ++n_chosen;
        }
        if (n_chosen != t) {
            rc = OR_ERANK;
        } else {
            uint64_t new_weight = 0;
            for (int i = 0; i < t; ++i) {
                uint64_t *o = after + (size_t)i * stride;
                for (int k = 0; k < t; ++k)
                    if ((chosen[(size_t)i * mw + k / 64] >> (k % 64)) & 1u)
                        for (int wi = 0; wi < stride; ++wi)
                            o[wi] ^= rows[(size_t)k * stride + wi];
                for (int wi = 0; wi < w; ++wi)
                    new_weight += (uint64_t)__builtin_popcountll(o[wi]);
            }
            if (!(new_weight > old_weight ||
                  (require_improvement && new_weight == old_weight))) {
                memcpy(rows, after, sizeof(uint64_t) * (size_t)t * (size_t)stride);
                *out_applied = 1;
                *out_improvement = old_weight - new_weight;
            }
        }
    }
    free(pm);
    free(pool);
    free(piv);
    free(used);
    free(chosen);
    free(after);
    free(v);
    return rc;
}
