// Seeded synthetic inputs for the benchmark workloads (host code).
//
// Generator: XorShift64 (include/bmmpy/core/detail/xorshift64.hpp:10-40);
// Bernoulli(p) bits are `next_unit_double() < p`, filled row-major with columns
// inner.  Random invertible GF(2) matrices take each bit as the top bit of the
// splitmix64 finalizer of the next XorShift64 output and are rejected until
// full rank: XorShift64 is linear over GF(2), so bits read linearly off its
// outputs (the top bit is what `next_unit_double() < 0.5` tests) make every
// matrix a linear image of the 64-bit state, of rank at most about 64.
// Planted matrices are transform * sparse over GF(2) (BitMatrix::mul
// semantics, src/core/bit_matrix_algebra.cpp:27-50).
#include "common.hpp"

#include <cstring>
#include <vector>

namespace bmmhip {
namespace {

struct Xs64 {
    uint64_t s;
    explicit Xs64(uint64_t seed) : s(seed == 0 ? 0x9E3779B97F4A7C15ull : seed) {}
    uint64_t next() {
        uint64_t x = s;
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        s = x;
        return x;
    }
    uint64_t index(uint64_t ub) { return ub == 0 ? 0 : next() % ub; }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

void fill_mixed_half(Xs64& g, uint64_t* rows, size_t n_rows, size_t stride, size_t cols) {
    for (size_t r = 0; r < n_rows; ++r) {
        uint64_t* row = rows + r * stride;
        std::memset(row, 0, stride * 8);
        for (size_t c = 0; c < cols; ++c)
            if (mix64(g.next()) >> 63)
                row[c >> 6] |= 1ull << (c & 63);
    }
}

void fill(Xs64& g, uint64_t* rows, size_t n_rows, size_t stride, size_t cols, double p) {
    for (size_t r = 0; r < n_rows; ++r) {
        uint64_t* row = rows + r * stride;
        std::memset(row, 0, stride * 8);
        for (size_t c = 0; c < cols; ++c)
            if (g.unit() < p)
                row[c >> 6] |= 1ull << (c & 63);
    }
}

size_t gf2_rank(std::vector<uint64_t> m, size_t n, size_t stride, size_t cols) {
    size_t rank = 0;
    for (size_t c = 0; c < cols && rank < n; ++c) {
        size_t piv = rank;
        while (piv < n && !((m[piv * stride + (c >> 6)] >> (c & 63)) & 1u))
            ++piv;
        if (piv == n)
            continue;
        if (piv != rank)
            for (size_t w = 0; w < stride; ++w)
                std::swap(m[piv * stride + w], m[rank * stride + w]);
        for (size_t r = rank + 1; r < n; ++r)
            if ((m[r * stride + (c >> 6)] >> (c & 63)) & 1u)
                for (size_t w = 0; w < stride; ++w)
                    m[r * stride + w] ^= m[rank * stride + w];
        ++rank;
    }
    return rank;
}

} // namespace
} // namespace bmmhip

using namespace bmmhip;

extern "C" {

int bmm_gen_bernoulli(uint64_t seed, size_t n_rows, size_t cols, double p, uint64_t* out) {
    return guarded([&] {
        if (!out && n_rows && cols)
            fail(BMM_ERR_INVALID, "gen_bernoulli: null output");
        Xs64 g(seed);
        fill(g, out, n_rows, (cols + 63) / 64, cols, p);
    });
}

int bmm_gen_planted(uint64_t seed, size_t n_rows, size_t cols, double p, uint64_t* sparse,
                    uint64_t* mixed, uint64_t* transform) {
    return guarded([&] {
        if (!sparse || !mixed || !transform)
            fail(BMM_ERR_INVALID, "gen_planted: null output");
        Xs64 g(seed);
        const size_t w = (cols + 63) / 64, wn = (n_rows + 63) / 64;
        fill(g, sparse, n_rows, w, cols, p);
        std::vector<uint64_t> m(n_rows * wn);
        do {
            fill_mixed_half(g, m.data(), n_rows, wn, n_rows);
        } while (gf2_rank(m, n_rows, wn, n_rows) != n_rows);
        std::memcpy(transform, m.data(), m.size() * 8);
        for (size_t i = 0; i < n_rows; ++i) {
            uint64_t* o = mixed + i * w;
            std::memset(o, 0, w * 8);
            for (size_t k = 0; k < n_rows; ++k)
                if ((m[i * wn + (k >> 6)] >> (k & 63)) & 1u)
                    for (size_t x = 0; x < w; ++x)
                        o[x] ^= sparse[k * w + x];
        }
    });
}

int bmm_gen_batch(uint64_t seed, size_t n_mat, int32_t* rows, int32_t* cols, double* density,
                  uint64_t* words, size_t words_cap, int64_t* word_offset, size_t* n_words) {
    return guarded([&] {
        Xs64 g(seed);
        size_t off = 0;
        for (size_t m = 0; m < n_mat; ++m) {
            const int r = 8 + (int)g.index(21);
            const int c = 64 * (1 + (int)g.index(32));
            const double d = 0.1 + 0.4 * g.unit();
            const size_t w = (size_t)c / 64;
            if (words && off + (size_t)r * w > words_cap)
                fail(BMM_ERR_INVALID, "gen_batch: words_cap too small");
            rows[m] = r;
            cols[m] = c;
            density[m] = d;
            word_offset[m] = (int64_t)off;
            if (words) {
                fill(g, words + off, (size_t)r, w, (size_t)c, d);
            } else {
                for (size_t i = 0; i < (size_t)r * (size_t)c; ++i)
                    (void)g.unit();
            }
            off += (size_t)r * w;
        }
        *n_words = off;
    });
}

} // extern "C"
