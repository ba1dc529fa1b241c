// SASelector (src/search/sa_selector.cpp:10-84, sa_selector_pairwise.cpp,
// sa_selector_higher_order.cpp).  The annealing depends on the order of the
// XorShift64 draws, so its steps follow the reference exactly: the random
// initial partition, the temperature probe, per iteration one window slot and
// one pool slot, acceptance of improving moves and of others with
// probability exp(delta / T) while T > min_temperature, geometric cooling, and
// the best state by (score, then lexicographically smallest sorted rows).
// The pairwise synergy table (2 popcount(r_i & r_j)) comes from the GPU.
#include "bmmpy.hpp"

#include "bmmpy_hip.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace bmmpy {
namespace {

struct Xs64 {
    std::uint64_t s;
    explicit Xs64(std::uint64_t seed) : s(seed == 0 ? 0x9E3779B97F4A7C15ull : seed) {}
    std::uint64_t next() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    }
    std::size_t index(std::size_t ub) { return ub == 0 ? 0 : (std::size_t)(next() % ub); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Restart {
    std::vector<std::size_t> rows;
    std::int64_t score = std::numeric_limits<std::int64_t>::min();
    std::size_t accepted = 0;
    std::size_t iterations_run = 0;
    std::size_t best_iteration = 0;
};

bool better(std::int64_t a, const std::vector<std::size_t>& ra, std::int64_t b,
            const std::vector<std::size_t>& rb) {
    return a != b ? a > b : ra < rb;
}

std::vector<std::size_t> sorted(std::vector<std::size_t> v) {
    std::sort(v.begin(), v.end());
    return v;
}

void partition(std::size_t n, std::size_t k, Xs64& g, std::vector<std::size_t>& window,
               std::vector<std::size_t>& pool) {
    pool.resize(n);
    std::iota(pool.begin(), pool.end(), std::size_t{0});
    for (std::size_t i = 0; i < k; ++i)
        std::swap(pool[i], pool[i + g.index(n - i)]);
    window.assign(pool.begin(), pool.begin() + (std::ptrdiff_t)k);
    pool.erase(pool.begin(), pool.begin() + (std::ptrdiff_t)k);
}

// delta(idx_w, idx_p) scores swapping window[idx_w] for pool[idx_p];
// commit(idx_w, idx_p) is called before the swap when a move is accepted.
template <typename Delta, typename Commit>
Restart anneal(const SASelectorConfig& c, Xs64& g, std::vector<std::size_t>& window,
               std::vector<std::size_t>& pool, std::int64_t score, Delta&& delta,
               Commit&& commit) {
    Restart best;
    best.rows = sorted(window);
    best.score = score;
    best.iterations_run = pool.empty() ? 0 : c.iterations;
    if (pool.empty() || c.iterations == 0)
        return best;
    double temperature = 1.0;
    {
        const std::size_t samples = std::max<std::size_t>(c.temperature_probe_samples, 1);
        double neg_sum = 0.0;
        std::size_t neg = 0;
        for (std::size_t q = 0; q < samples; ++q) {
            const std::size_t iw = g.index(window.size());
            const std::size_t ip = g.index(pool.size());
            const std::int64_t d = delta(iw, ip);
            if (d < 0) {
                neg_sum += (double)(-d);
                ++neg;
            }
        }
        if (neg != 0)
            temperature = std::max(-(neg_sum / (double)neg) / std::log(c.initial_acceptance_probability),
                                   c.min_temperature);
    }
    std::size_t accepted = 0;
    for (std::size_t iter = 0; iter < c.iterations; ++iter) {
        const std::size_t iw = g.index(window.size());
        const std::size_t ip = g.index(pool.size());
        const std::int64_t d = delta(iw, ip);
        bool accept = d > 0;
        if (!accept && temperature > c.min_temperature)
            accept = g.unit() < std::exp((double)d / temperature);
        if (accept) {
            commit(iw, ip);
            std::swap(window[iw], pool[ip]);
            score += d;
            ++accepted;
            std::vector<std::size_t> rows = sorted(window);
            if (better(score, rows, best.score, best.rows)) {
                best.rows = std::move(rows);
                best.score = score;
                best.accepted = accepted;
                best.best_iteration = iter + 1;
            }
        }
        temperature = std::max(c.min_temperature, temperature * c.cooling_rate);
    }
    return best;
}

// Pairwise synergy 2 popcount(r_i & r_j) (0 on the diagonal), from the GPU.
std::vector<std::uint32_t> synergy_table(const BitMatrix& m) {
    const std::size_t n = m.rows(), words = m.words_per_row();
    std::vector<std::uint64_t> packed(n * words);
    for (std::size_t r = 0; r < n; ++r)
        std::copy(m.row_words(r), m.row_words(r) + words, packed.begin() + (std::ptrdiff_t)(r * words));
    std::vector<std::uint32_t> t(n * n, 0);
    throw_status(bmm_hip_pair_overlaps(packed.data(), n, words, t.data()));
    for (std::uint32_t& v : t) {
        if (v > std::numeric_limits<std::uint32_t>::max() / 2)
            throw std::overflow_error("SASelector: pairwise synergy exceeds uint32 range");
        v *= 2;
    }
    return t;
}

Restart pairwise_restart(std::size_t n, std::size_t k, const SASelectorConfig& c, Xs64& g,
                         const std::vector<std::uint32_t>& T) {
    std::vector<std::size_t> window, pool;
    partition(n, k, g, window, pool);
    std::int64_t score = 0;
    for (std::size_t i = 0; i < window.size(); ++i)
        for (std::size_t j = i + 1; j < window.size(); ++j)
            score += T[window[i] * n + window[j]];
    auto delta = [&](std::size_t iw, std::size_t ip) {
        const std::size_t out = window[iw], in = pool[ip];
        std::int64_t d = 0;
        for (std::size_t i = 0; i < window.size(); ++i)
            if (i != iw)
                d += (std::int64_t)T[in * n + window[i]] - (std::int64_t)T[out * n + window[i]];
        return d;
    };
    return anneal(c, g, window, pool, score, delta, [](std::size_t, std::size_t) {});
}

std::int64_t cube_term(std::uint32_t count) {
    const std::int64_t x = count;
    return x * (x - 1) * (x - 1);
}

Restart higher_order_restart(const BitMatrix& m, std::size_t k, const SASelectorConfig& c,
                             Xs64& g) {
    std::vector<std::size_t> window, pool;
    partition(m.rows(), k, g, window, pool);
    const std::size_t words = m.words_per_row(), cols = m.cols();
    const std::uint64_t tail = cols % 64 ? (std::uint64_t{1} << (cols % 64)) - 1 : ~std::uint64_t{0};
    auto word = [&](std::size_t row, std::size_t w) {
        const std::uint64_t x = m.row_words(row)[w];
        return w + 1 == words ? x & tail : x;
    };
    std::vector<std::uint32_t> counts(cols, 0);
    for (std::size_t r : window)
        for (std::size_t w = 0; w < words; ++w)
            for (std::uint64_t x = word(r, w); x; x &= x - 1)
                ++counts[w * 64 + (std::size_t)__builtin_ctzll(x)];
    std::int64_t score = 0;
    for (std::uint32_t n : counts)
        score += cube_term(n);
    auto delta = [&](std::size_t iw, std::size_t ip) {
        const std::size_t out = window[iw], in = pool[ip];
        std::int64_t d = 0;
        for (std::size_t w = 0; w < words; ++w) {
            const std::uint64_t o = word(out, w), i = word(in, w);
            for (std::uint64_t x = o & ~i; x; x &= x - 1) {
                const std::uint32_t n = counts[w * 64 + (std::size_t)__builtin_ctzll(x)];
                d += cube_term(n - 1) - cube_term(n);
            }
            for (std::uint64_t x = i & ~o; x; x &= x - 1) {
                const std::uint32_t n = counts[w * 64 + (std::size_t)__builtin_ctzll(x)];
                d += cube_term(n + 1) - cube_term(n);
            }
        }
        return d;
    };
    auto commit = [&](std::size_t iw, std::size_t ip) {
        const std::size_t out = window[iw], in = pool[ip];
        for (std::size_t w = 0; w < words; ++w) {
            const std::uint64_t o = word(out, w), i = word(in, w);
            for (std::uint64_t x = o & ~i; x; x &= x - 1)
                --counts[w * 64 + (std::size_t)__builtin_ctzll(x)];
            for (std::uint64_t x = i & ~o; x; x &= x - 1)
                ++counts[w * 64 + (std::size_t)__builtin_ctzll(x)];
        }
    };
    return anneal(c, g, window, pool, score, delta, commit);
}

} // namespace

SASelectionResult SASelector::select(const BitMatrix& matrix, std::size_t window_size) const {
    const SASelectorConfig& c = _config;
    if (c.restarts == 0)
        throw std::invalid_argument("SASelector: restarts must be >= 1");
    if (!(c.initial_acceptance_probability > 0.0 && c.initial_acceptance_probability < 1.0))
        throw std::invalid_argument("SASelector: initial_acceptance_probability must be in (0, 1)");
    if (c.cooling_rate <= 0.0 || c.cooling_rate > 1.0)
        throw std::invalid_argument("SASelector: cooling_rate must be in (0, 1]");
    if (c.min_temperature <= 0.0)
        throw std::invalid_argument("SASelector: min_temperature must be > 0");
    if (window_size > matrix.rows())
        throw std::invalid_argument("SASelector: window_size exceeds matrix row count");
    SASelectionResult out;
    out.seed = c.seed == 0 ? 0x9E3779B97F4A7C15ull : c.seed;
    if (window_size == 0)
        return out;
    if (c.cooling_policy != CoolingPolicyKind::AdaptiveGeometric)
        throw std::invalid_argument("SASelector: unsupported cooling policy");
    if (c.score_policy != WindowScorePolicyKind::PairwiseSynergy &&
        c.score_policy != WindowScorePolicyKind::HigherOrderSynergy)
        throw std::invalid_argument("SASelector: unsupported score policy");
    std::vector<std::uint32_t> table;
    if (c.score_policy == WindowScorePolicyKind::PairwiseSynergy)
        table = synergy_table(matrix);
    Xs64 g(out.seed);
    bool has_best = false;
    for (std::size_t restart = 0; restart < c.restarts; ++restart) {
        const Restart r = c.score_policy == WindowScorePolicyKind::PairwiseSynergy
                              ? pairwise_restart(matrix.rows(), window_size, c, g, table)
                              : higher_order_restart(matrix, window_size, c, g);
        if (!has_best || better(r.score, r.rows, out.score, out.rows)) {
            out.rows = r.rows;
            out.score = r.score;
            out.accepted_moves = r.accepted;
            out.iterations_run = r.iterations_run;
            out.best_iteration = r.best_iteration;
            out.restart_index = restart;
            has_best = true;
        }
    }
    return out;
}

RowWindow SASelector::select_window(BitMatrix& matrix, std::size_t window_size) const {
    return matrix.row_window(select(matrix, window_size).rows);
}

RowWindow SASelector::select_window(const BitMatrix& matrix, std::size_t window_size) const {
    return matrix.row_window(select(matrix, window_size).rows);
}

} // namespace bmmpy
