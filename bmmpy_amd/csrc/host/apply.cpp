// GreedyApplier (src/apply/greedy_applier.cpp:13-109), row-XOR replay and
// GlobalGreedyApplier (src/apply/global_greedy_applier.cpp:15-105).
//
// GreedyApplier's choices depend only on candidate weights and the initial
// row popcounts, so the host decides the whole row-XOR sequence first and the
// GPU replays it over the window's rows in one launch.  GlobalGreedyApplier's
// choice depends only on the masks and weights as well; its transform is
// applied on the GPU as one row combination.
#include "bmmpy.hpp"

#include "bmmpy_hip.h"
#include "pivot.hpp"

#include <algorithm>
#include <array>
#include <limits>

namespace bmmpy {

std::uint64_t replay_row_xors(RowWindow& window, const std::vector<std::uint16_t>& ops) {
    if (!window.is_mutable())
        throw std::logic_error("RowWindow is not mutable");
    std::vector<std::uint64_t> rows = window.packed_rows();
    std::uint64_t wt = 0;
    throw_status(bmm_hip_replay(rows.data(), window.size(), window.words_per_row(), ops.data(),
                                ops.size() / 2, &wt));
    window.unpack_rows(rows);
    return wt;
}

std::uint64_t GreedyApplier::next_random() noexcept {
    std::uint64_t x = _prng_state;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    _prng_state = x;
    return x;
}

ApplyResult GreedyApplier::apply(RowWindow& window, const std::vector<Candidate>& candidates) {
    const std::size_t n = window.size();
    if (n > 64)
        throw std::invalid_argument("row window size must be <= 64");
    std::array<std::uint64_t, 64> c{};
    for (std::size_t i = 0; i < n; ++i)
        c[i] = 1ull << i;
    std::vector<std::uint64_t> cur(n);
    for (std::size_t i = 0; i < n; ++i)
        cur[i] = window.row_popcount(i);

    std::vector<const Candidate*> order;
    order.reserve(candidates.size());
    for (const Candidate& cand : candidates)
        order.push_back(&cand);
    if (_stochastic)
        for (std::size_t i = order.size(); i > 1; --i)
            std::swap(order[i - 1], order[(std::size_t)(next_random() % i)]);
    std::stable_sort(order.begin(), order.end(),
                     [](const Candidate* a, const Candidate* b) { return a->weight < b->weight; });

    ApplyResult result;
    std::vector<std::uint16_t> ops;
    for (const Candidate* cand : order) {
        const std::uint64_t m = cand->mask_u64();
        std::uint64_t a = 0;
        for (std::size_t i = 0; i < n; ++i)
            if (__builtin_popcountll(m & c[i]) & 1)
                a |= 1ull << i;
        if (a == 0)
            continue;
        std::array<std::size_t, 64> best{};
        std::size_t nbest = 0;
        std::uint64_t best_gain = 0;
        for (std::size_t r = 0; r < n; ++r) {
            if (!((a >> r) & 1u))
                continue;
            const std::uint64_t old = cur[r], w = cand->weight;
            if (old <= w)
                continue;
            const std::uint64_t gain = old - w;
            if (gain < _min_gain)
                continue;
            if (gain > best_gain) {
                best_gain = gain;
                best[0] = r;
                nbest = 1;
            } else if (gain == best_gain) {
                best[nbest++] = r;
            }
        }
        if (nbest == 0)
            continue;
        const std::size_t chosen = _stochastic ? best[(std::size_t)(next_random() % nbest)] : best[0];
        for (std::size_t i = 0; i < n; ++i)
            if (i != chosen && ((a >> i) & 1u)) {
                ops.push_back((std::uint16_t)i);
                ops.push_back((std::uint16_t)chosen);
            }
        cur[chosen] = cand->weight;
        for (std::size_t j = 0; j < n; ++j)
            if (j != chosen && ((a >> j) & 1u))
                c[j] ^= c[chosen];
        ++result.applied_count;
        result.weight_improvement += best_gain;
    }
    if (!ops.empty())
        replay_row_xors(window, ops);
    return result;
}

namespace {

struct PoolEntry {
    std::vector<std::uint64_t> mask;
    std::uint32_t weight;
};

// Candidate mask words widened or trimmed to `words`; a set bit at or beyond
// row `bits` is rejected (include/bmmpy/core/detail/xor_basis.hpp:73-110).
std::vector<std::uint64_t> normalized_mask(const std::vector<std::uint64_t>& mask,
                                           std::size_t bits, std::size_t words) {
    for (std::size_t w = 0; w < mask.size(); ++w) {
        const std::uint64_t allowed =
            w * 64 >= bits ? 0ull : (bits - w * 64 >= 64 ? ~0ull : (1ull << (bits - w * 64)) - 1);
        if (mask[w] & ~allowed)
            throw std::invalid_argument("mask contains bits outside the supported row range");
    }
    std::vector<std::uint64_t> out(words, 0);
    std::copy_n(mask.begin(), std::min(words, mask.size()), out.begin());
    return out;
}

} // namespace

ApplyResult GlobalGreedyApplier::apply(RowWindow& window,
                                       const std::vector<Candidate>& candidates) const {
    ApplyResult result;
    const std::size_t t = window.size();
    if (t == 0)
        return result;
    const std::size_t mw = (t + 63) / 64;

    // Pool: candidates in input order, then the unit rows; stable by weight
    // (global_greedy_applier.cpp:31-54).
    std::vector<PoolEntry> pool;
    pool.reserve(candidates.size() + t);
    for (const Candidate& c : candidates)
        pool.push_back(PoolEntry{normalized_mask(c.mask, t, mw), c.weight});
    for (std::size_t r = 0; r < t; ++r) {
        const std::uint64_t pc = window.row_popcount(r);
        if (pc > std::numeric_limits<std::uint32_t>::max())
            throw std::invalid_argument(
                "GlobalGreedyApplier: row weight exceeds supported u32 candidate range");
        std::vector<std::uint64_t> unit(mw, 0);
        unit[r / 64] = 1ull << (r % 64);
        pool.push_back(PoolEntry{std::move(unit), (std::uint32_t)pc});
    }
    std::stable_sort(pool.begin(), pool.end(),
                     [](const PoolEntry& a, const PoolEntry& b) { return a.weight < b.weight; });

    // Greedy basis in coefficient space (global_greedy_applier.cpp:72-86).
    PivotRows basis(mw);
    std::vector<std::uint64_t> chosen;   // t x mw mask words, in pick order
    chosen.reserve(t * mw);
    std::size_t n_chosen = 0;
    for (const PoolEntry& e : pool) {
        if (!basis.try_insert(e.mask))
            continue;
        chosen.insert(chosen.end(), e.mask.begin(), e.mask.end());
        if (++n_chosen == t)
            break;
    }
    if (n_chosen != t)
        throw std::runtime_error(
            "GlobalGreedyApplier: failed to assemble a full invertible transform");

    // after = T * before on the GPU (global_greedy_applier.cpp:88-96).
    const std::vector<std::uint64_t> before = window.packed_rows();
    const std::size_t words = window.words_per_row();
    const std::uint64_t old_weight = window.weight();
    std::vector<std::uint64_t> after(t * words);
    std::uint64_t new_weight = 0;
    throw_status(bmm_hip_row_combine(before.data(), t, words, chosen.data(), mw, t, after.data(),
                                     &new_weight));
    if (new_weight > old_weight || (_require_improvement && new_weight == old_weight))
        return result;
    window.unpack_rows(after);
    result.applied_count = 1;
    result.weight_improvement = old_weight - new_weight;
    return result;
}

} // namespace bmmpy
