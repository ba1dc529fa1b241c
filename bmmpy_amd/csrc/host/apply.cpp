// GreedyApplier (src/apply/greedy_applier.cpp:13-109) and row-XOR replay.
//
// The applier's choices depend only on candidate weights and the initial row
// popcounts, so the host decides the whole row-XOR sequence first and the GPU
// replays it over the window's rows in one launch.
#include "bmmpy.hpp"

#include "bmmpy_hip.h"

#include <algorithm>
#include <array>

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

} // namespace bmmpy
