// ExactBasisSolver (src/solver/exact_basis_solver.cpp:58-157) over the GPU
// candidate stream: the reference enumerates every nonzero mask in Gray
// order, counting-sorts by weight (stably) and inserts the masks' row
// combinations into a pivot basis in that order until the window's rank is
// reached.  bmm_hip_weight_order returns the first k masks of exactly that
// order; k grows (x8) only when the basis is not complete within them.
#include "bmmpy.hpp"

#include "bmmpy_hip.h"
#include "pivot.hpp"

#include <algorithm>
#include <stdexcept>

namespace bmmpy {
ExactBasisResult ExactBasisSolver::solve(const RowWindow& window) const {
    ExactBasisResult result;
    result.input_rows = window.size();
    result.cols = window.cols();
    const std::size_t rows = window.size();
    if (rows == 0) {
        result.transform_matrix = BitMatrix(0, 0);
        result.basis_matrix = BitMatrix(0, window.cols());
        return result;
    }
    if (rows > _config.max_rows)
        throw std::invalid_argument(
            "ExactBasisSolver: window has more rows than the configured max_rows cap");
    if (rows > 32)
        throw std::invalid_argument("ExactBasisSolver: this implementation supports at most 32 rows");
    const std::uint64_t states = (std::uint64_t{1} << rows) - 1;
    if (states > _config.max_states)
        throw std::invalid_argument(
            "ExactBasisSolver: state count exceeds the configured max_states cap");
    if (states * sizeof(std::uint32_t) > _config.max_storage_bytes)
        throw std::invalid_argument(
            "ExactBasisSolver: required packed-mask storage exceeds max_storage_bytes");

    const std::size_t words = window.words_per_row();
    BitMatrix materialized(rows, window.cols());
    for (std::size_t r = 0; r < rows; ++r)
        std::copy(window.row_words(r), window.row_words(r) + words, materialized.row_words(r));
    const std::size_t target_rank = materialized.rank();
    result.rank = target_rank;
    result.enumerated_states = states;
    if (target_rank == 0) {
        result.transform_matrix = BitMatrix(0, rows);
        result.basis_matrix = BitMatrix(0, window.cols());
        return result;
    }

    const std::vector<std::uint64_t> packed = window.packed_rows();
    std::uint64_t k = std::min<std::uint64_t>(states, std::max<std::uint64_t>(1024, 16 * rows));
    for (;;) {
        std::vector<std::uint32_t> masks((std::size_t)k), weights((std::size_t)k);
        std::size_t n = 0;
        throw_status(bmm_hip_weight_order(packed.data(), rows, words, window.cols(), (std::size_t)k,
                                          masks.data(), weights.data(), &n));
        PivotRows basis(words);
        std::vector<std::uint64_t> v(words);
        result.basis_masks.clear();
        result.basis_weights.clear();
        result.total_weight = 0;
        for (std::size_t q = 0; q < n && result.basis_masks.size() < target_rank; ++q) {
            std::fill(v.begin(), v.end(), 0);
            for (std::uint32_t m = masks[q]; m; m &= m - 1) {
                const std::uint64_t* src = window.row_words((std::size_t)__builtin_ctz(m));
                for (std::size_t w = 0; w < words; ++w)
                    v[w] ^= src[w];
            }
            if (!basis.try_insert(v))
                continue;
            result.basis_masks.push_back(masks[q]);
            result.basis_weights.push_back(weights[q]);
            result.total_weight += weights[q];
        }
        if (result.basis_masks.size() == target_rank)
            break;
        if (k == states)
            throw std::runtime_error(
                "ExactBasisSolver: failed to assemble a full basis from enumerated candidates");
        k = std::min<std::uint64_t>(states, k * 8);
    }

    BitMatrix transform(target_rank, rows);
    for (std::size_t q = 0; q < target_rank; ++q)
        for (std::uint32_t m = result.basis_masks[q]; m; m &= m - 1)
            transform.set(q, (std::size_t)__builtin_ctz(m), true);
    result.basis_matrix = transform.mul(materialized);
    result.transform_matrix = std::move(transform);
    return result;
}

} // namespace bmmpy
