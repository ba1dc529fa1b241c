// Incremental GF(2) independence test over multi-word vectors, keyed by the
// highest set bit (the reference's detail::PivotBasis,
// include/bmmpy/core/detail/xor_basis.hpp:123-160): a vector reduced to zero
// by the kept rows is dependent.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace bmmpy {

class PivotRows {
public:
    explicit PivotRows(std::size_t words) : _words(words) {}

    bool try_insert(std::vector<std::uint64_t> v) {
        for (;;) {
            const int hb = highest_bit(v);
            if (hb < 0)
                return false;
            auto it = std::lower_bound(_pivots.begin(), _pivots.end(), hb);
            if (it == _pivots.end() || *it != hb) {
                const std::size_t pos = (std::size_t)(it - _pivots.begin());
                _pivots.insert(it, hb);
                _rows.insert(_rows.begin() + (std::ptrdiff_t)pos, std::move(v));
                return true;
            }
            const auto& r = _rows[(std::size_t)(it - _pivots.begin())];
            for (std::size_t w = 0; w < _words; ++w)
                v[w] ^= r[w];
        }
    }

private:
    int highest_bit(const std::vector<std::uint64_t>& v) const {
        for (std::size_t w = _words; w > 0; --w)
            if (v[w - 1])
                return (int)((w - 1) * 64 + 63 - (std::size_t)__builtin_clzll(v[w - 1]));
        return -1;
    }

    std::size_t _words;
    std::vector<int> _pivots;
    std::vector<std::vector<std::uint64_t>> _rows;
};

} // namespace bmmpy
