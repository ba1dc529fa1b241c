// Searchers: the reference's argument handling and early returns, with every
// search executed on the GPU through the C ABI.
#include "bmmpy.hpp"

#include "bmmpy_hip.h"

#include <algorithm>
#include <climits>
#include <limits>

namespace bmmpy {

namespace {

std::uint64_t tail_mask(std::size_t cols) {
    const std::size_t tb = cols % 64;
    return tb == 0 ? ~0ull : ((1ull << tb) - 1);
}

bool device_available() {
    bmm_device_info info{};
    return bmm_hip_device_info(&info) == BMM_OK && info.device_count > 0;
}

std::vector<Candidate> to_candidates(const std::vector<bmm_topk_entry>& e, std::size_t n) {
    std::vector<Candidate> out;
    out.reserve(n);
    for (std::size_t i = 0; i < n; ++i)
        out.push_back(Candidate::from_u64(e[i].mask, e[i].weight));
    return out;
}

// Result capacity: min(k, 2^t - 1) records.
std::size_t out_capacity(std::size_t k, std::size_t t) {
    if (t >= 63)
        return k;
    return std::min<std::size_t>(k, (std::size_t{1} << t) - 1);
}

} // namespace

// --------------------------------------------------------------- FWHT ----
std::vector<Candidate> FwhtSearch::search(const RowWindow& window) {
    const std::size_t t = window.size();
    if (t == 0 || t > _config.max_rows || _config.max_candidates == 0)
        return {};
    if (t > Candidate::k_word_bits)
        throw std::invalid_argument("FwhtSearch: window_rows size must be <= 64");
    if (window.words_per_row() == 0)
        return {};
    if (window.cols() > (std::size_t)std::numeric_limits<std::int32_t>::max())
        throw std::overflow_error("FwhtSearch: total weight exceeds int32 range");
    const std::size_t k = _config.max_candidates;
    const auto rows = window.packed_rows();
    std::vector<bmm_topk_entry> out(out_capacity(k, t));
    std::size_t n = 0;
    throw_status(bmm_hip_fwht_topk(rows.data(), t, window.words_per_row(), window.cols(), k,
                                   out.data(), &n));
    return to_candidates(out, n);
}

// ------------------------------------------------------ split window ----
CompactSplitWindow build_compact_split_window(const RowWindow& window, std::size_t low_bits) {
    CompactSplitWindow out;
    out.t = window.size();
    out.low_bits = low_bits;
    if (low_bits > out.t)
        throw std::invalid_argument(
            "build_compact_split_window: low_bits must be <= number of rows");
    if (out.t > 64)
        throw std::invalid_argument("build_compact_split_window: row count must be <= 64");
    out.high_bits = out.t - low_bits;
    const std::size_t w = window.words_per_row();
    if (out.t == 0 || w == 0)
        return out;
    const auto rows = window.row_ptrs();
    std::vector<std::uint64_t> pats;
    for (std::size_t wi = 0; wi < w; ++wi) {
        std::uint64_t active = 0;
        for (const std::uint64_t* r : rows)
            active |= r[wi];
        if (wi + 1 == w)
            active &= tail_mask(window.cols());
        while (active) {
            const int bit = __builtin_ctzll(active);
            std::uint64_t p = 0;
            for (std::size_t r = 0; r < out.t; ++r)
                p |= ((rows[r][wi] >> bit) & 1ull) << r;
            pats.push_back(p);
            active &= active - 1;
        }
    }
    if (pats.size() > (std::size_t)std::numeric_limits<std::int32_t>::max())
        throw std::overflow_error("build_compact_split_window: total weight exceeds int32 range");
    std::sort(pats.begin(), pats.end());
    const std::uint64_t lowm = low_bits == 0 ? 0 : (low_bits >= 64 ? ~0ull : ((1ull << low_bits) - 1));
    for (std::size_t i = 0; i < pats.size();) {
        std::size_t j = i;
        while (j < pats.size() && pats[j] == pats[i])
            ++j;
        out.q.push_back(low_bits >= 64 ? 0 : (pats[i] >> low_bits));
        out.r.push_back(pats[i] & lowm);
        out.multiplicity.push_back((std::int32_t)(j - i));
        i = j;
    }
    out.total_weight = (std::int32_t)pats.size();
    return out;
}

namespace {
std::vector<Candidate> run_mitm(const CompactSplitWindow& prep, std::size_t k) {
    std::vector<bmm_topk_entry> out(out_capacity(k, prep.t));
    std::size_t n = 0;
    throw_status(bmm_hip_mitm_topk(prep.q.data(), prep.r.data(), prep.multiplicity.data(),
                                   prep.q.size(), prep.low_bits, prep.high_bits,
                                   prep.total_weight, k, out.data(), &n));
    return to_candidates(out, n);
}
} // namespace

// --------------------------------------------------------------- MITM ----
std::vector<Candidate> MitmFwhtSearch::search(const RowWindow& window) {
    if (_config.max_candidates == 0)
        return {};
    const std::size_t t = window.size();
    if (t == 0)
        return {};
    if (t > Candidate::k_word_bits)
        throw std::invalid_argument("MitmFwhtSearch: window_rows size must be <= 64");
    // Split rule and limits of src/search/mitm_fwht_search.cpp:56-60, 283-289.
    const std::size_t high = t <= 26 ? t / 2 : std::min<std::size_t>(t / 2, 13);
    const std::size_t low = t - high;
    if (low > 30)
        throw std::invalid_argument("MitmFwhtSearch: low half exceeds supported width");
    if (low >= 31 || high >= 31)
        throw std::invalid_argument("MitmFwhtSearch: split dimensions must be < 31");
    const CompactSplitWindow prep = build_compact_split_window(window, low);
    if (prep.multiplicity.empty())
        return {};
    return run_mitm(prep, _config.max_candidates);
}

// -------------------------------------------------------- brute force ----
std::vector<Candidate> BruteforceSearch::search(const RowWindow& window) {
    if (_config.max_candidates == 0)
        return {};
    const std::size_t t = window.size();
    if (t == 0)
        return {};
    if (t > Candidate::k_word_bits)
        throw std::invalid_argument("BruteforceSearch: window_rows size must be <= 64");
    if (window.words_per_row() == 0)
        return {};
    if (window.cols() > (std::size_t)std::numeric_limits<std::uint32_t>::max())
        throw std::overflow_error("BruteforceSearch: window width exceeds candidate weight range");
    if (_config.chunk_bits != 0 && (_config.chunk_bits > t || _config.chunk_bits >= 64))
        throw std::invalid_argument(
            "BruteforceSearch: chunk_bits must be in [1, min(t, 63)] or 0 for auto");
    const std::size_t k = _config.max_candidates;
    const auto rows = window.packed_rows();
    std::vector<bmm_topk_entry> out(out_capacity(k, t));
    std::size_t n = 0;
    throw_status(bmm_hip_bruteforce_topk(rows.data(), t, window.words_per_row(),
                                         _config.chunk_bits, k, out.data(), &n));
    return to_candidates(out, n);
}

// ------------------------------------------------- reference GPU names ---
namespace {
// resolve_low_bits of src/search/cuda_mitm_fwht_search.cpp:77-116 with the
// MI355X shape (16-bit LDS entries up to 2^15, 256 CUs x 8 blocks).
std::size_t resolve_low_bits(std::size_t t, std::size_t configured) {
    const std::size_t max_low = std::min<std::size_t>(t - 1, 15);
    if (configured != 0)
        return configured;
    const std::size_t target_blocks = 256 * 8;
    std::size_t min_high = 0;
    while ((std::size_t{1} << min_high) < target_blocks)
        ++min_high;
    std::size_t best = std::min<std::size_t>(max_low, 10);
    double best_cost = std::numeric_limits<double>::infinity();
    for (std::size_t low = 1; low <= max_low; ++low) {
        if (max_low >= 8 && low < 8)
            continue;
        const std::size_t high = t - low;
        double cost = 512.0 / (double)(std::size_t{1} << low) + (double)low;
        if (high < min_high)
            cost += 32.0 * (double)(min_high - high);
        if (cost < best_cost) {
            best_cost = cost;
            best = low;
        }
    }
    return best;
}
} // namespace

std::vector<Candidate> CudaMitmFwhtSearch::search(const RowWindow& window) {
    if (_config.max_candidates == 0)
        return {};
    const std::size_t t = window.size();
    if (t == 0)
        return {};
    if (t > Candidate::k_word_bits)
        throw std::invalid_argument(
            "CudaMitmFwhtSearch: too many rows in the window (must be <= 64)");
    if (t < k_min_rows || t > k_max_rows)
        throw std::invalid_argument(
            "CudaMitmFwhtSearch: window_rows size is outside the supported GPU range");
    const std::size_t low = resolve_low_bits(t, _config.low_bits);
    if (low == 0 || low >= t)
        throw std::invalid_argument(
            "CudaMitmFwhtSearch: low_bits must be in [1, t - 1] or 0 for auto");
    if (low > 15)
        throw std::invalid_argument(
            "CudaMitmFwhtSearch: low_bits exceeds the specialized GPU template range");
    if (window.words_per_row() == 0)
        return {};
    const CompactSplitWindow prep = build_compact_split_window(window, low);
    if (prep.multiplicity.empty())
        return {};
    if (prep.total_weight > 32767)
        throw std::invalid_argument(
            "CudaMitmFwhtSearch: total_weight must be <= 32767 for int16-safe GPU mode");
    if (!device_available())
        throw std::runtime_error("CudaMitmFwhtSearch: no HIP device is available");
    if (_config.max_candidates > 128)
        throw std::invalid_argument("run_cuda_mitm_fwht_search: max_candidates must be <= 128");
    return run_mitm(prep, _config.max_candidates);
}

std::vector<Candidate> CudaBruteforceSearch::search(const RowWindow& window) {
    if (_config.max_candidates == 0)
        return {};
    const std::size_t t = window.size();
    if (t == 0)
        return {};
    if (t > Candidate::k_word_bits)
        throw std::invalid_argument(
            "CudaBruteforceSearch: too many rows in the window (must be <= 64)");
    if (t < k_min_rows || t > k_max_rows)
        throw std::invalid_argument(
            "CudaBruteforceSearch: window_rows size is outside the supported GPU range");
    if (_config.max_candidates > 128)
        throw std::invalid_argument("CudaBruteforceSearch: max_candidates must be <= 128");
    const std::size_t w = window.words_per_row();
    if (w == 0)
        return {};
    if (w != 8 && w != 16 && w != 64)
        throw std::invalid_argument("CudaBruteforceSearch: only widths 512, 1024, and 4096 are "
                                    "supported in the current GPU path");
    if (_config.chunk_bits != 0 && (_config.chunk_bits >= t || _config.chunk_bits >= 64))
        throw std::invalid_argument(
            "CudaBruteforceSearch: chunk_bits must be in [1, t - 1] or 0 for auto");
    if (!device_available())
        throw std::runtime_error("CudaBruteforceSearch: no HIP device is available");
    const auto rows = window.packed_rows();
    std::vector<bmm_topk_entry> out(_config.max_candidates);
    std::size_t n = 0;
    throw_status(bmm_hip_bruteforce_topk(rows.data(), t, w, _config.chunk_bits,
                                         _config.max_candidates, out.data(), &n));
    return to_candidates(out, n);
}

} // namespace bmmpy
