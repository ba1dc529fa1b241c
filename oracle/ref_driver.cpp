// Checker-only shim: a C entry layer over the reference's own CPU library, so
// that tests and fixture generation can call the reference search, FWHT and
// apply routines through ctypes.  Built by oracle/Makefile (target `ref`) from
// the reference sources where they lie; the output goes to oracle/_ref/ and is
// never shipped or called by the product path.
//
// Every function takes rows as `t` contiguous rows of `W` uint64 words
// (unpadded, LSB-first bit order, bit c at word c/64, bit c%64) plus `cols`.
// Return value: number of results written (>= 0), or -1 on a C++ exception
// whose message is copied into `err` (up to err_len bytes).

#include "bmmpy/apply/global_greedy_applier.hpp"
#include "bmmpy/apply/greedy_applier.hpp"
#include "bmmpy/core/bit_matrix.hpp"
#include "bmmpy/core/row_window.hpp"
#include "bmmpy/math/comb.hpp"
#include "bmmpy/math/fwht.hpp"
#include "bmmpy/search/bruteforce_search.hpp"
#include "bmmpy/search/fwht_search.hpp"
#include "bmmpy/search/mitm_fwht_search.hpp"
#include "bmmpy/search/sa_selector.hpp"
#include "bmmpy/search/split_window_prep.hpp"
#include "bmmpy/solver/exact_basis_solver.hpp"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <exception>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

namespace {

bmmpy::BitMatrix load(const std::uint64_t* rows, int t, int W, int cols) {
    bmmpy::BitMatrix m(static_cast<std::size_t>(t), static_cast<std::size_t>(cols));
    for (int r = 0; r < t; ++r) {
        std::uint64_t* dst = m.row_words(static_cast<std::size_t>(r));
        std::memcpy(dst, rows + static_cast<std::size_t>(r) * W, sizeof(std::uint64_t) * W);
    }
    return m;
}

void set_err(char* err, int err_len, const char* msg) {
    if (err == nullptr || err_len <= 0)
        return;
    std::strncpy(err, msg, static_cast<std::size_t>(err_len - 1));
    err[err_len - 1] = '\0';
}

std::vector<std::size_t> all_rows(int t) {
    std::vector<std::size_t> v(static_cast<std::size_t>(t));
    std::iota(v.begin(), v.end(), std::size_t{0});
    return v;
}

int emit(const std::vector<bmmpy::Candidate>& c, std::uint64_t* out_mask, std::uint32_t* out_w,
         std::size_t cap) {
    std::size_t n = c.size() < cap ? c.size() : cap;
    for (std::size_t i = 0; i < n; ++i) {
        out_mask[i] = c[i].mask.empty() ? 0 : c[i].mask.front();
        out_w[i] = c[i].weight;
    }
    return static_cast<int>(c.size());
}

} // namespace

extern "C" {

int ref_fwht_search(const std::uint64_t* rows, int t, int W, int cols, std::size_t max_rows,
                    std::size_t k, std::uint64_t* out_mask, std::uint32_t* out_w, std::size_t cap,
                    char* err, int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        auto win = m.row_window(all_rows(t));
        bmmpy::FwhtSearch s(bmmpy::FwhtSearchConfig{max_rows, k});
        return emit(s.search(win), out_mask, out_w, cap);
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

int ref_mitm_search(const std::uint64_t* rows, int t, int W, int cols, std::size_t k,
                    std::uint64_t* out_mask, std::uint32_t* out_w, std::size_t cap, char* err,
                    int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        auto win = m.row_window(all_rows(t));
        bmmpy::MitmFwhtSearch s(bmmpy::MitmFwhtSearchConfig{1024, 20, std::size_t{1} << 16, k});
        return emit(s.search(win), out_mask, out_w, cap);
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

int ref_bruteforce_search(const std::uint64_t* rows, int t, int W, int cols, std::size_t k,
                          std::size_t chunk_bits, std::uint64_t* out_mask, std::uint32_t* out_w,
                          std::size_t cap, char* err, int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        auto win = m.row_window(all_rows(t));
        bmmpy::BruteforceSearch s(bmmpy::BruteforceSearchConfig{k, chunk_bits});
        return emit(s.search(win), out_mask, out_w, cap);
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

int ref_fwht_i32(std::int32_t* data, std::size_t n, char* err, int err_len) {
    try {
        bmmpy::fwht_inplace(data, n);
        return 0;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

int ref_fwht_i16(std::int16_t* data, std::size_t n, char* err, int err_len) {
    try {
        bmmpy::fwht_inplace(data, n);
        return 0;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// Writes up to `cap` (q, r, multiplicity) entries; returns the unique count and
// stores total_weight / high_bits through the out pointers.
int ref_compact_split(const std::uint64_t* rows, int t, int W, int cols, int low_bits,
                      std::int32_t* total, int* high_bits, std::uint64_t* q, std::uint64_t* r,
                      std::int32_t* mult, std::size_t cap, char* err, int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        auto win = m.row_window(all_rows(t));
        auto p = bmmpy::build_compact_split_window(win, static_cast<std::size_t>(low_bits));
        *total = p.total_weight;
        *high_bits = static_cast<int>(p.high_bits);
        for (std::size_t i = 0; i < p.q.size() && i < cap; ++i) {
            q[i] = p.q[i];
            r[i] = p.r[i];
            mult[i] = p.multiplicity[i];
        }
        return static_cast<int>(p.q.size());
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// Replays row-XOR ops (dst ^= src; no-op when dst == src) through
// RowWindow::row_xor and returns the window weight; rows are updated in place.
long long ref_replay(std::uint64_t* rows, int t, int W, int cols, const std::uint16_t* ops,
                     std::size_t n_ops, char* err, int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        auto win = m.row_window(all_rows(t));
        for (std::size_t i = 0; i < n_ops; ++i)
            win.row_xor(ops[2 * i + 1], ops[2 * i]);
        for (int rr = 0; rr < t; ++rr)
            std::memcpy(rows + static_cast<std::size_t>(rr) * W, m.row_words(rr),
                        sizeof(std::uint64_t) * W);
        return static_cast<long long>(m.weight());
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// GreedyApplier::apply over candidates (mask, weight); rows updated in place.
int ref_greedy_apply(std::uint64_t* rows, int t, int W, int cols, const std::uint64_t* masks,
                     const std::uint32_t* weights, std::size_t n_cand, std::uint64_t min_gain,
                     int stochastic, std::uint64_t seed, std::uint64_t* out_applied,
                     std::uint64_t* out_improvement, char* err, int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        auto win = m.row_window(all_rows(t));
        std::vector<bmmpy::Candidate> cands;
        for (std::size_t i = 0; i < n_cand; ++i)
            cands.push_back(bmmpy::Candidate::from_u64(masks[i], weights[i]));
        bmmpy::GreedyApplier ap(min_gain, stochastic != 0, seed);
        auto res = ap.apply(win, cands);
        *out_applied = res.applied_count;
        *out_improvement = res.weight_improvement;
        for (int rr = 0; rr < t; ++rr)
            std::memcpy(rows + static_cast<std::size_t>(rr) * W, m.row_words(rr),
                        sizeof(std::uint64_t) * W);
        return 0;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// SASelector::select on a t-row matrix.  stats = {accepted_moves,
// iterations_run, best_iteration, restart_index, seed}; returns the number
// of selected rows written to out_rows.
int ref_sa_select(const std::uint64_t* rows, int t, int W, int cols, std::size_t window_size,
                  std::size_t iterations, std::size_t restarts, std::uint64_t seed, int policy,
                  std::size_t probe_samples, double p0, double cooling, double min_t,
                  std::uint64_t* out_rows, std::uint64_t* stats, long long* score, char* err,
                  int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        bmmpy::SASelectorConfig c;
        c.iterations = iterations;
        c.restarts = restarts;
        c.seed = seed;
        c.score_policy = policy ? bmmpy::WindowScorePolicyKind::HigherOrderSynergy
                                : bmmpy::WindowScorePolicyKind::PairwiseSynergy;
        c.temperature_probe_samples = probe_samples;
        c.initial_acceptance_probability = p0;
        c.cooling_rate = cooling;
        c.min_temperature = min_t;
        const auto r = bmmpy::SASelector(c).select(m, window_size);
        for (std::size_t i = 0; i < r.rows.size(); ++i)
            out_rows[i] = r.rows[i];
        stats[0] = r.accepted_moves;
        stats[1] = r.iterations_run;
        stats[2] = r.best_iteration;
        stats[3] = r.restart_index;
        stats[4] = r.seed;
        *score = r.score;
        return static_cast<int>(r.rows.size());
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// ExactBasisSolver (default caps) on all t rows: writes up to `cap` basis
// masks / weights in selection order, the total weight and the basis matrix
// rows (rank x W words); returns the rank.
int ref_exact_basis(const std::uint64_t* rows, int t, int W, int cols, std::uint32_t* out_masks,
                    std::uint32_t* out_w, std::size_t cap, std::uint64_t* total,
                    std::uint64_t* basis_rows, char* err, int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        auto win = m.row_window(all_rows(t));
        bmmpy::ExactBasisSolver s;
        const auto r = s.solve(win);
        const std::size_t n = std::min(cap, r.basis_masks.size());
        for (std::size_t i = 0; i < n; ++i) {
            out_masks[i] = r.basis_masks[i];
            out_w[i] = r.basis_weights[i];
            std::memcpy(basis_rows + i * static_cast<std::size_t>(W),
                        r.basis_matrix.row_words(i), sizeof(std::uint64_t) * W);
        }
        *total = r.total_weight;
        return static_cast<int>(r.rank);
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// calc_scores_and_order, both score types; returns 0.
int ref_scores_order_i32(const std::int32_t* h, std::size_t m, std::int32_t n, std::int32_t* s,
                         std::int32_t* order, std::int32_t* cnt, std::int32_t* off, char* err,
                         int err_len) {
    try {
        bmmpy::calc_scores_and_order(h, m, n, s, order, cnt, off);
        return 0;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

int ref_scores_order_i16(const std::int16_t* h, std::size_t m, std::int32_t n, std::int16_t* s,
                         std::int32_t* order, std::int32_t* cnt, std::int32_t* off, char* err,
                         int err_len) {
    try {
        bmmpy::calc_scores_and_order(h, m, n, s, order, cnt, off);
        return 0;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// fixed_weight_masks_u32 / _u64: up to capacity masks into out (as u64);
// returns the count.
long long ref_fixed_weight_masks(std::uint32_t n, std::uint32_t k, int width, std::uint64_t* out,
                                 std::size_t capacity, char* err, int err_len) {
    try {
        std::vector<std::uint64_t> v;
        if (width == 32) {
            std::vector<std::uint32_t> v32;
            bmmpy::fixed_weight_masks_u32(n, k, v32);
            v.assign(v32.begin(), v32.end());
        } else {
            bmmpy::fixed_weight_masks_u64(n, k, v);
        }
        for (std::size_t i = 0; i < v.size() && i < capacity; ++i)
            out[i] = v[i];
        return (long long)v.size();
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// GlobalGreedyApplier::apply over candidates of mask_words words each; rows
// updated in place.
int ref_global_greedy(std::uint64_t* rows, int t, int W, int cols, const std::uint64_t* masks,
                      int mask_words, const std::uint32_t* weights, std::size_t n_cand,
                      int require_improvement, std::uint64_t* out_applied,
                      std::uint64_t* out_improvement, char* err, int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        auto win = m.row_window(all_rows(t));
        std::vector<bmmpy::Candidate> cands;
        for (std::size_t i = 0; i < n_cand; ++i)
            cands.emplace_back(
                bmmpy::Candidate::mask_type(masks + i * mask_words, masks + (i + 1) * mask_words),
                weights[i]);
        const auto res = bmmpy::GlobalGreedyApplier(require_improvement != 0).apply(win, cands);
        *out_applied = res.applied_count;
        *out_improvement = res.weight_improvement;
        for (int rr = 0; rr < t; ++rr)
            std::memcpy(rows + static_cast<std::size_t>(rr) * W, m.row_words(rr),
                        sizeof(std::uint64_t) * W);
        return 0;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// BitMatrix::hash, save_binary and save_text of t rows (words as given, so
// stray bits past `cols` reach the writers as they would in the reference).
// Lengths are returned even when the buffers are too small (then nothing is
// copied); returns 0.
int ref_matrix_io(const std::uint64_t* rows, int t, int W, int cols, char* hex65,
                  unsigned char* bin, std::size_t bin_cap, std::size_t* bin_len, char* text,
                  std::size_t text_cap, std::size_t* text_len, char* err, int err_len) {
    try {
        auto m = load(rows, t, W, cols);
        const std::string h = m.hash();
        std::memcpy(hex65, h.c_str(), h.size() + 1);
        std::ostringstream b, x;
        m.save_binary(b);
        m.save_text(x);
        const std::string bs = b.str(), xs = x.str();
        *bin_len = bs.size();
        *text_len = xs.size();
        if (bs.size() <= bin_cap)
            std::memcpy(bin, bs.data(), bs.size());
        if (xs.size() <= text_cap)
            std::memcpy(text, xs.data(), xs.size());
        return 0;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// BitMatrix::load_binary / load_text (kind 0 / 1) from a byte buffer; writes
// rows x words_per_row words (up to cap) and the shape; returns 0.
int ref_matrix_load(const unsigned char* buf, std::size_t len, int kind, std::uint64_t* out,
                    std::size_t cap, std::uint64_t* shape, char* err, int err_len) {
    try {
        std::istringstream in(std::string(reinterpret_cast<const char*>(buf), len));
        const bmmpy::BitMatrix m =
            kind == 0 ? bmmpy::BitMatrix::load_binary(in) : bmmpy::BitMatrix::load_text(in);
        shape[0] = m.rows();
        shape[1] = m.cols();
        const std::size_t w = m.words_per_row();
        if (m.rows() * w <= cap)
            for (std::size_t r = 0; r < m.rows(); ++r)
                std::memcpy(out + r * w, m.row_words(r), sizeof(std::uint64_t) * w);
        return 0;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

// BitMatrix::mul and rank: out = a (ar x ac) * b (ac x bc), rows of aw / bw
// words; returns rank(a).
int ref_mul_rank(const std::uint64_t* a, int ar, int aw, int ac, const std::uint64_t* b, int bw,
                 int bc, std::uint64_t* out, char* err, int err_len) {
    try {
        auto A = load(a, ar, aw, ac);
        auto B = load(b, ac, bw, bc);
        const auto P = A.mul(B);
        for (int r = 0; r < ar; ++r)
            std::memcpy(out + static_cast<std::size_t>(r) * bw, P.row_words(r),
                        sizeof(std::uint64_t) * bw);
        return static_cast<int>(A.rank());
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return -1;
    }
}

} // extern "C"
