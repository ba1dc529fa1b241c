// C++ host API mirroring the reference's public types for the row-XOR
// weight-minimization path: BitMatrix / RowWindow (include/bmmpy/core),
// Candidate (include/bmmpy/types/candidate.hpp), the Searcher plugin
// interface and its searchers (include/bmmpy/search), the appliers
// (include/bmmpy/apply) and RuntimeFeatures (include/bmmpy/stub.hpp).
// Every search and replay runs on the GPU through the C ABI in bmmpy_hip.h;
// there is no CPU compute path behind these classes.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iosfwd>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace bmmpy {

// ------------------------------------------------------------- errors ----
// include/bmmpy/core/bit_matrix.hpp:14.
enum class MatrixErr { SizeTooLarge, AllocationFailed, DimensionMismatch };

// RuntimeError subclass raised for matrix shape problems
// (python/bindings/bit_matrix.cpp:94).
class MatrixError : public std::runtime_error {
public:
    explicit MatrixError(MatrixErr code);
    MatrixErr code() const noexcept { return _code; }

private:
    MatrixErr _code;
};

// Throws the reference exception type matching a BMM_* status.
void throw_status(int status);

// --------------------------------------------------------- BitMatrix -----
class RowWindow;

// Row-major packed GF(2) matrix: bit c of a row at word c/64, bit c%64; row
// stride padded to a multiple of 4 words; storage 32-byte aligned and zeroed
// (src/core/bit_matrix.cpp:41-72).
class BitMatrix {
public:
    static constexpr std::size_t k_word_bits = 64;
    static constexpr std::size_t k_alignment = 32;

    BitMatrix() noexcept = default;
    BitMatrix(std::size_t rows, std::size_t cols);
    ~BitMatrix() noexcept;
    BitMatrix(const BitMatrix& other);
    BitMatrix& operator=(const BitMatrix& other);
    BitMatrix(BitMatrix&& other) noexcept;
    BitMatrix& operator=(BitMatrix&& other) noexcept;

    std::size_t rows() const noexcept { return _rows; }
    std::size_t cols() const noexcept { return _cols; }
    std::size_t stride_words() const noexcept { return _stride; }
    std::size_t words_per_row() const noexcept { return (_cols + 63) / 64; }
    std::size_t total_words() const noexcept { return _rows * _stride; }
    const std::uint64_t* data() const noexcept { return _data; }
    std::uint64_t* data() noexcept { return _data; }

    void copy_from_words(const std::uint64_t* src, std::size_t count);
    const std::uint64_t* row_words(std::size_t row) const;
    std::uint64_t* row_words(std::size_t row);

    void set(std::size_t row, std::size_t col, bool value);
    bool get(std::size_t row, std::size_t col) const;

    // No-op when target == source (src/core/bit_matrix_rows.cpp:12-20).
    void row_xor(std::size_t target_row, std::size_t source_row);
    void row_xor_from(std::size_t target_row, const BitMatrix& source, std::size_t source_row);
    void swap_rows(std::size_t r1, std::size_t r2);

    std::uint64_t row_popcount(std::size_t row) const;
    std::uint64_t weight() const;
    std::size_t rank() const;
    std::size_t total_bytes() const noexcept { return total_words() * 8; }
    BitMatrix mul(const BitMatrix& other) const;
    // XOR of equal shapes (src/core/bit_matrix_algebra.cpp:13-25).
    void xor_assign(const BitMatrix& other);
    BitMatrix add(const BitMatrix& other) const;
    // Square-and-multiply over GPU products (bit_matrix_algebra.cpp:52-74).
    BitMatrix power(std::uint32_t exp) const;
    BitMatrix extract_rows_by_indices(const std::vector<std::size_t>& indices) const;
    void insert_rows_by_indices(const BitMatrix& source, const std::vector<std::size_t>& indices);
    static BitMatrix identity(std::size_t n);

    // Text and BMMPYB1 binary formats (src/core/bit_matrix_io.cpp) and the
    // canonical SHA-256 (src/core/bit_matrix_hash.cpp): hex digest of
    // le64(rows) | le64(cols) | every row's words little-endian, the last word
    // masked to the logical columns.
    void save_text(std::ostream& out) const;
    void save_binary(std::ostream& out) const;
    void save_text(const std::string& path) const;
    void save_binary(const std::string& path) const;
    static BitMatrix load_text(std::istream& in);
    static BitMatrix load_binary(std::istream& in);
    static BitMatrix load_text(const std::string& path);
    static BitMatrix load_binary(const std::string& path);
    std::string hash() const;

    RowWindow row_window(const std::vector<std::size_t>& rows);
    RowWindow row_window(const std::vector<std::size_t>& rows) const;

private:
    void allocate(std::size_t words);
    void release() noexcept;

    std::uint64_t* _data = nullptr;
    std::size_t _rows = 0;
    std::size_t _cols = 0;
    std::size_t _stride = 0;
};

// Non-owning ordered view of unique matrix rows; local row r = mask bit r
// (src/core/row_window.cpp:119-137).
class RowWindow {
public:
    RowWindow(BitMatrix& matrix, std::vector<std::size_t> rows);
    RowWindow(const BitMatrix& matrix, std::vector<std::size_t> rows);

    const BitMatrix& matrix() const noexcept { return *_matrix; }
    bool is_mutable() const noexcept { return _mutable != nullptr; }
    std::size_t size() const noexcept { return _rows.size(); }
    std::size_t cols() const noexcept { return _matrix->cols(); }
    std::size_t words_per_row() const noexcept { return _matrix->words_per_row(); }
    std::size_t stride_words() const noexcept { return _matrix->stride_words(); }
    const std::vector<std::size_t>& global_rows() const noexcept { return _rows; }
    std::size_t global_row(std::size_t local_row) const;
    std::vector<const std::uint64_t*> row_ptrs() const;
    const std::uint64_t* row_words(std::size_t local_row) const;
    std::uint64_t row_popcount(std::size_t local_row) const;
    std::uint64_t weight() const;
    // Columns set in at least one row of the window: the popcount of the OR
    // of its rows (src/core/row_window.cpp:67-88).
    std::uint64_t total_weight() const;
    // Writes rows.rows() == size() rows over the window's rows, in order
    // (src/core/row_window.cpp:109-115).
    void assign_materialized(const BitMatrix& rows);
    bool get(std::size_t local_row, std::size_t col) const;
    void set(std::size_t local_row, std::size_t col, bool value);
    void row_xor(std::size_t target_local_row, std::size_t source_local_row);

    // Rows packed contiguously (t x words_per_row, unpadded) for the C ABI.
    std::vector<std::uint64_t> packed_rows() const;
    // Writes t x words_per_row packed words back into the window's rows.
    void unpack_rows(const std::vector<std::uint64_t>& packed);

private:
    BitMatrix& require_mutable();
    const BitMatrix* _matrix;
    BitMatrix* _mutable;
    std::vector<std::size_t> _rows;
};

// ---------------------------------------------------------- Candidate ----
struct Candidate {
    using word_type = std::uint64_t;
    using mask_type = std::vector<word_type>;
    static constexpr std::size_t k_word_bits = 64;

    mask_type mask;
    std::uint32_t weight = 0;

    Candidate() = default;
    Candidate(mask_type m, std::uint32_t w) : mask(std::move(m)), weight(w) {}
    static Candidate from_u64(word_type m, std::uint32_t w) { return Candidate(mask_type{m}, w); }
    static Candidate make_unit(std::size_t row_count, std::size_t row, std::uint32_t w);

    bool has_row(std::size_t i) const noexcept;
    std::uint32_t mask_popcount() const noexcept;
    std::vector<std::size_t> selected_rows() const;
    // Throws std::logic_error if higher words are set (candidate.hpp:160-167).
    word_type mask_u64() const;
};

bool operator==(const Candidate& a, const Candidate& b);

// ----------------------------------------------------------- Searcher ----
class Searcher {
public:
    virtual ~Searcher() = default;
    virtual std::vector<Candidate> search(const RowWindow& window) = 0;
    virtual const char* name() const noexcept = 0;
};

// Upper bound of max_candidates served by the GPU top-k (larger k raises
// std::invalid_argument).
constexpr std::size_t k_gpu_topk_max = 4096;

struct FwhtSearchConfig {
    std::size_t max_rows = 16;
    std::size_t max_candidates = 64;
};

// src/search/fwht_search.cpp:34-132 on the GPU.
class FwhtSearch final : public Searcher {
public:
    explicit FwhtSearch(FwhtSearchConfig config = {}) : _config(config) {}
    std::vector<Candidate> search(const RowWindow& window) override;
    const char* name() const noexcept override { return "fwht"; }
    const FwhtSearchConfig& config() const noexcept { return _config; }

private:
    FwhtSearchConfig _config;
};

struct MitmFwhtSearchConfig {
    std::size_t reserve_unique_patterns = 1024;
    std::size_t reserve_left_rows = 20;
    std::size_t reserve_right_states = std::size_t{1} << 16;
    std::size_t max_candidates = 64;
};

// src/search/mitm_fwht_search.cpp:267-322 on the GPU (same limits: the
// reference split must keep low <= 30 and both halves < 31).
class MitmFwhtSearch final : public Searcher {
public:
    explicit MitmFwhtSearch(MitmFwhtSearchConfig config = {}) : _config(config) {}
    std::vector<Candidate> search(const RowWindow& window) override;
    const char* name() const noexcept override { return "mitm_fwht"; }
    const MitmFwhtSearchConfig& config() const noexcept { return _config; }

private:
    MitmFwhtSearchConfig _config;
};

struct BruteforceSearchConfig {
    std::size_t max_candidates = 128;
    std::size_t chunk_bits = 0;
};

// src/search/bruteforce_search.cpp:300-330 on the GPU (popcount over whole
// words, tail bits included).
class BruteforceSearch final : public Searcher {
public:
    explicit BruteforceSearch(BruteforceSearchConfig config = {}) : _config(config) {}
    std::vector<Candidate> search(const RowWindow& window) override;
    const char* name() const noexcept override { return "bruteforce"; }
    const BruteforceSearchConfig& config() const noexcept { return _config; }

private:
    BruteforceSearchConfig _config;
};

struct CudaMitmFwhtSearchConfig {
    std::size_t max_candidates = 64;
    std::size_t low_bits = 0;
};

// The reference's GPU split-FWHT searcher (src/search/cuda_mitm_fwht_search.cpp:
// 123-180), same validation limits, served by the HIP kernels.
class CudaMitmFwhtSearch final : public Searcher {
public:
    explicit CudaMitmFwhtSearch(CudaMitmFwhtSearchConfig config = {}) : _config(config) {}
    std::vector<Candidate> search(const RowWindow& window) override;
    const char* name() const noexcept override { return "cuda_mitm_fwht"; }
    const CudaMitmFwhtSearchConfig& config() const noexcept { return _config; }

private:
    static constexpr std::size_t k_min_rows = 16;
    static constexpr std::size_t k_max_rows = 64;
    CudaMitmFwhtSearchConfig _config;
};

struct CudaBruteforceSearchConfig {
    std::size_t max_candidates = 128;
    std::size_t chunk_bits = 0;
};

// The reference's GPU brute force (src/search/cuda_bruteforce_search.cpp:38-103),
// same validation limits, served by the HIP kernels.
class CudaBruteforceSearch final : public Searcher {
public:
    explicit CudaBruteforceSearch(CudaBruteforceSearchConfig config = {}) : _config(config) {}
    std::vector<Candidate> search(const RowWindow& window) override;
    const char* name() const noexcept override { return "cuda_bruteforce"; }
    const CudaBruteforceSearchConfig& config() const noexcept { return _config; }

private:
    static constexpr std::size_t k_min_rows = 16;
    static constexpr std::size_t k_max_rows = 64;
    CudaBruteforceSearchConfig _config;
};

// CompactSplitWindow (include/bmmpy/search/split_window_prep.hpp:10-20).
struct CompactSplitWindow {
    std::size_t t = 0;
    std::size_t low_bits = 0;
    std::size_t high_bits = 0;
    std::int32_t total_weight = 0;
    std::vector<std::uint64_t> q;
    std::vector<std::uint64_t> r;
    std::vector<std::int32_t> multiplicity;
};

// Host-side histogram of active-column patterns, sorted by pattern
// (src/search/split_window_prep.cpp:42-137).
CompactSplitWindow build_compact_split_window(const RowWindow& window, std::size_t low_bits);

// ------------------------------------------------------------ FWHT -------
void fwht_inplace(std::int32_t* data, std::size_t size);
void fwht_inplace(std::int16_t* data, std::size_t size);
// include/bmmpy/math/fwht.hpp:11-25 (bmm_hip_scores_order_i16 / _i32).
void calc_scores_and_order(const std::int16_t* h, std::size_t m, std::int32_t n,
                           std::int16_t* s_by_mask, std::int32_t* order, std::int32_t* cnt,
                           std::int32_t* off);
void calc_scores_and_order(const std::int32_t* h, std::size_t m, std::int32_t n,
                           std::int32_t* s_by_mask, std::int32_t* order, std::int32_t* cnt,
                           std::int32_t* off);

// fixed_weight_masks_u32 / _u64 (include/bmmpy/math/comb.hpp:7-12): out =
// every n-bit mask with k set bits, increasing (computed on the GPU).
void fixed_weight_masks_u32(std::uint32_t n, std::uint32_t k, std::vector<std::uint32_t>& out);
void fixed_weight_masks_u64(std::uint32_t n, std::uint32_t k, std::vector<std::uint64_t>& out);

// ---------------------------------------------------------- appliers -----
struct ApplyResult {
    std::size_t applied_count = 0;
    std::uint64_t weight_improvement = 0;
};

// src/apply/greedy_applier.cpp:13-109: the choice of target rows is made on
// the host from candidate weights; the resulting row-XOR sequence is replayed
// on the GPU.
class GreedyApplier {
public:
    GreedyApplier(std::uint64_t min_gain, bool stochastic = false, std::uint64_t seed = 0) noexcept
        : _min_gain(min_gain), _stochastic(stochastic),
          _prng_state(seed == 0 ? 0xDEADBULL : seed) {}
    ApplyResult apply(RowWindow& window, const std::vector<Candidate>& candidates);

private:
    std::uint64_t next_random() noexcept;
    std::uint64_t _min_gain;
    bool _stochastic;
    std::uint64_t _prng_state;
};

// src/apply/global_greedy_applier.cpp:59-105: the pool (candidates plus unit
// rows weighted by their popcounts, stably sorted by weight) and the pivot
// greedy run on the host over the mask words; the chosen square transform is
// applied to the window's rows on the GPU (bmm_hip_row_combine), which also
// returns the new weight.
class GlobalGreedyApplier {
public:
    explicit GlobalGreedyApplier(bool require_improvement = true) noexcept
        : _require_improvement(require_improvement) {}
    const char* name() const noexcept { return "global_greedy"; }
    bool require_improvement() const noexcept { return _require_improvement; }
    ApplyResult apply(RowWindow& window, const std::vector<Candidate>& candidates) const;

private:
    bool _require_improvement;
};

// Replays row-XOR ops (ops[2i] = source, ops[2i+1] = target, local rows) on the
// GPU; returns the window weight afterwards.
std::uint64_t replay_row_xors(RowWindow& window, const std::vector<std::uint16_t>& ops);

// ---------------------------------------------------------- SASelector ---
enum class WindowScorePolicyKind { PairwiseSynergy = 0, HigherOrderSynergy = 1 };
enum class CoolingPolicyKind { AdaptiveGeometric = 0 };

struct SASelectorConfig {
    std::size_t iterations = 10000;
    std::size_t restarts = 8;
    std::uint64_t seed = 0;
    WindowScorePolicyKind score_policy = WindowScorePolicyKind::PairwiseSynergy;
    CoolingPolicyKind cooling_policy = CoolingPolicyKind::AdaptiveGeometric;
    std::size_t temperature_probe_samples = 64;
    double initial_acceptance_probability = 0.8;
    double cooling_rate = 0.99;
    double min_temperature = 1e-6;
};

struct SASelectionResult {
    std::vector<std::size_t> rows;
    std::int64_t score = 0;
    std::size_t accepted_moves = 0;
    std::size_t iterations_run = 0;
    std::size_t best_iteration = 0;
    std::size_t restart_index = 0;
    std::uint64_t seed = 0;
};

// Simulated-annealing window selection (src/search/sa_selector.cpp:10-75):
// same configuration checks, restarts, RNG stream and acceptance rule.  The
// pairwise synergy table 2 popcount(r_i & r_j) is built once on the GPU
// (the reference rebuilds it per restart; it does not depend on the RNG);
// the annealing itself is sequential and runs on the host.
class SASelector {
public:
    explicit SASelector(SASelectorConfig config = {}) : _config(config) {}
    SASelectionResult select(const BitMatrix& matrix, std::size_t window_size) const;
    RowWindow select_window(BitMatrix& matrix, std::size_t window_size) const;
    RowWindow select_window(const BitMatrix& matrix, std::size_t window_size) const;
    const char* name() const noexcept { return "sa"; }
    const SASelectorConfig& config() const noexcept { return _config; }

private:
    SASelectorConfig _config;
};

// ------------------------------------------------------------ solver -----
struct ExactBasisSolverConfig {
    std::size_t max_rows = 24;
    std::uint64_t max_states = (std::uint64_t{1} << 24) - 1;
    std::size_t max_storage_bytes = 128u * 1024u * 1024u;
};

struct ExactBasisResult {
    std::size_t input_rows = 0;
    std::size_t cols = 0;
    std::size_t rank = 0;
    std::uint64_t enumerated_states = 0;
    std::uint64_t total_weight = 0;
    std::vector<std::uint32_t> basis_masks;
    std::vector<std::uint32_t> basis_weights;
    BitMatrix transform_matrix;
    BitMatrix basis_matrix;
};

// Minimum-total-weight basis of a window's row space
// (src/solver/exact_basis_solver.cpp:58-157): the stream of nonzero row
// combinations in (weight, Gray index) order comes from the GPU (one
// transform of the column-pattern histogram gives every weight); the greedy
// pivot-basis selection over that stream runs on the host.
class ExactBasisSolver {
public:
    explicit ExactBasisSolver(ExactBasisSolverConfig config = {}) : _config(config) {}
    const char* name() const noexcept { return "exact_basis"; }
    ExactBasisResult solve(const RowWindow& window) const;
    const ExactBasisSolverConfig& config() const noexcept { return _config; }

private:
    ExactBasisSolverConfig _config;
};

// ----------------------------------------------------------- runtime -----
struct RuntimeFeatures {
    bool avx2_compiled = false;
    bool avx2_supported = false;
    bool avx2_enabled = false;
    bool cuda_compiled = false;
    bool cuda_available = false;
    bool hip_compiled = true;
    bool hip_available = false;
    bool parallel_compiled = true;
    bool parallel_enabled = true;
    int max_threads = 1;
    int device_count = 0;
    int compute_units = 0;
    std::string device_name;
    std::string device_arch;
    std::string bit_ops_backend = "hip";
    std::string fwht_backend = "hip";
};

RuntimeFeatures get_runtime_features();
const char* get_version() noexcept;

} // namespace bmmpy
