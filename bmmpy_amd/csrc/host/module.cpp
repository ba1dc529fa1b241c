// pybind11 module `bmmpy_amd._bmmpy`: the reference's nanobind surface
// (python/bindings/*.cpp) for the row-XOR weight-minimization path, with the
// same class names, config structs, attribute names and exception types.
#include "bmmpy.hpp"

#include <pybind11/pybind11.h>
#include <pybind11/numpy.h>
#include <pybind11/stl.h>

#include <sstream>

namespace py = pybind11;
using namespace bmmpy;

namespace {

// The text format (src/core/bit_matrix_io.cpp:34-47), as the reference's
// to_text / __str__.
std::string matrix_to_text(const BitMatrix& m) {
    std::ostringstream out;
    m.save_text(out);
    return out.str();
}

BitMatrix matrix_from_text(const std::string& text) {
    std::istringstream in(text);
    return BitMatrix::load_text(in);
}

// str or os.PathLike -> filesystem path string.
std::string fspath(const py::object& p) {
    return py::module_::import("os").attr("fspath")(p).cast<std::string>();
}

std::vector<std::string> matrix_to_rows(const BitMatrix& m) {
    std::vector<std::string> out;
    for (std::size_t r = 0; r < m.rows(); ++r) {
        std::string s;
        for (std::size_t c = 0; c < m.cols(); ++c)
            s.push_back(m.get(r, c) ? '1' : '0');
        out.push_back(std::move(s));
    }
    return out;
}

BitMatrix matrix_from_rows(const std::vector<std::string>& rows) {
    const std::size_t cols = rows.empty() ? 0 : rows.front().size();
    BitMatrix m(rows.size(), cols);
    for (std::size_t r = 0; r < rows.size(); ++r) {
        if (rows[r].size() != cols)
            throw std::invalid_argument("BitMatrix.from_rows: all rows must have equal width");
        for (std::size_t c = 0; c < cols; ++c) {
            if (rows[r][c] == '1')
                m.set(r, c, true);
            else if (rows[r][c] != '0')
                throw std::invalid_argument(
                    "BitMatrix.from_rows: rows must contain only '0' and '1'");
        }
    }
    return m;
}

std::pair<std::size_t, std::size_t> index2(const py::tuple& t) {
    if (t.size() != 2)
        throw std::invalid_argument("BitMatrix index must be (row, col)");
    return {t[0].cast<std::size_t>(), t[1].cast<std::size_t>()};
}

// Packed words as a (rows, words_per_row) uint64 array (copy).
py::array_t<std::uint64_t> matrix_words(const BitMatrix& m) {
    py::array_t<std::uint64_t> a({m.rows(), m.words_per_row()});
    auto v = a.mutable_unchecked<2>();
    for (std::size_t r = 0; r < m.rows(); ++r)
        for (std::size_t w = 0; w < m.words_per_row(); ++w)
            v(r, w) = m.row_words(r)[w];
    return a;
}

BitMatrix matrix_from_words(py::array_t<std::uint64_t, py::array::c_style | py::array::forcecast> a,
                            std::size_t cols) {
    if (a.ndim() != 2)
        throw std::invalid_argument("from_words: expected a 2-D uint64 array");
    const std::size_t rows = a.shape(0), w = a.shape(1);
    if (w != (cols + 63) / 64)
        throw MatrixError(MatrixErr::DimensionMismatch);
    BitMatrix m(rows, cols);
    auto v = a.unchecked<2>();
    const std::uint64_t tail = cols % 64 ? ((1ull << (cols % 64)) - 1) : ~0ull;
    for (std::size_t r = 0; r < rows; ++r)
        for (std::size_t i = 0; i < w; ++i)
            m.row_words(r)[i] = i + 1 == w ? (v(r, i) & tail) : v(r, i);
    return m;
}

template <typename T> py::array_t<T> fwht_copy(py::array_t<T, py::array::c_style | py::array::forcecast> a) {
    py::array_t<T> out(a.size());
    std::copy(a.data(), a.data() + a.size(), out.mutable_data());
    fwht_inplace(out.mutable_data(), (std::size_t)out.size());
    return out;
}

std::string candidate_repr(const Candidate& c) {
    std::ostringstream s;
    s << "Candidate(weight=" << c.weight << ", rows=[";
    auto rows = c.selected_rows();
    for (std::size_t i = 0; i < rows.size(); ++i)
        s << (i ? ", " : "") << rows[i];
    s << "])";
    return s.str();
}

} // namespace

PYBIND11_MODULE(_bmmpy, m) {
    m.doc() = "bmmpy row-XOR weight minimization on MI355X (HIP kernels behind a C ABI)";

    static py::exception<MatrixError> matrix_error(m, "MatrixError", PyExc_RuntimeError);
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p)
                std::rethrow_exception(p);
        } catch (const MatrixError& e) {
            matrix_error(e.what());
        } catch (const std::logic_error& e) {
            // invalid_argument / out_of_range derive from logic_error and are
            // translated by pybind11's defaults; other logic errors surface
            // as RuntimeError like nanobind does.
            if (dynamic_cast<const std::invalid_argument*>(&e))
                throw;
            if (dynamic_cast<const std::out_of_range*>(&e))
                throw;
            if (dynamic_cast<const std::length_error*>(&e))
                throw;
            if (dynamic_cast<const std::domain_error*>(&e))
                throw;
            PyErr_SetString(PyExc_RuntimeError, e.what());
        }
    });

    py::enum_<MatrixErr>(m, "MatrixErr")
        .value("SizeTooLarge", MatrixErr::SizeTooLarge)
        .value("AllocationFailed", MatrixErr::AllocationFailed)
        .value("DimensionMismatch", MatrixErr::DimensionMismatch);

    py::class_<BitMatrix>(m, "BitMatrix")
        .def(py::init<>())
        .def(py::init<std::size_t, std::size_t>(), py::arg("rows"), py::arg("cols"))
        .def_property_readonly("rows", &BitMatrix::rows)
        .def_property_readonly("cols", &BitMatrix::cols)
        .def_property_readonly("shape",
                               [](const BitMatrix& b) { return py::make_tuple(b.rows(), b.cols()); })
        .def_property_readonly("stride_words", &BitMatrix::stride_words)
        .def_property_readonly("words_per_row", &BitMatrix::words_per_row)
        .def_property_readonly("total_words", &BitMatrix::total_words)
        .def_property_readonly("total_bytes", &BitMatrix::total_bytes)
        .def("__len__", &BitMatrix::rows)
        .def("__str__", &matrix_to_text)
        .def("__repr__",
             [](const BitMatrix& b) {
                 return "BitMatrix(rows=" + std::to_string(b.rows()) +
                        ", cols=" + std::to_string(b.cols()) +
                        ", weight=" + std::to_string(b.weight()) + ")";
             })
        .def("copy", [](const BitMatrix& b) { return BitMatrix(b); })
        .def(
            "copy_from_words",
            [](BitMatrix& b, const std::vector<std::uint64_t>& words) {
                b.copy_from_words(words.data(), words.size());
            },
            py::arg("words"))
        .def("add", &BitMatrix::add, py::arg("other"))
        .def("power", &BitMatrix::power, py::arg("exp"))
        .def("extract_rows_by_indices", &BitMatrix::extract_rows_by_indices, py::arg("indices"))
        .def("insert_rows_by_indices", &BitMatrix::insert_rows_by_indices, py::arg("source"),
             py::arg("indices"))
        .def(
            "save_text_file",
            [](const BitMatrix& b, const py::object& path) { b.save_text(fspath(path)); },
            py::arg("path"))
        .def(
            "save_binary_file",
            [](const BitMatrix& b, const py::object& path) { b.save_binary(fspath(path)); },
            py::arg("path"))
        .def("hash", &BitMatrix::hash)
        .def("__getitem__",
             [](const BitMatrix& b, const py::tuple& t) {
                 auto [r, c] = index2(t);
                 return b.get(r, c);
             })
        .def("__setitem__",
             [](BitMatrix& b, const py::tuple& t, bool v) {
                 auto [r, c] = index2(t);
                 b.set(r, c, v);
             })
        .def("get", &BitMatrix::get, py::arg("row"), py::arg("col"))
        .def("set", &BitMatrix::set, py::arg("row"), py::arg("col"), py::arg("value"))
        .def("row_xor", &BitMatrix::row_xor, py::arg("target_row"), py::arg("source_row"))
        .def("row_xor_from", &BitMatrix::row_xor_from, py::arg("target_row"), py::arg("source"),
             py::arg("source_row"))
        .def("mul", &BitMatrix::mul, py::arg("other"))
        .def("__matmul__", &BitMatrix::mul, py::arg("other"))
        .def("row_popcount", &BitMatrix::row_popcount, py::arg("row"))
        .def("weight", &BitMatrix::weight)
        .def("swap_rows", &BitMatrix::swap_rows, py::arg("r1"), py::arg("r2"))
        .def("rank", &BitMatrix::rank)
        .def("to_rows", &matrix_to_rows)
        .def("to_text", &matrix_to_text)
        .def("to_words", &matrix_words)
        .def(
            "row_window",
            [](BitMatrix& b, const std::vector<std::size_t>& rows) { return b.row_window(rows); },
            py::arg("rows"), py::keep_alive<0, 1>())
        .def_static("from_rows", &matrix_from_rows, py::arg("rows"))
        .def_static("from_words", &matrix_from_words, py::arg("words"), py::arg("cols"))
        .def_static("from_text", &matrix_from_text, py::arg("text"))
        .def_static(
            "load_text_file",
            [](const py::object& path) { return BitMatrix::load_text(fspath(path)); },
            py::arg("path"))
        .def_static(
            "load_binary_file",
            [](const py::object& path) { return BitMatrix::load_binary(fspath(path)); },
            py::arg("path"))
        .def_static("identity", &BitMatrix::identity, py::arg("n"));

    py::class_<RowWindow>(m, "RowWindow")
        .def(py::init([](BitMatrix& b, const std::vector<std::size_t>& rows) {
                 return RowWindow(b, rows);
             }),
             py::arg("matrix"), py::arg("rows"), py::keep_alive<1, 2>())
        .def("__len__", &RowWindow::size)
        .def_property_readonly("size", &RowWindow::size)
        .def_property_readonly("cols", &RowWindow::cols)
        .def_property_readonly("words_per_row", &RowWindow::words_per_row)
        .def_property_readonly("stride_words", &RowWindow::stride_words)
        .def_property_readonly("global_rows", &RowWindow::global_rows)
        .def_property_readonly("rows", &RowWindow::global_rows)
        .def("materialize",
             [](const RowWindow& w) {
                 BitMatrix out(w.size(), w.cols());
                 for (std::size_t i = 0; i < w.size(); ++i)
                     std::copy(w.row_words(i), w.row_words(i) + w.words_per_row(),
                               out.row_words(i));
                 return out;
             })
        .def_property_readonly("is_mutable", &RowWindow::is_mutable)
        .def("global_row", &RowWindow::global_row, py::arg("local_row"))
        .def("row_popcount", &RowWindow::row_popcount, py::arg("local_row"))
        .def("weight", &RowWindow::weight)
        .def_property_readonly("total_weight", &RowWindow::total_weight)
        .def_property_readonly("shape",
                               [](const RowWindow& w) { return py::make_tuple(w.size(), w.cols()); })
        .def("__repr__",
             [](const RowWindow& w) {
                 return "RowWindow(size=" + std::to_string(w.size()) +
                        ", cols=" + std::to_string(w.cols()) + ")";
             })
        .def("assign_materialized", &RowWindow::assign_materialized, py::arg("rows"))
        .def("get", &RowWindow::get, py::arg("local_row"), py::arg("col"))
        .def("set", &RowWindow::set, py::arg("local_row"), py::arg("col"), py::arg("value"))
        .def("row_xor", &RowWindow::row_xor, py::arg("target_local_row"),
             py::arg("source_local_row"));

    py::class_<Candidate>(m, "Candidate")
        .def(py::init<>())
        .def(py::init<Candidate::mask_type, std::uint32_t>(), py::arg("mask"), py::arg("weight"))
        .def("__len__", &Candidate::mask_popcount)
        .def_readwrite("mask", &Candidate::mask)
        .def_readwrite("weight", &Candidate::weight)
        .def("__repr__", &candidate_repr)
        .def("__contains__", &Candidate::has_row, py::arg("row"))
        .def("__eq__", [](const Candidate& a, const Candidate& b) {
            return a.weight == b.weight && a.mask == b.mask;
        })
        .def("has_row", &Candidate::has_row, py::arg("row"))
        .def_property_readonly("rows", &Candidate::selected_rows)
        .def("mask_popcount", &Candidate::mask_popcount)
        .def("mask_u64", &Candidate::mask_u64)
        .def("selected_rows", &Candidate::selected_rows)
        .def("__iter__",
             [](const Candidate& c) {
                 return py::iter(py::cast(c.selected_rows()));
             })
        .def_static("from_u64", &Candidate::from_u64, py::arg("mask"), py::arg("weight"))
        .def_static("from_words",
                    [](const Candidate::mask_type& mask, std::uint32_t weight) {
                        return Candidate(mask, weight);
                    },
                    py::arg("mask"), py::arg("weight"));

    py::class_<FwhtSearchConfig>(m, "FwhtSearchConfig")
        .def(py::init<>())
        .def_readwrite("max_rows", &FwhtSearchConfig::max_rows)
        .def_readwrite("max_candidates", &FwhtSearchConfig::max_candidates);
    py::class_<MitmFwhtSearchConfig>(m, "MitmFwhtSearchConfig")
        .def(py::init<>())
        .def_readwrite("reserve_unique_patterns", &MitmFwhtSearchConfig::reserve_unique_patterns)
        .def_readwrite("reserve_left_rows", &MitmFwhtSearchConfig::reserve_left_rows)
        .def_readwrite("reserve_right_states", &MitmFwhtSearchConfig::reserve_right_states)
        .def_readwrite("max_candidates", &MitmFwhtSearchConfig::max_candidates);
    py::class_<BruteforceSearchConfig>(m, "BruteforceSearchConfig")
        .def(py::init<>())
        .def_readwrite("max_candidates", &BruteforceSearchConfig::max_candidates)
        .def_readwrite("chunk_bits", &BruteforceSearchConfig::chunk_bits);
    py::class_<CudaMitmFwhtSearchConfig>(m, "CudaMitmFwhtSearchConfig")
        .def(py::init<>())
        .def_readwrite("max_candidates", &CudaMitmFwhtSearchConfig::max_candidates)
        .def_readwrite("low_bits", &CudaMitmFwhtSearchConfig::low_bits);
    py::class_<CudaBruteforceSearchConfig>(m, "CudaBruteforceSearchConfig")
        .def(py::init<>())
        .def_readwrite("max_candidates", &CudaBruteforceSearchConfig::max_candidates)
        .def_readwrite("chunk_bits", &CudaBruteforceSearchConfig::chunk_bits);

    py::class_<Searcher>(m, "Searcher")
        .def("search", &Searcher::search, py::arg("window"))
        .def("name", &Searcher::name);
    py::class_<FwhtSearch, Searcher>(m, "FwhtSearch")
        .def(py::init<FwhtSearchConfig>(), py::arg("config") = FwhtSearchConfig{});
    py::class_<MitmFwhtSearch, Searcher>(m, "MitmFwhtSearch")
        .def(py::init<MitmFwhtSearchConfig>(), py::arg("config") = MitmFwhtSearchConfig{});
    py::class_<BruteforceSearch, Searcher>(m, "BruteforceSearch")
        .def(py::init<BruteforceSearchConfig>(), py::arg("config") = BruteforceSearchConfig{});
    py::class_<CudaMitmFwhtSearch, Searcher>(m, "CudaMitmFwhtSearch")
        .def(py::init<CudaMitmFwhtSearchConfig>(), py::arg("config") = CudaMitmFwhtSearchConfig{});
    py::class_<CudaBruteforceSearch, Searcher>(m, "CudaBruteforceSearch")
        .def(py::init<CudaBruteforceSearchConfig>(),
             py::arg("config") = CudaBruteforceSearchConfig{});

    py::class_<CompactSplitWindow>(m, "CompactSplitWindow")
        .def_readonly("t", &CompactSplitWindow::t)
        .def_readonly("low_bits", &CompactSplitWindow::low_bits)
        .def_readonly("high_bits", &CompactSplitWindow::high_bits)
        .def_readonly("total_weight", &CompactSplitWindow::total_weight)
        .def_readonly("q", &CompactSplitWindow::q)
        .def_readonly("r", &CompactSplitWindow::r)
        .def_readonly("multiplicity", &CompactSplitWindow::multiplicity);
    m.def("build_compact_split_window", &build_compact_split_window, py::arg("window"),
          py::arg("low_bits"));

    py::class_<ApplyResult>(m, "ApplyResult")
        .def(py::init<>())
        .def_readwrite("applied_count", &ApplyResult::applied_count)
        .def_readwrite("weight_improvement", &ApplyResult::weight_improvement)
        .def("__repr__", [](const ApplyResult& r) {
            return "ApplyResult(applied_count=" + std::to_string(r.applied_count) +
                   ", weight_improvement=" + std::to_string(r.weight_improvement) + ")";
        });
    py::class_<GreedyApplier>(m, "GreedyApplier")
        .def(py::init<std::uint64_t, bool, std::uint64_t>(), py::arg("min_gain"),
             py::arg("stochastic") = false, py::arg("seed") = 0)
        .def("apply", &GreedyApplier::apply, py::arg("window"), py::arg("candidates"));
    py::class_<GlobalGreedyApplier>(m, "GlobalGreedyApplier")
        .def(py::init<bool>(), py::arg("require_improvement") = true)
        .def("name", &GlobalGreedyApplier::name)
        .def("apply", &GlobalGreedyApplier::apply, py::arg("window"), py::arg("candidates"));
    m.def("replay_row_xors", &replay_row_xors, py::arg("window"), py::arg("ops"));

    py::enum_<WindowScorePolicyKind>(m, "WindowScorePolicyKind")
        .value("PairwiseSynergy", WindowScorePolicyKind::PairwiseSynergy)
        .value("HigherOrderSynergy", WindowScorePolicyKind::HigherOrderSynergy);
    py::enum_<CoolingPolicyKind>(m, "CoolingPolicyKind")
        .value("AdaptiveGeometric", CoolingPolicyKind::AdaptiveGeometric);
    py::class_<SASelectorConfig>(m, "SASelectorConfig")
        .def(py::init<>())
        .def_readwrite("iterations", &SASelectorConfig::iterations)
        .def_readwrite("restarts", &SASelectorConfig::restarts)
        .def_readwrite("seed", &SASelectorConfig::seed)
        .def_readwrite("score_policy", &SASelectorConfig::score_policy)
        .def_readwrite("cooling_policy", &SASelectorConfig::cooling_policy)
        .def_readwrite("temperature_probe_samples", &SASelectorConfig::temperature_probe_samples)
        .def_readwrite("initial_acceptance_probability",
                       &SASelectorConfig::initial_acceptance_probability)
        .def_readwrite("cooling_rate", &SASelectorConfig::cooling_rate)
        .def_readwrite("min_temperature", &SASelectorConfig::min_temperature);
    py::class_<SASelectionResult>(m, "SASelectionResult")
        .def(py::init<>())
        .def_readwrite("rows", &SASelectionResult::rows)
        .def_readwrite("score", &SASelectionResult::score)
        .def_readwrite("accepted_moves", &SASelectionResult::accepted_moves)
        .def_readwrite("iterations_run", &SASelectionResult::iterations_run)
        .def_readwrite("best_iteration", &SASelectionResult::best_iteration)
        .def_readwrite("restart_index", &SASelectionResult::restart_index)
        .def_readwrite("seed", &SASelectionResult::seed);
    py::class_<SASelector>(m, "SASelector")
        .def(py::init<SASelectorConfig>(), py::arg("config") = SASelectorConfig{})
        .def("name", &SASelector::name)
        .def("select", &SASelector::select, py::arg("matrix"), py::arg("window_size"))
        .def(
            "select_window",
            [](const SASelector& s, BitMatrix& matrix, std::size_t window_size) {
                return s.select_window(matrix, window_size);
            },
            py::arg("matrix"), py::arg("window_size"), py::keep_alive<0, 2>());

    py::class_<ExactBasisSolverConfig>(m, "ExactBasisSolverConfig")
        .def(py::init<>())
        .def_readwrite("max_rows", &ExactBasisSolverConfig::max_rows)
        .def_readwrite("max_states", &ExactBasisSolverConfig::max_states)
        .def_readwrite("max_storage_bytes", &ExactBasisSolverConfig::max_storage_bytes);
    py::class_<ExactBasisResult>(m, "ExactBasisResult")
        .def(py::init<>())
        .def_readwrite("input_rows", &ExactBasisResult::input_rows)
        .def_readwrite("cols", &ExactBasisResult::cols)
        .def_readwrite("rank", &ExactBasisResult::rank)
        .def_readwrite("enumerated_states", &ExactBasisResult::enumerated_states)
        .def_readwrite("total_weight", &ExactBasisResult::total_weight)
        .def_readwrite("basis_masks", &ExactBasisResult::basis_masks)
        .def_readwrite("basis_weights", &ExactBasisResult::basis_weights)
        .def_readwrite("transform_matrix", &ExactBasisResult::transform_matrix)
        .def_readwrite("basis_matrix", &ExactBasisResult::basis_matrix);
    py::class_<ExactBasisSolver>(m, "ExactBasisSolver")
        .def(py::init<ExactBasisSolverConfig>(), py::arg("config") = ExactBasisSolverConfig{})
        .def("name", &ExactBasisSolver::name)
        .def("solve", &ExactBasisSolver::solve, py::arg("window"));

    py::class_<RuntimeFeatures>(m, "RuntimeFeatures")
        .def_readonly("avx2_compiled", &RuntimeFeatures::avx2_compiled)
        .def_readonly("avx2_supported", &RuntimeFeatures::avx2_supported)
        .def_readonly("avx2_available", &RuntimeFeatures::avx2_supported)
        .def_readonly("avx2_enabled", &RuntimeFeatures::avx2_enabled)
        .def_readonly("cuda_compiled", &RuntimeFeatures::cuda_compiled)
        .def_readonly("cuda_available", &RuntimeFeatures::cuda_available)
        .def_readonly("hip_compiled", &RuntimeFeatures::hip_compiled)
        .def_readonly("hip_available", &RuntimeFeatures::hip_available)
        .def_readonly("parallel_compiled", &RuntimeFeatures::parallel_compiled)
        .def_readonly("parallel_enabled", &RuntimeFeatures::parallel_enabled)
        .def_readonly("max_threads", &RuntimeFeatures::max_threads)
        .def_readonly("device_count", &RuntimeFeatures::device_count)
        .def_readonly("compute_units", &RuntimeFeatures::compute_units)
        .def_readonly("device_name", &RuntimeFeatures::device_name)
        .def_readonly("device_arch", &RuntimeFeatures::device_arch)
        .def_readonly("bit_ops_backend", &RuntimeFeatures::bit_ops_backend)
        .def_readonly("fwht_backend", &RuntimeFeatures::fwht_backend);
    m.def("get_runtime_features", &get_runtime_features);
    m.def("get_version", &get_version);
    m.def("add", [](int a, int b) { return a + b; }, py::arg("a"), py::arg("b"));
    m.def("fwht_i32", &fwht_copy<std::int32_t>, py::arg("values"));
    m.def("fwht_i16", &fwht_copy<std::int16_t>, py::arg("values"));
    // Lists, as the reference's bindings return (python/bindings/math.cpp:14-34).
    m.def(
        "fixed_weight_masks_u32",
        [](std::uint32_t n, std::uint32_t k) {
            std::vector<std::uint32_t> out;
            fixed_weight_masks_u32(n, k, out);
            return out;
        },
        py::arg("n"), py::arg("k"));
    m.def(
        "fixed_weight_masks_u64",
        [](std::uint32_t n, std::uint32_t k) {
            std::vector<std::uint64_t> out;
            fixed_weight_masks_u64(n, k, out);
            return out;
        },
        py::arg("n"), py::arg("k"));
}
