// BitMatrix / RowWindow / Candidate (host data structures; layout as
// src/core/bit_matrix.cpp:41-72 and src/core/row_window.cpp).
#include "bmmpy.hpp"

#include "bmmpy_hip.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <unordered_set>

namespace bmmpy {

namespace {
const char* matrix_err_text(MatrixErr c) {
    // src/core/bit_matrix.cpp:10-21.
    switch (c) {
    case MatrixErr::SizeTooLarge: return "Matrix dimensions too large";
    case MatrixErr::AllocationFailed: return "Memory allocation failed";
    case MatrixErr::DimensionMismatch: return "Matrix dimensions mismatch";
    }
    return "Unknown matrix error";
}
} // namespace

MatrixError::MatrixError(MatrixErr code) : std::runtime_error(matrix_err_text(code)), _code(code) {}

void throw_status(int status) {
    if (status == BMM_OK)
        return;
    const std::string msg = bmm_hip_last_error();
    switch (status) {
    case BMM_ERR_INVALID: throw std::invalid_argument(msg);
    case BMM_ERR_OVERFLOW: throw std::overflow_error(msg);
    case BMM_ERR_RANGE: throw std::out_of_range(msg);
    default: throw std::runtime_error(msg);
    }
}

// ------------------------------------------------------------ BitMatrix --
BitMatrix::BitMatrix(std::size_t rows, std::size_t cols) : _rows(rows), _cols(cols) {
    const std::size_t words = (cols + 63) / 64;
    _stride = (words + 3) / 4 * 4;
    if (rows != 0 && _stride > SIZE_MAX / 8 / rows)
        throw MatrixError(MatrixErr::SizeTooLarge);
    allocate(_rows * _stride);
}

void BitMatrix::allocate(std::size_t words) {
    _data = nullptr;
    if (words == 0)
        return;
    void* p = std::aligned_alloc(k_alignment, (words * 8 + k_alignment - 1) / k_alignment *
                                                  k_alignment);
    if (!p)
        throw MatrixError(MatrixErr::AllocationFailed);
    std::memset(p, 0, words * 8);
    _data = static_cast<std::uint64_t*>(p);
}

void BitMatrix::release() noexcept {
    std::free(_data);
    _data = nullptr;
}

BitMatrix::~BitMatrix() noexcept { release(); }

BitMatrix::BitMatrix(const BitMatrix& o) : _rows(o._rows), _cols(o._cols), _stride(o._stride) {
    allocate(total_words());
    if (_data)
        std::memcpy(_data, o._data, total_words() * 8);
}

BitMatrix& BitMatrix::operator=(const BitMatrix& o) {
    if (this != &o) {
        BitMatrix tmp(o);
        *this = std::move(tmp);
    }
    return *this;
}

BitMatrix::BitMatrix(BitMatrix&& o) noexcept
    : _data(o._data), _rows(o._rows), _cols(o._cols), _stride(o._stride) {
    o._data = nullptr;
    o._rows = o._cols = o._stride = 0;
}

BitMatrix& BitMatrix::operator=(BitMatrix&& o) noexcept {
    if (this != &o) {
        release();
        _data = o._data;
        _rows = o._rows;
        _cols = o._cols;
        _stride = o._stride;
        o._data = nullptr;
        o._rows = o._cols = o._stride = 0;
    }
    return *this;
}

void BitMatrix::copy_from_words(const std::uint64_t* src, std::size_t count) {
    if (count != total_words())
        throw MatrixError(MatrixErr::DimensionMismatch);
    if (count)
        std::memcpy(_data, src, count * 8);
}

const std::uint64_t* BitMatrix::row_words(std::size_t row) const {
    if (row >= _rows)
        throw std::out_of_range("row out of bounds");
    return _data + row * _stride;
}

std::uint64_t* BitMatrix::row_words(std::size_t row) {
    if (row >= _rows)
        throw std::out_of_range("row out of bounds");
    return _data + row * _stride;
}

void BitMatrix::set(std::size_t row, std::size_t col, bool value) {
    if (row >= _rows || col >= _cols)
        throw std::out_of_range("index out of bounds");
    std::uint64_t& w = _data[row * _stride + col / 64];
    const std::uint64_t bit = 1ull << (col % 64);
    w = value ? (w | bit) : (w & ~bit);
}

bool BitMatrix::get(std::size_t row, std::size_t col) const {
    if (row >= _rows || col >= _cols)
        throw std::out_of_range("index out of bounds");
    return (_data[row * _stride + col / 64] >> (col % 64)) & 1u;
}

void BitMatrix::row_xor(std::size_t target, std::size_t source) {
    if (target >= _rows || source >= _rows)
        throw std::out_of_range("row out of bounds");
    if (target == source)
        return;
    std::uint64_t* d = _data + target * _stride;
    const std::uint64_t* s = _data + source * _stride;
    for (std::size_t w = 0; w < _stride; ++w)
        d[w] ^= s[w];
}

void BitMatrix::row_xor_from(std::size_t target, const BitMatrix& source, std::size_t src_row) {
    if (target >= _rows || src_row >= source._rows)
        throw std::out_of_range("row out of bounds");
    if (_cols != source._cols || _stride != source._stride)
        throw MatrixError(MatrixErr::DimensionMismatch);
    std::uint64_t* d = _data + target * _stride;
    const std::uint64_t* s = source._data + src_row * _stride;
    for (std::size_t w = 0; w < _stride; ++w)
        d[w] ^= s[w];
}

void BitMatrix::swap_rows(std::size_t r1, std::size_t r2) {
    if (r1 >= _rows || r2 >= _rows)
        throw std::out_of_range("row out of bounds");
    if (r1 == r2)
        return;
    std::swap_ranges(_data + r1 * _stride, _data + (r1 + 1) * _stride, _data + r2 * _stride);
}

std::uint64_t BitMatrix::row_popcount(std::size_t row) const {
    const std::uint64_t* r = row_words(row);
    std::uint64_t n = 0;
    for (std::size_t w = 0; w < words_per_row(); ++w)
        n += (std::uint64_t)__builtin_popcountll(r[w]);
    return n;
}

std::uint64_t BitMatrix::weight() const {
    std::uint64_t n = 0;
    for (std::size_t r = 0; r < _rows; ++r)
        n += row_popcount(r);
    return n;
}

std::size_t BitMatrix::rank() const {
    BitMatrix t(*this);
    std::size_t rank = 0;
    for (std::size_t c = 0; c < _cols && rank < _rows; ++c) {
        std::size_t p = rank;
        while (p < _rows && !t.get(p, c))
            ++p;
        if (p == _rows)
            continue;
        t.swap_rows(rank, p);
        for (std::size_t r = rank + 1; r < _rows; ++r)
            if (t.get(r, c))
                t.row_xor(r, rank);
        ++rank;
    }
    return rank;
}

// GF(2) product on the GPU (src/core/bit_matrix_algebra.cpp:27-50): row i of
// the result is the XOR of the rows of `other` selected by row i of this
// matrix, which bmm_hip_row_combine takes as its mask words.
BitMatrix BitMatrix::mul(const BitMatrix& other) const {
    if (_cols != other._rows)
        throw MatrixError(MatrixErr::DimensionMismatch);
    BitMatrix out(_rows, other._cols);
    const std::size_t mw = words_per_row(), ow = other.words_per_row();
    if (_rows == 0 || ow == 0)
        return out;
    std::vector<std::uint64_t> masks(_rows * mw), rows(other._rows * ow), prod(_rows * ow);
    for (std::size_t i = 0; i < _rows; ++i)
        std::memcpy(masks.data() + i * mw, row_words(i), mw * 8);
    for (std::size_t k = 0; k < other._rows; ++k)
        std::memcpy(rows.data() + k * ow, other.row_words(k), ow * 8);
    std::uint64_t wt = 0;
    throw_status(bmm_hip_row_combine(rows.data(), other._rows, ow, masks.data(), mw, _rows,
                                     prod.data(), &wt));
    for (std::size_t i = 0; i < _rows; ++i)
        std::memcpy(out.row_words(i), prod.data() + i * ow, ow * 8);
    return out;
}

void BitMatrix::xor_assign(const BitMatrix& other) {
    if (_rows != other._rows || _cols != other._cols)
        throw MatrixError(MatrixErr::DimensionMismatch);
    for (std::size_t r = 0; r < _rows; ++r)
        row_xor_from(r, other, r);
}

BitMatrix BitMatrix::add(const BitMatrix& other) const {
    BitMatrix out(*this);
    out.xor_assign(other);
    return out;
}

BitMatrix BitMatrix::power(std::uint32_t exp) const {
    if (_rows != _cols)
        throw MatrixError(MatrixErr::DimensionMismatch);
    BitMatrix result = identity(_rows), base(*this);
    while (exp) {
        if (exp & 1u)
            result = result.mul(base);
        exp >>= 1;
        if (exp)
            base = base.mul(base);
    }
    return result;
}

// src/core/bit_matrix_rows.cpp:97-131 (whole strides are copied).
BitMatrix BitMatrix::extract_rows_by_indices(const std::vector<std::size_t>& indices) const {
    BitMatrix out(indices.size(), _cols);
    for (std::size_t d = 0; d < indices.size(); ++d) {
        if (indices[d] >= _rows)
            throw std::out_of_range("source row out of bounds");
        std::memcpy(out._data + d * _stride, _data + indices[d] * _stride, _stride * 8);
    }
    return out;
}

void BitMatrix::insert_rows_by_indices(const BitMatrix& source,
                                       const std::vector<std::size_t>& indices) {
    if (source._rows != indices.size() || _cols != source._cols || _stride != source._stride)
        throw MatrixError(MatrixErr::DimensionMismatch);
    for (std::size_t s = 0; s < indices.size(); ++s) {
        if (indices[s] >= _rows)
            throw std::out_of_range("target row out of bounds");
        std::memcpy(_data + indices[s] * _stride, source._data + s * _stride, _stride * 8);
    }
}

BitMatrix BitMatrix::identity(std::size_t n) {
    BitMatrix m(n, n);
    for (std::size_t i = 0; i < n; ++i)
        m.set(i, i, true);
    return m;
}

RowWindow BitMatrix::row_window(const std::vector<std::size_t>& rows) {
    return RowWindow(*this, rows);
}

RowWindow BitMatrix::row_window(const std::vector<std::size_t>& rows) const {
    return RowWindow(*this, rows);
}

// ------------------------------------------------------------ RowWindow --
namespace {
// Uniqueness first, then bounds (src/core/row_window.cpp:119-137).
void check_rows(const BitMatrix& m, const std::vector<std::size_t>& rows) {
    std::unordered_set<std::size_t> seen;
    for (std::size_t r : rows)
        if (!seen.insert(r).second)
            throw std::invalid_argument("RowWindow rows must be unique");
    for (std::size_t r : rows)
        if (r >= m.rows())
            throw std::out_of_range("row window row out of bounds");
}
} // namespace

RowWindow::RowWindow(BitMatrix& matrix, std::vector<std::size_t> rows)
    : _matrix(&matrix), _mutable(&matrix), _rows(std::move(rows)) {
    check_rows(matrix, _rows);
}

RowWindow::RowWindow(const BitMatrix& matrix, std::vector<std::size_t> rows)
    : _matrix(&matrix), _mutable(nullptr), _rows(std::move(rows)) {
    check_rows(matrix, _rows);
}

std::size_t RowWindow::global_row(std::size_t local_row) const {
    if (local_row >= _rows.size())
        throw std::out_of_range("RowWindow: local row out of range");
    return _rows[local_row];
}

std::vector<const std::uint64_t*> RowWindow::row_ptrs() const {
    std::vector<const std::uint64_t*> p;
    p.reserve(_rows.size());
    for (std::size_t r : _rows)
        p.push_back(_matrix->row_words(r));
    return p;
}

const std::uint64_t* RowWindow::row_words(std::size_t local_row) const {
    return _matrix->row_words(global_row(local_row));
}

std::uint64_t RowWindow::row_popcount(std::size_t local_row) const {
    return _matrix->row_popcount(global_row(local_row));
}

std::uint64_t RowWindow::weight() const {
    std::uint64_t n = 0;
    for (std::size_t r : _rows)
        n += _matrix->row_popcount(r);
    return n;
}

std::uint64_t RowWindow::total_weight() const {
    const std::size_t W = words_per_row();
    const std::size_t tail = cols() % 64;
    std::uint64_t n = 0;
    for (std::size_t w = 0; w < W; ++w) {
        std::uint64_t any = 0;
        for (std::size_t r : _rows)
            any |= _matrix->row_words(r)[w];
        if (w + 1 == W && tail)
            any &= (1ull << tail) - 1;
        n += (std::uint64_t)__builtin_popcountll(any);
    }
    return n;
}

void RowWindow::assign_materialized(const BitMatrix& rows) {
    if (rows.rows() != _rows.size())
        throw MatrixError(MatrixErr::DimensionMismatch);
    require_mutable().insert_rows_by_indices(rows, _rows);
}

bool RowWindow::get(std::size_t local_row, std::size_t col) const {
    return _matrix->get(global_row(local_row), col);
}

BitMatrix& RowWindow::require_mutable() {
    if (!_mutable)
        throw std::logic_error("RowWindow: window over a const matrix is read-only");
    return *_mutable;
}

void RowWindow::set(std::size_t local_row, std::size_t col, bool value) {
    require_mutable().set(global_row(local_row), col, value);
}

void RowWindow::row_xor(std::size_t target, std::size_t source) {
    require_mutable().row_xor(global_row(target), global_row(source));
}

std::vector<std::uint64_t> RowWindow::packed_rows() const {
    const std::size_t w = words_per_row();
    std::vector<std::uint64_t> out(_rows.size() * w);
    for (std::size_t i = 0; i < _rows.size(); ++i)
        std::memcpy(out.data() + i * w, _matrix->row_words(_rows[i]), w * 8);
    return out;
}

void RowWindow::unpack_rows(const std::vector<std::uint64_t>& packed) {
    BitMatrix& m = require_mutable();
    const std::size_t w = words_per_row();
    if (packed.size() != _rows.size() * w)
        throw MatrixError(MatrixErr::DimensionMismatch);
    for (std::size_t i = 0; i < _rows.size(); ++i)
        std::memcpy(m.row_words(_rows[i]), packed.data() + i * w, w * 8);
}

// ------------------------------------------------------------ Candidate --
Candidate Candidate::make_unit(std::size_t row_count, std::size_t row, std::uint32_t w) {
    mask_type m((row_count + 63) / 64, 0);
    m[row / 64] |= 1ull << (row % 64);
    return Candidate(std::move(m), w);
}

bool Candidate::has_row(std::size_t i) const noexcept {
    const std::size_t wi = i / 64;
    return wi < mask.size() && ((mask[wi] >> (i % 64)) & 1u);
}

std::uint32_t Candidate::mask_popcount() const noexcept {
    std::uint32_t n = 0;
    for (word_type w : mask)
        n += (std::uint32_t)__builtin_popcountll(w);
    return n;
}

std::vector<std::size_t> Candidate::selected_rows() const {
    std::vector<std::size_t> out;
    for (std::size_t wi = 0; wi < mask.size(); ++wi)
        for (word_type w = mask[wi]; w; w &= w - 1)
            out.push_back(wi * 64 + (std::size_t)__builtin_ctzll(w));
    return out;
}

Candidate::word_type Candidate::mask_u64() const {
    for (std::size_t i = 1; i < mask.size(); ++i)
        if (mask[i] != 0)
            throw std::logic_error("Candidate::mask_u64: mask does not fit in u64");
    return mask.empty() ? 0 : mask.front();
}

bool operator==(const Candidate& a, const Candidate& b) {
    return a.weight == b.weight && a.mask_u64() == b.mask_u64();
}

} // namespace bmmpy
