// BitMatrix text / BMMPYB1 binary formats (src/core/bit_matrix_io.cpp) and
// the canonical matrix hash (src/core/bit_matrix_hash.cpp:30-58).
//
// Binary layout: the 8 bytes "BMMPYB1\0", rows and cols as little-endian
// u64, then for each row its ceil(cols / 64) words as stored (little-endian),
// without the stride padding.  The hash is the SHA-256 hex digest of
// le64(rows) | le64(cols) | each row's words, the last word of a row masked to
// the logical columns, so padding and stray tail bits never change it.
//
// SHA-256 follows FIPS 180-4: the round constants and initial state are the
// first 32 fraction bits of the cube roots of the first 64 primes and of the
// square roots of the first 8 primes, derived below with exact integer roots.
#include "bmmpy.hpp"

#include <array>
#include <cstring>
#include <fstream>
#include <istream>
#include <limits>
#include <ostream>
#include <sstream>

namespace bmmpy {
namespace {

// ------------------------------------------------------------ SHA-256 --
using u128 = unsigned __int128;

// floor(x^(1/k)) for k = 2, 3 by bisection on 128-bit integers.
std::uint64_t iroot(u128 x, int k) {
    std::uint64_t lo = 0, hi = std::uint64_t(1) << (k == 2 ? 60 : 40);
    while (lo < hi) {
        const std::uint64_t mid = lo + (hi - lo + 1) / 2;
        u128 p = mid;
        for (int i = 1; i < k; ++i)
            p *= mid;
        if (p <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

struct ShaConstants {
    std::array<std::uint32_t, 64> k{};
    std::array<std::uint32_t, 8> h0{};
    ShaConstants() {
        int found = 0;
        for (std::uint32_t p = 2; found < 64; ++p) {
            bool prime = true;
            for (std::uint32_t d = 2; d * d <= p; ++d)
                if (p % d == 0) {
                    prime = false;
                    break;
                }
            if (!prime)
                continue;
            // frac(p^(1/3)) * 2^32 = floor((p * 2^96)^(1/3)) mod 2^32.
            k[found] = (std::uint32_t)iroot((u128)p << 96, 3);
            // frac(p^(1/2)) * 2^32 = floor((p * 2^64)^(1/2)) mod 2^32.
            if (found < 8)
                h0[found] = (std::uint32_t)iroot((u128)p << 64, 2);
            ++found;
        }
    }
};

const ShaConstants& sha_constants() {
    static const ShaConstants c;
    return c;
}

inline std::uint32_t rotr(std::uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

class Sha256 {
public:
    Sha256() : _h(sha_constants().h0) {}

    void update(const unsigned char* p, std::size_t n) {
        _len += n;
        while (n) {
            const std::size_t take = std::min<std::size_t>(64 - _fill, n);
            std::memcpy(_buf.data() + _fill, p, take);
            _fill += take;
            p += take;
            n -= take;
            if (_fill == 64) {
                block(_buf.data());
                _fill = 0;
            }
        }
    }

    void update_le64(std::uint64_t v) {
        unsigned char b[8];
        for (int i = 0; i < 8; ++i)
            b[i] = (unsigned char)(v >> (8 * i));
        update(b, 8);
    }

    std::string hex() {
        const std::uint64_t bits = _len * 8;
        const unsigned char one = 0x80, zero = 0;
        update(&one, 1);
        while (_fill != 56)
            update(&zero, 1);
        unsigned char be[8];
        for (int i = 0; i < 8; ++i)
            be[i] = (unsigned char)(bits >> (56 - 8 * i));
        update(be, 8);
        static const char* digits = "0123456789abcdef";
        std::string out;
        for (std::uint32_t w : _h)
            for (int i = 28; i >= 0; i -= 4)
                out.push_back(digits[(w >> i) & 15u]);
        return out;
    }

private:
    void block(const unsigned char* b) {
        const auto& K = sha_constants().k;
        std::uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = (std::uint32_t)b[4 * i] << 24 | (std::uint32_t)b[4 * i + 1] << 16 |
                   (std::uint32_t)b[4 * i + 2] << 8 | (std::uint32_t)b[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            const std::uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const std::uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        std::uint32_t a = _h[0], bb = _h[1], c = _h[2], d = _h[3], e = _h[4], f = _h[5],
                      g = _h[6], h = _h[7];
        for (int i = 0; i < 64; ++i) {
            const std::uint32_t t1 =
                h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + K[i] + w[i];
            const std::uint32_t t2 =
                (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & bb) ^ (a & c) ^ (bb & c));
            h = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = bb;
            bb = a;
            a = t1 + t2;
        }
        _h[0] += a;
        _h[1] += bb;
        _h[2] += c;
        _h[3] += d;
        _h[4] += e;
        _h[5] += f;
        _h[6] += g;
        _h[7] += h;
    }

    std::array<std::uint32_t, 8> _h;
    std::array<unsigned char, 64> _buf{};
    std::size_t _fill = 0;
    std::uint64_t _len = 0;
};

// ----------------------------------------------------------------- I/O --
constexpr char kMagic[8] = {'B', 'M', 'M', 'P', 'Y', 'B', '1', '\0'};

[[noreturn]] void io_error(const char* msg) { throw std::runtime_error(msg); }

void put_le64(std::ostream& out, std::uint64_t v) {
    char b[8];
    for (int i = 0; i < 8; ++i)
        b[i] = (char)(unsigned char)(v >> (8 * i));
    out.write(b, 8);
    if (!out)
        io_error("failed to write matrix binary data");
}

std::uint64_t get_le64(std::istream& in) {
    unsigned char b[8];
    in.read(reinterpret_cast<char*>(b), 8);
    if (!in)
        io_error("failed to read matrix binary data");
    std::uint64_t v = 0;
    for (int i = 0; i < 8; ++i)
        v |= (std::uint64_t)b[i] << (8 * i);
    return v;
}

bool is_space(char c) {
    return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == '\f' || c == '\v';
}

} // namespace

std::string BitMatrix::hash() const {
    Sha256 h;
    h.update_le64(_rows);
    h.update_le64(_cols);
    const std::size_t w = words_per_row();
    const std::uint64_t last = _cols % 64 ? (std::uint64_t(1) << (_cols % 64)) - 1 : ~0ull;
    for (std::size_t r = 0; r < _rows && w; ++r) {
        const std::uint64_t* row = row_words(r);
        for (std::size_t q = 0; q < w; ++q)
            h.update_le64(q + 1 == w ? row[q] & last : row[q]);
    }
    return h.hex();
}

void BitMatrix::save_text(std::ostream& out) const {
    out << _rows << ' ' << _cols << '\n';
    if (!out)
        io_error("failed to write matrix text header");
    std::string line(_cols + 1, '\n');
    for (std::size_t r = 0; r < _rows; ++r) {
        for (std::size_t c = 0; c < _cols; ++c)
            line[c] = get(r, c) ? '1' : '0';
        out.write(line.data(), (std::streamsize)line.size());
        if (!out)
            io_error("failed to write matrix text data");
    }
}

void BitMatrix::save_binary(std::ostream& out) const {
    out.write(kMagic, 8);
    if (!out)
        io_error("failed to write matrix binary header");
    put_le64(out, _rows);
    put_le64(out, _cols);
    const std::size_t w = words_per_row();
    for (std::size_t r = 0; r < _rows; ++r)
        for (std::size_t q = 0; q < w; ++q)
            put_le64(out, row_words(r)[q]);
}

void BitMatrix::save_text(const std::string& path) const {
    std::ofstream out(path);
    if (!out)
        io_error("failed to open matrix text file for writing");
    save_text(out);
}

void BitMatrix::save_binary(const std::string& path) const {
    std::ofstream out(path, std::ios::binary);
    if (!out)
        io_error("failed to open matrix binary file for writing");
    save_binary(out);
}

BitMatrix BitMatrix::load_text(std::istream& in) {
    std::size_t rows = 0, cols = 0;
    if (!(in >> rows >> cols))
        io_error("failed to read matrix text header");
    BitMatrix m(rows, cols);
    if (cols != 0 && rows > std::numeric_limits<std::size_t>::max() / cols)
        throw MatrixError(MatrixErr::SizeTooLarge);
    const std::size_t total = rows * cols;
    std::size_t i = 0;
    char ch = 0;
    while (i < total && in.get(ch)) {
        if (ch == '0' || ch == '1') {
            if (ch == '1')
                m.set(i / cols, i % cols, true);
            ++i;
        } else if (!is_space(ch)) {
            io_error("matrix text contains invalid character");
        }
    }
    if (i != total)
        io_error("failed to read matrix text payload");
    while (in.get(ch))
        if (!is_space(ch))
            io_error("matrix text contains trailing data");
    return m;
}

BitMatrix BitMatrix::load_binary(std::istream& in) {
    char magic[8];
    in.read(magic, 8);
    if (!in)
        io_error("failed to read matrix binary header");
    if (std::memcmp(magic, kMagic, 8) != 0)
        io_error("invalid matrix binary format");
    const std::uint64_t rows = get_le64(in), cols = get_le64(in);
    if (rows > std::numeric_limits<std::size_t>::max() ||
        cols > std::numeric_limits<std::size_t>::max())
        throw MatrixError(MatrixErr::SizeTooLarge);
    BitMatrix m((std::size_t)rows, (std::size_t)cols);
    const std::size_t w = m.words_per_row();
    std::vector<unsigned char> buf(w * 8);
    for (std::size_t r = 0; r < m.rows(); ++r) {
        in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)buf.size());
        if (!in)
            io_error("failed to read matrix binary payload");
        std::uint64_t* row = m.row_words(r);
        for (std::size_t q = 0; q < w; ++q) {
            std::uint64_t v = 0;
            for (int i = 0; i < 8; ++i)
                v |= (std::uint64_t)buf[q * 8 + i] << (8 * i);
            row[q] = v;
        }
    }
    return m;
}

BitMatrix BitMatrix::load_text(const std::string& path) {
    std::ifstream in(path);
    if (!in)
        io_error("failed to open matrix text file for reading");
    return load_text(in);
}

BitMatrix BitMatrix::load_binary(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in)
        io_error("failed to open matrix binary file for reading");
    return load_binary(in);
}

} // namespace bmmpy
