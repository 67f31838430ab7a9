// MatrixMarket banner: the drop-in for include/opsparse/io/matrix_market.hpp:11-65
// (struct MatrixMarketBanner with fields matrix / storage / type / symmetry,
// tokenize(tokens, str, delimiters) and a template read_mm_banner over any
// stream with std::getline).  Header-only, as the reference's is.  Errors
// throw std::runtime_error.  Supported: "matrix coordinate" banners with
// real / integer / pattern / complex values and general / symmetric /
// skew-symmetric symmetry (hermitian, "vector" and "array" are rejected, as
// the reference rejects them).
#ifndef OPSPARSE_IO_MATRIX_MARKET_HPP_
#define OPSPARSE_IO_MATRIX_MARKET_HPP_

#include <cctype>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <vector>

namespace opsparse::io {

struct MatrixMarketBanner {
    std::string matrix;    // object: "matrix" ("vector" is rejected)
    std::string storage;   // "coordinate" ("array" is rejected)
    std::string type;      // "real", "integer", "pattern" or "complex"
    std::string symmetry;  // "general", "symmetric" or "skew-symmetric" (lower-cased)
};

// Replaces `tokens` with the maximal runs of `str` that contain none of
// the characters in `delimiters`.
inline void tokenize(std::vector<std::string> &tokens, const std::string &str,
                     const std::string &delimiters = "\n\r\t ")
{
    tokens.clear();
    std::string::size_type i = 0;
    const std::string::size_type n = str.size();
    while (i < n) {
        const std::string::size_type b = str.find_first_not_of(delimiters, i);
        if (b == std::string::npos) break;
        std::string::size_type e = str.find_first_of(delimiters, b);
        if (e == std::string::npos) e = n;
        tokens.emplace_back(str, b, e - b);
        i = e;
    }
}

namespace detail {
inline bool one_of(const std::string &v, std::initializer_list<const char *> set)
{
    for (const char *s : set)
        if (v == s) return true;
    return false;
}
}  // namespace detail

// Reads the first line of `input` ("%%MatrixMarket matrix coordinate
// <type> <symmetry>") into `banner`, validating every field.
template <typename Stream>
inline void read_mm_banner(Stream &input, MatrixMarketBanner &banner)
{
    std::string line;
    std::getline(input, line);
    std::vector<std::string> tok;
    tokenize(tok, line);
    if (tok.size() != 5 || tok[0] != "%%MatrixMarket") throw std::runtime_error("MatrixMarket banner not found");
    banner.matrix = tok[1];
    banner.storage = tok[2];
    banner.type = tok[3];
    banner.symmetry = tok[4];
    // Keywords are case-insensitive in the format; the symmetry is kept
    // lower-cased because callers compare it.
    for (char &ch : banner.symmetry) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    if (banner.matrix != "matrix") throw std::runtime_error("MatrixMarket: unsupported object " + banner.matrix);
    if (banner.storage != "coordinate")
        throw std::runtime_error("MatrixMarket: unsupported storage format " + banner.storage);
    if (!detail::one_of(banner.type, {"real", "integer", "pattern", "complex"}))
        throw std::runtime_error("MatrixMarket: invalid data type " + banner.type);
    if (!detail::one_of(banner.symmetry, {"general", "symmetric", "skew-symmetric"}))
        throw std::runtime_error("MatrixMarket: unsupported symmetry " + banner.symmetry);
}

}  // namespace opsparse::io

#endif  // OPSPARSE_IO_MATRIX_MARKET_HPP_
