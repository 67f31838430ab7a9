// MatrixMarket coordinate banner: the drop-in for
// include/opsparse/io/matrix_market.hpp:30-65 (object / format / type /
// symmetry fields; parse errors throw std::runtime_error).
#ifndef OPSPARSE_IO_MATRIX_MARKET_HPP_
#define OPSPARSE_IO_MATRIX_MARKET_HPP_

#include <istream>
#include <string>
#include <vector>

namespace opsparse::io {

struct MatrixMarketBanner {
    std::string storage;   // "matrix"
    std::string format;    // "coordinate"
    std::string type;      // "real", "integer", "pattern", "complex"
    std::string symmetry;  // "general", "symmetric", "skew-symmetric", "hermitian"
};

// Splits on whitespace.
void tokenize(std::vector<std::string> &tokens, const std::string &line);
// Reads and validates the first line ("%%MatrixMarket matrix coordinate ...").
void read_mm_banner(std::istream &in, MatrixMarketBanner &banner);

}  // namespace opsparse::io

#endif  // OPSPARSE_IO_MATRIX_MARKET_HPP_
