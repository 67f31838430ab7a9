// Tunables of the gfx950 pipeline.  The reference's constants
// (include/opsparse/utils/config.hpp:6-33: 8 bins, 32-lane warps, hash
// multiplier 107) are re-derived here for 64-lane wavefronts and 160 KB of
// LDS; the per-bin table sizes and lane groups are listed in
// opsparse_amd/csrc/launch.hpp next to the kernels that use them.
#ifndef OPSPARSE_UTILS_CONFIG_HPP_
#define OPSPARSE_UTILS_CONFIG_HPP_

namespace opsparse {

inline constexpr int kNumBin = 8;
inline constexpr int kWaveSize = 64;
inline constexpr int kHashScale = 107;
// Upper bounds (inclusive) of row products per symbolic bin and of row nnz
// per numeric bin; the last bin is unbounded.
inline constexpr int kSymbolicBinBound[kNumBin - 1] = {16, 32, 64, 256, 1024, 4096, 16384};
inline constexpr int kNumericBinBound[kNumBin - 1] = {16, 32, 64, 128, 256, 1024, 4096};

template <typename Int>
inline constexpr Int div_up(Int a, Int b)
{
    return (a + b - 1) / b;
}

}  // namespace opsparse

#endif  // OPSPARSE_UTILS_CONFIG_HPP_
