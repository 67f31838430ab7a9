// Helpers of the benchmark drivers: the drop-in for the reference's
// benchmarks/bench_common.hpp:10-48 (resolve_matrix_path, load_pair,
// parse_args in namespace opsparse::bench), same argument meaning.
//
// resolve_matrix_path keeps the reference's layout for bare names
// (ER / G500 / SuiteSparse folders under ../matrix) and, as an addition,
// takes any name that ends in ".mtx" or contains a '/' as a path.
#ifndef OPSPARSE_AMD_DRIVERS_BENCH_COMMON_HPP_
#define OPSPARSE_AMD_DRIVERS_BENCH_COMMON_HPP_

#include <string>

#include "opsparse/core_ds/csr.hpp"

namespace opsparse::bench {

inline bool ends_with(const std::string &s, const std::string &tail)
{
    return s.size() >= tail.size() && s.compare(s.size() - tail.size(), tail.size(), tail) == 0;
}

inline std::string resolve_matrix_path(const std::string &name)
{
    if (ends_with(name, ".mtx") || name.find('/') != std::string::npos) return name;
    if (name.find("ER") != std::string::npos) return "../matrix/ER/" + name + ".mtx";
    if (name.find("G500") != std::string::npos) return "../matrix/G500/" + name + ".mtx";
    return "../matrix/suite_sparse/" + name + "/" + name + ".mtx";
}

// Loads A and B for C = A * B.  The same name gives B = A.  When the inner
// dimensions differ, the larger side is cut to the smaller one with the
// sub-matrix constructor: B keeps its first A.N rows, or A its first B.M
// columns (bench_common.hpp:17-34).
inline void load_pair(const std::string &mat1, const std::string &mat2, CSR &A, CSR &B)
{
    A.construct(resolve_matrix_path(mat1));
    if (mat1 == mat2) {
        B = A;
        return;
    }
    B.construct(resolve_matrix_path(mat2));
    if (A.N < B.M) {
        CSR cut(B, A.N, B.N, 0, 0);
        B = cut;
    } else if (A.N > B.M) {
        CSR cut(A, A.M, B.M, 0, 0);
        A = cut;
    }
}

// No argument: the reference's default matrix "can_24"; one: A = B; two: A, B.
inline void parse_args(int argc, char **argv, std::string &mat1, std::string &mat2)
{
    mat1 = mat2 = "can_24";
    if (argc >= 2) mat1 = mat2 = argv[1];
    if (argc >= 3) mat2 = argv[2];
}

}  // namespace opsparse::bench

#endif  // OPSPARSE_AMD_DRIVERS_BENCH_COMMON_HPP_
