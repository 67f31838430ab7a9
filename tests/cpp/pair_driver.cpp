// Host-only caller of the drop-in CSR helpers: bench::load_pair (the
// reference's benchmarks/bench_common.hpp:17-34) and the CSR sub-matrix
// constructor (csr.cu:110-147).  No HIP call is made.
//
// usage: pair_driver pair <A.mtx> <B.mtx> <out.bin>
//        pair_driver sub <A.mtx> <M_> <N_> <M_start> <N_start> <out.bin>
// Writes each resulting CSR as int32 M, N, nnz; rpt; col; double val.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "bench_common.hpp"

using namespace opsparse;

static void put(FILE *f, const CSR &X)
{
    const int hdr[3] = {X.M, X.N, X.nnz};
    std::fwrite(hdr, sizeof(int), 3, f);
    std::fwrite(X.rpt, sizeof(int), static_cast<size_t>(X.M) + 1, f);
    std::fwrite(X.col, sizeof(int), static_cast<size_t>(X.nnz), f);
    std::fwrite(X.val, sizeof(double), static_cast<size_t>(X.nnz), f);
}

int main(int argc, char **argv)
{
    if (argc == 5 && std::strcmp(argv[1], "pair") == 0) {
        CSR A, B;
        bench::load_pair(argv[2], argv[3], A, B);
        FILE *f = std::fopen(argv[4], "wb");
        if (!f) return 1;
        put(f, A);
        put(f, B);
        std::fclose(f);
        return 0;
    }
    if (argc == 8 && std::strcmp(argv[1], "sub") == 0) {
        CSR A(argv[2]);
        try {
            CSR S(A, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6]));
            FILE *f = std::fopen(argv[7], "wb");
            if (!f) return 1;
            put(f, S);
            std::fclose(f);
        } catch (const std::out_of_range &) {
            std::puts("out_of_range");
            return 3;
        }
        return 0;
    }
    std::fprintf(stderr, "usage: pair_driver pair A.mtx B.mtx out | sub A.mtx M N M0 N0 out\n");
    return 2;
}
