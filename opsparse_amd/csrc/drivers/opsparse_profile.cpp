// Per-phase profile driver: C = A * B for MatrixMarket inputs, one warm-up
// call and the mean of 10 timed calls, printed with SpgemmTimings::print.
// Counterpart of the reference's benchmarks/opsparse_profile.cu:10-48
// (same arguments, loop and output).
//
// usage: opsparse_profile [A] [B]   (names or .mtx paths; default can_24)
#include <cstdio>
#include <string>

#include "bench_common.hpp"
#include "opsparse/core_algo/spgemm/spgemm.hpp"
#include "opsparse/numeric/compute_flop.hpp"
#include "opsparse/profile/spgemm_timings.hpp"
#include "opsparse/system/hip_common.hpp"

int main(int argc, char **argv)
{
    using namespace opsparse;
    std::string mat1, mat2;
    bench::parse_args(argc, argv, mat1, mat2);
    CSR A, B;
    bench::load_pair(mat1, mat2, A, B);
    A.H2D();
    B.H2D();
    const long total_flop = compute_flop(A, B);
    hip_runtime_warmup();
    SpgemmMeta meta;
    CSR C;
    {
        SpgemmTimings warm;
        opsparse_spgemm(A, B, C, meta, warm);
        C.release();
    }
    constexpr int kIters = 10;
    SpgemmTimings timing, mean;
    for (int i = 0; i < kIters; i++) {
        opsparse_spgemm(A, B, C, meta, timing);
        mean += timing;
        if (i + 1 < kIters) C.release();
    }
    mean /= kIters;
    std::printf("%s ", mat1.c_str());
    mean.print(2.0 * static_cast<double>(total_flop));
    A.release();
    B.release();
    C.release();
    return 0;
}
