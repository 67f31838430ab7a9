// Reference-style caller of the C++ drop-in API, written the way the
// reference's benchmarks/opsparse_profile.cu drives OpSparse: load a
// MatrixMarket file into opsparse::CSR, copy it to the device, run
// opsparse_spgemm / _reg / _safe on C = A*A, copy C back.  Writes C as
// binary (int32 M, N, nnz; rpt; col; double val) for the test to check.
//
// usage: dropin_driver <matrix.mtx> <out.bin>
#include <cstdio>
#include <cstdlib>

#include "opsparse/core_algo/spgemm/spgemm.hpp"
#include "opsparse/numeric/compute_flop.hpp"

using namespace opsparse;

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <matrix.mtx> <out.bin>\n", argv[0]);
        return 2;
    }
    CSR A(argv[1]);
    A.H2D();
    const long flop = compute_flop(A, A);
    SpgemmMeta meta;
    SpgemmTimings timing;
    CSR C;
    opsparse_spgemm(A, A, C, meta, timing);
    C.D2H();

    CSR C_reg;
    SpgemmTimings t_reg;
    opsparse_spgemm_reg(A, A, C_reg, meta, t_reg);
    C_reg.D2H();
    if (!(C == C_reg)) {
        std::fprintf(stderr, "opsparse_spgemm and opsparse_spgemm_reg differ\n");
        return 1;
    }
    CSR C_safe;
    SpgemmTimings t_safe;
    SpgemmPerf perf;
    const SpgemmStatus st = opsparse_spgemm_safe(A, A, C_safe, meta, t_safe, perf);
    if (st != SpgemmStatus::ok) {
        std::fprintf(stderr, "opsparse_spgemm_safe returned %d\n", static_cast<int>(st));
        return 1;
    }
    C_safe.D2H();
    if (!(C == C_safe)) {
        std::fprintf(stderr, "opsparse_spgemm and opsparse_spgemm_safe differ\n");
        return 1;
    }
    // The workspace's public mirrors, as the reference's SpgemmMeta has
    // them after a call: numeric bin sizes and offsets, max row nnz and
    // nnz(C) on the host and (after d2h_all) the same from the device.
    mint rows = 0;
    for (int b = 0; b < kNumBin; b++) {
        if (meta.bin_offset[b] != rows) {
            std::fprintf(stderr, "bin_offset[%d] = %d, expected %d\n", b, meta.bin_offset[b], rows);
            return 1;
        }
        rows += meta.bin_size[b];
    }
    if (rows > A.M || *meta.total_nnz != C.nnz || *meta.max_row_nnz <= 0 || !meta.stream || !meta.d_bin_size) {
        std::fprintf(stderr, "meta mirrors: rows %d of %d, total_nnz %d (C %d), max_row_nnz %d\n", rows, A.M,
                     *meta.total_nnz, C.nnz, *meta.max_row_nnz);
        return 1;
    }
    mint saved[kNumBin + 2];
    for (int i = 0; i < kNumBin + 2; i++) saved[i] = meta.bin_size[i];
    for (int i = 0; i < kNumBin + 2; i++) meta.bin_size[i] = -7;
    meta.d2h_all(0);
    OPSPARSE_CHECK_HIP(hipStreamSynchronize(meta.stream[0]));
    for (int i = 0; i < kNumBin + 2; i++) {
        if (meta.bin_size[i] != saved[i]) {
            std::fprintf(stderr, "device mirror %d = %d, host %d\n", i, meta.bin_size[i], saved[i]);
            return 1;
        }
    }
    std::printf("products %ld nnz(C) %d total %.6f s\n", flop, C.nnz, timing.total);

    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    const int hdr[3] = {C.M, C.N, C.nnz};
    std::fwrite(hdr, sizeof(int), 3, f);
    std::fwrite(C.rpt, sizeof(int), static_cast<size_t>(C.M) + 1, f);
    std::fwrite(C.col, sizeof(int), static_cast<size_t>(C.nnz), f);
    std::fwrite(C.val, sizeof(double), static_cast<size_t>(C.nnz), f);
    std::fclose(f);
    C.release();
    C_reg.release();
    C_safe.release();
    return 0;
}
