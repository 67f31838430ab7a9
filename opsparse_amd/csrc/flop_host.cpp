// Host product count with OpenMP (C ABI + the drop-in C++ overloads).
// Restates src/opsparse/numeric/compute_flop.cu:33-65: per-row sum of the
// B-row lengths of the row's A-nonzeros, 64-bit total.
#include <omp.h>

#include <cstdint>
#include <vector>

#include "opsparse/core_ds/csr.hpp"
#include "opsparse/numeric/compute_flop.hpp"
#include "opsparse_amd.h"

extern "C" int64_t ops_compute_flop_host(int32_t M, const int32_t *a_rpt, const int32_t *a_col, const int32_t *b_rpt,
                                         int64_t *row_flop_out)
{
    if (M <= 0 || !a_rpt || !a_col || !b_rpt) return 0;
    int64_t total = 0;
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int32_t i = 0; i < M; i++) {
        int64_t s = 0;
        for (int32_t j = a_rpt[i]; j < a_rpt[i + 1]; j++) s += static_cast<int64_t>(b_rpt[a_col[j] + 1] - b_rpt[a_col[j]]);
        if (row_flop_out) row_flop_out[i] = s;
        total += s;
    }
    return total;
}

namespace opsparse {

long compute_flop(const mint *row_ptr_a, const mint *col_idx_a, const mint *row_ptr_b, mint M, mint *row_flop_out)
{
    if (!row_flop_out) return static_cast<long>(ops_compute_flop_host(M, row_ptr_a, col_idx_a, row_ptr_b, nullptr));
    std::vector<int64_t> tmp(static_cast<size_t>(M > 0 ? M : 0));
    const long total = static_cast<long>(ops_compute_flop_host(M, row_ptr_a, col_idx_a, row_ptr_b, tmp.data()));
    for (mint i = 0; i < M; i++) row_flop_out[i] = static_cast<mint>(tmp[i]);
    return total;
}

long compute_flop(const CSR &A, const CSR &B, mint *row_flop) { return compute_flop(A.rpt, A.col, B.rpt, A.M, row_flop); }

long compute_flop(const CSR &A, const CSR &B) { return compute_flop(A.rpt, A.col, B.rpt, A.M, nullptr); }

}  // namespace opsparse
