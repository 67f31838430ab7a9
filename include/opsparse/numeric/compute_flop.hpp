// Products per row ("flop"): row_flop[i] = sum over A(i,k) != 0 of nnz(B row k),
// 64-bit total returned.  Host OpenMP versions with the reference's
// signatures (include/opsparse/numeric/compute_flop.hpp:11-13); the device
// counterpart of k_compute_flop is ops_compute_flop_device in opsparse_amd.h.
#ifndef OPSPARSE_NUMERIC_COMPUTE_FLOP_HPP_
#define OPSPARSE_NUMERIC_COMPUTE_FLOP_HPP_

#include "opsparse/system/hip_common.hpp"

namespace opsparse {

class CSR;

long compute_flop(const mint *row_ptr_a, const mint *col_idx_a, const mint *row_ptr_b, mint M, mint *row_flop_out);
long compute_flop(const CSR &A, const CSR &B, mint *row_flop);
long compute_flop(const CSR &A, const CSR &B);

}  // namespace opsparse

#endif  // OPSPARSE_NUMERIC_COMPUTE_FLOP_HPP_
