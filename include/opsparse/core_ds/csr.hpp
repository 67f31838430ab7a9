// CSR matrix with host and device mirrors: the drop-in for the reference's
// opsparse::CSR (include/opsparse/core_ds/csr.hpp:10-37), same members and
// methods.  Zero-based, int32 row pointers and columns, columns strictly
// ascending per row, FP64 values.
#ifndef OPSPARSE_CORE_DS_CSR_HPP_
#define OPSPARSE_CORE_DS_CSR_HPP_

#include <string>

#include "opsparse/system/hip_common.hpp"

namespace opsparse {

class CSR {
   public:
    mint M = 0;
    mint N = 0;
    mint nnz = 0;
    mint *rpt = nullptr;
    mint *col = nullptr;
    mdouble *val = nullptr;

    mint *d_rpt = nullptr;
    mint *d_col = nullptr;
    mdouble *d_val = nullptr;

    CSR() = default;
    explicit CSR(const std::string &mtx_file);
    CSR(const CSR &A);
    // Sub-matrix rows [M_start, M_start+M_), columns [N_start, N_start+N_).
    CSR(const CSR &A, mint M_, mint N_, mint M_start, mint N_start);
    ~CSR();

    void hrelease();
    void drelease();
    void release();
    void D2H();
    void H2D();
    // Structure exact; values within 1e-9 absolute or relative (csr.cu:149-183).
    bool operator==(const CSR &A);
    CSR &operator=(const CSR &A);
    void construct(const std::string &mtx_file);
};

}  // namespace opsparse

#endif  // OPSPARSE_CORE_DS_CSR_HPP_
