// Per-call workspace of the drop-in API.  Keeps the name and the public
// bin-count mirrors of the reference's SpgemmMeta
// (include/opsparse/core_algo/spgemm/meta.hpp:15-60) but owns a persistent
// ops_workspace (HIP stream pool, events, growable device buffers) that is
// created on first use and reused by later calls, instead of allocating
// and destroying streams and buffers inside every call (meta.cu:16-65).
#ifndef OPSPARSE_CORE_ALGO_SPGEMM_META_HPP_
#define OPSPARSE_CORE_ALGO_SPGEMM_META_HPP_

#include <cstddef>

#include "opsparse/system/hip_common.hpp"
#include "opsparse/utils/config.hpp"

struct ops_workspace;

namespace opsparse {

class CSR;

class SpgemmMeta {
   public:
    mint M = 0;
    mint N = 0;
    // Host mirrors filled after each call: rows per symbolic / numeric bin,
    // max row nnz and nnz(C) (the reference's bin_size, max_row_nnz, total_nnz).
    mint bin_size[kNumBin] = {};
    mint numeric_bin_size[kNumBin] = {};
    mint max_row_nnz = 0;
    long long total_nnz = 0;
    // Streams of the pool (valid after the first call).
    hipStream_t *stream = nullptr;

    SpgemmMeta() = default;
    SpgemmMeta(const SpgemmMeta &) = delete;
    SpgemmMeta &operator=(const SpgemmMeta &) = delete;
    ~SpgemmMeta();

    // Creates the workspace if needed (idempotent).
    void allocate(CSR &C);
    // Destroys the workspace (the next call recreates it).
    void release();
    ops_workspace *workspace();

   private:
    ops_workspace *ws_ = nullptr;
};

}  // namespace opsparse

#endif  // OPSPARSE_CORE_ALGO_SPGEMM_META_HPP_
