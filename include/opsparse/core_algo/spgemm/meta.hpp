// Workspace of the drop-in API.  Keeps the name, the public members and the
// helper methods of the reference's SpgemmMeta
// (include/opsparse/core_algo/spgemm/meta.hpp:17-57) but owns a persistent
// ops_workspace (HIP stream pool, events, growable device buffers) that is
// created on first use and reused by later calls, instead of allocating
// and destroying streams and buffers inside every call (meta.cu:16-65).
//
// After each call the bin mirrors describe its numeric bins, as the
// reference's do after its numeric phase: bin_size / bin_offset (rows per
// bin and their exclusive prefix), max_row_nnz and total_nnz on the host,
// and the same four arrays on the device (d_bin_size, ...); d_bins is the
// workspace's list of row ids grouped by bin (valid until the next call).
// The reference's CUB scratch and global-hash memory pool have no
// counterpart here (hand-written scan; long rows need no global hash
// table): d_cub_storage and d_global_mem_pool stay null.
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
    // combined device / host memory (one allocation each) that the mirrors
    // below point into: bin_size[kNumBin], bin_offset[kNumBin],
    // max_row_nnz[1], total_nnz[1]
    mint *d_combined_mem = nullptr;
    mint *combined_mem = nullptr;

    mint M = 0;
    mint N = 0;

    mint *d_bins = nullptr;         // size M
    mint *d_bin_size = nullptr;     // size kNumBin
    mint *d_bin_offset = nullptr;   // size kNumBin
    mint *d_max_row_nnz = nullptr;  // size 1
    mint *d_total_nnz = nullptr;    // size 1
    mint *d_cub_storage = nullptr;  // no counterpart: null

    mint *bin_size = nullptr;     // size kNumBin
    mint *bin_offset = nullptr;   // size kNumBin
    mint *max_row_nnz = nullptr;  // size 1
    mint *total_nnz = nullptr;    // size 1; -1 when nnz(C) exceeds INT32_MAX (see total_nnz64)

    std::size_t cub_storage_size = 0;
    hipStream_t *stream = nullptr;  // the workspace's kNumBin pool streams

    mint *d_global_mem_pool = nullptr;  // no counterpart: null
    std::size_t global_mem_pool_size = 0;
    bool global_mem_pool_malloced = false;

    // Additions: rows per symbolic bin of the last call and its 64-bit nnz(C).
    mint symbolic_bin_size[kNumBin] = {};
    long long total_nnz64 = 0;

    SpgemmMeta() = default;
    SpgemmMeta(const SpgemmMeta &) = delete;
    SpgemmMeta &operator=(const SpgemmMeta &) = delete;
    explicit SpgemmMeta(CSR &C);
    ~SpgemmMeta();

    // Allocates C.d_rpt (M + 1 entries) for C of A.M rows, as the reference
    // does before its symbolic phase (meta.cu:67-72); the call itself
    // allocates C's arrays, so this is only for callers that fill C.d_rpt.
    void allocate_rpt(CSR &C);
    // Creates the workspace and the mirrors if needed (idempotent).
    void allocate(CSR &C);
    // Destroys the workspace and the mirrors (the next call recreates them).
    void release();

    // The reference's mirror helpers (meta.cu:74-105), on pool stream
    // stream_idx: zero the device bin sizes / every device mirror, copy
    // them to the host, or copy the host bin offsets to the device.
    void memset_bin_size(mint stream_idx);
    void memset_all(mint stream_idx);
    void d2h_bin_size(mint stream_idx);
    void d2h_all(mint stream_idx);
    void h2d_bin_offset(mint stream_idx);

    ops_workspace *workspace();
    // Fills the mirrors from the workspace after a call (host and device).
    void update_from_last_call(const mint *numeric_rows, const mint *symbolic_rows, mint max_nnz, long long nnz);

   private:
    ops_workspace *ws_ = nullptr;
    hipStream_t pool_[kNumBin] = {};
    hipStream_t at(mint stream_idx) const;
};

}  // namespace opsparse

#endif  // OPSPARSE_CORE_ALGO_SPGEMM_META_HPP_
