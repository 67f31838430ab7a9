// Packing of batch instances, shared by ops_batch_pack and the batched
// executor (ops_exec_create_batch), whose graph replays the pack kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "opsparse_amd.h"

namespace opsamd {

// The instances of one side of a batch (the A_i or the B_i): a device table
// of their descriptors with row, column and entry bases, and the packed sizes.
struct PackedSide {
    void *inst = nullptr;  // device table, one entry per instance with rows
    int nlive = 0;
    int rows = 0, cols = 0, nnz = 0;
    void release();
};

// Checks the host descriptors, reads each instance's rpt[0] and rpt[M] from
// the device (synchronises `s`) and builds the table.  row_base and col_base
// (count + 1 host entries, may be null) receive the packed bases.  Packed
// block-diagonally by default; `stack` keeps every instance's columns (all
// column bases 0, as many columns as the widest instance).  Returns an ops
// status.
int pack_side_prepare(int count, const ops_csr *instances, PackedSide &side, int64_t *row_base, int64_t *col_base,
                      hipStream_t s, bool stack = false);

// Packs the side's instances into prpt (rows + 1), pcol and pval (nnz each)
// with kernels only (no host synchronisation; may be captured in a graph).
void pack_side_launch(const PackedSide &side, bool f64, int *prpt, int *pcol, void *pval, hipStream_t s);

}  // namespace opsamd
