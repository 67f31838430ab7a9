// Drop-in entry points: same signatures as the reference
// (include/opsparse/core_algo/spgemm/spgemm.hpp:12-22), built on the C ABI
// in opsparse_amd.h.  A and B must have their device arrays (CSR::H2D);
// C's device arrays are allocated by the call (free with C.release()).
#ifndef OPSPARSE_CORE_ALGO_SPGEMM_SPGEMM_HPP_
#define OPSPARSE_CORE_ALGO_SPGEMM_SPGEMM_HPP_

#include "opsparse/core_algo/spgemm/meta.hpp"
#include "opsparse/core_ds/csr.hpp"
#include "opsparse/profile/perf.hpp"
#include "opsparse/profile/spgemm_timings.hpp"

namespace opsparse {

// Full pipeline with per-phase host timing (synchronises after each phase).
void opsparse_spgemm(const CSR &A, const CSR &B, CSR &C, SpgemmMeta &meta, SpgemmTimings &timing);

// Same result with no per-phase synchronisation; only total is timed.
void opsparse_spgemm_reg(const CSR &A, const CSR &B, CSR &C, SpgemmMeta &meta, SpgemmTimings &timing);

// Never throws: pre-checks A.N == B.M and the int32 product count on the
// host arrays, classifies failures, records memory use.
SpgemmStatus opsparse_spgemm_safe(const CSR &A, const CSR &B, CSR &C, SpgemmMeta &meta, SpgemmTimings &timing,
                                  SpgemmPerf &perf);

}  // namespace opsparse

#endif  // OPSPARSE_CORE_ALGO_SPGEMM_SPGEMM_HPP_
