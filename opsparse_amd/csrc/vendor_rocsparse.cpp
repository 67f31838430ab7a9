// Benchmark-only vendor comparator: rocSPARSE's generic SpGEMM on the same
// device CSR inputs, timed with the drop-in call's convention (host wall
// clock around a complete product, C's arrays allocated inside the timed
// call and freed outside it).  It mirrors the reference's cuSPARSE
// comparator (cusparse_spgemm.cu:41-66, cusparse_reg.cu:10-44) and is built
// as its own library (libvendor_rocsparse.so): nothing in libopsparse_amd
// calls it, so it is never a dispatch path.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <time.h>

#include <cstdint>
#include <cstdio>

#include "opsparse_amd.h"

namespace {

double now_s()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC_RAW, &ts);
    return static_cast<double>(ts.tv_sec) + static_cast<double>(ts.tv_nsec) * 1e-9;
}

#define RS(x)                                                                      \
    do {                                                                           \
        const rocsparse_status s_ = (x);                                           \
        if (s_ != rocsparse_status_success) {                                      \
            std::fprintf(stderr, "rocsparse error %d at %s\n", int(s_), #x);       \
            return -1;                                                             \
        }                                                                          \
    } while (0)
#define HC(x)                                                                      \
    do {                                                                           \
        const hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "hip error %s at %s\n", hipGetErrorString(e_), #x); \
            return -1;                                                             \
        }                                                                          \
    } while (0)

// One complete C = A*B: buffer query, nnz stage, C allocation, compute.
// On success C's arrays are returned for the caller to check and free.
int one_product(rocsparse_handle h, const ops_csr *A, const ops_csr *B, rocsparse_datatype dt, size_t vbytes,
                int64_t *nnz_out, int32_t **c_rpt, int32_t **c_col, void **c_val)
{
    rocsparse_spmat_descr ma, mb, mc, md;
    RS(rocsparse_create_csr_descr(&ma, A->M, A->N, A->nnz, A->rpt, A->col, A->val, rocsparse_indextype_i32,
                                  rocsparse_indextype_i32, rocsparse_index_base_zero, dt));
    RS(rocsparse_create_csr_descr(&mb, B->M, B->N, B->nnz, B->rpt, B->col, B->val, rocsparse_indextype_i32,
                                  rocsparse_indextype_i32, rocsparse_index_base_zero, dt));
    int32_t *rpt = nullptr;
    HC(hipMalloc(&rpt, sizeof(int32_t) * (static_cast<size_t>(A->M) + 1)));
    RS(rocsparse_create_csr_descr(&mc, A->M, B->N, 0, rpt, nullptr, nullptr, rocsparse_indextype_i32,
                                  rocsparse_indextype_i32, rocsparse_index_base_zero, dt));
    // D is unused (beta = nullptr) but the interface takes a descriptor.
    RS(rocsparse_create_csr_descr(&md, A->M, B->N, 0, rpt, nullptr, nullptr, rocsparse_indextype_i32,
                                  rocsparse_indextype_i32, rocsparse_index_base_zero, dt));
    const double one64 = 1.0;
    const float one32 = 1.0f;
    const void *alpha = dt == rocsparse_datatype_f64_r ? static_cast<const void *>(&one64)
                                                        : static_cast<const void *>(&one32);
    size_t bytes = 0;
    RS(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, alpha, ma, mb, nullptr, md, mc, dt,
                        rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_buffer_size, &bytes, nullptr));
    void *buf = nullptr;
    HC(hipMalloc(&buf, bytes > 0 ? bytes : 1));
    RS(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, alpha, ma, mb, nullptr, md, mc, dt,
                        rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_nnz, &bytes, buf));
    int64_t rows, cols, nnz;
    RS(rocsparse_spmat_get_size(mc, &rows, &cols, &nnz));
    int32_t *col = nullptr;
    void *val = nullptr;
    HC(hipMalloc(&col, sizeof(int32_t) * static_cast<size_t>(nnz > 0 ? nnz : 1)));
    HC(hipMalloc(&val, vbytes * static_cast<size_t>(nnz > 0 ? nnz : 1)));
    RS(rocsparse_csr_set_pointers(mc, rpt, col, val));
    RS(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, alpha, ma, mb, nullptr, md, mc, dt,
                        rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_compute, &bytes, buf));
    HC(hipDeviceSynchronize());
    HC(hipFree(buf));
    RS(rocsparse_destroy_spmat_descr(ma));
    RS(rocsparse_destroy_spmat_descr(mb));
    RS(rocsparse_destroy_spmat_descr(mc));
    RS(rocsparse_destroy_spmat_descr(md));
    *nnz_out = nnz;
    *c_rpt = rpt;
    *c_col = col;
    *c_val = val;
    return 0;
}

}  // namespace

extern "C" {

// Times rocSPARSE SpGEMM C = A*B on device CSR inputs (dtype OPS_F64 or
// OPS_F32): one warm-up product, then `reps` products, each bracketed by a
// device synchronisation; returns the mean seconds per product in *sec and
// nnz(C) in *nnz.  If out is not null, the last product's C is returned
// there (device arrays from hipMalloc, freed by the caller with hipFree);
// otherwise it is freed.  Returns 0 on success, -1 on any error.
int vendor_rocsparse_spgemm(const ops_csr *A, const ops_csr *B, int dtype, int reps, double *sec, int64_t *nnz,
                            ops_csr *out)
{
    if (!A || !B || !sec || !nnz || reps < 1 || A->N != B->M) return -1;
    const rocsparse_datatype dt = dtype == OPS_F32 ? rocsparse_datatype_f32_r : rocsparse_datatype_f64_r;
    const size_t vbytes = dtype == OPS_F32 ? sizeof(float) : sizeof(double);
    rocsparse_handle h;
    RS(rocsparse_create_handle(&h));
    double total = 0.0;
    int64_t z = 0;
    int32_t *rpt = nullptr, *col = nullptr;
    void *val = nullptr;
    for (int it = 0; it <= reps; it++) {
        HC(hipDeviceSynchronize());
        const double t0 = now_s();
        if (one_product(h, A, B, dt, vbytes, &z, &rpt, &col, &val) != 0) return -1;
        const double t = now_s() - t0;
        if (it > 0) total += t;
        const bool keep = out && it == reps;
        if (!keep) {
            HC(hipFree(rpt));
            HC(hipFree(col));
            HC(hipFree(val));
        }
    }
    RS(rocsparse_destroy_handle(h));
    *sec = total / reps;
    *nnz = z;
    if (out) {
        out->M = A->M;
        out->N = B->N;
        out->nnz = static_cast<int32_t>(z);
        out->rpt = rpt;
        out->col = col;
        out->val = val;
    }
    return 0;
}

}  // extern "C"
