// Block-diagonal packing of independent CSR matrices (ops_batch_pack).
//
// One block per non-empty instance copies the instance's rows into its
// slice of the packed arrays, rebasing row offsets and shifting columns by
// the instance's column base.  The packed product then runs every instance
// through one plan: rows of all instances are binned together, so thousands
// of small products share a handful of kernel launches instead of paying a
// launch sequence and host synchronisation each.
//
// Each instance's entry count is taken from its own row offsets
// (rpt[M] - rpt[0], read from the device before packing) and must equal the
// descriptor's nnz; instances with no rows are skipped (their rpt is never
// read and may be null).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "opsparse_amd.h"
#include "outmem.hpp"

namespace {

struct InstDesc {
    const int *rpt;
    const int *col;
    const void *val;
    int M;
    int row_base;
    long long nnz_base;
    int col_base;
};

// ends[2i], ends[2i+1] = rpt[0], rpt[M] of instance i (M > 0 for every one).
__global__ void k_row_ends(const InstDesc *__restrict__ inst, int n, int *__restrict__ ends)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const InstDesc d = inst[i];
    ends[2 * i] = d.rpt[0];
    ends[2 * i + 1] = d.rpt[d.M];
}

template <typename V>
__global__ __launch_bounds__(256) void k_pack(const InstDesc *__restrict__ inst, int *__restrict__ prpt,
                                              int *__restrict__ pcol, V *__restrict__ pval)
{
    const InstDesc d = inst[blockIdx.x];
    const int off0 = d.rpt[0];
    for (int r = threadIdx.x; r < d.M; r += blockDim.x)
        prpt[d.row_base + r] = static_cast<int>(d.nnz_base + (d.rpt[r] - off0));
    const int nnz = d.rpt[d.M] - off0;
    const V *v = static_cast<const V *>(d.val);
    for (int j = threadIdx.x; j < nnz; j += blockDim.x) {
        pcol[d.nnz_base + j] = d.col[off0 + j] + d.col_base;
        pval[d.nnz_base + j] = v[off0 + j];
    }
}

__global__ void k_set_last(int *prpt, int M, int nnz)
{
    if (threadIdx.x == 0) prpt[M] = nnz;
}

}  // namespace

extern "C" int ops_batch_pack(int32_t count, const ops_csr *instances, ops_dtype dt, ops_csr *packed,
                              int64_t *row_base, int64_t *col_base, void *stream)
{
    if (!packed || count < 0 || (count > 0 && !instances)) return OPS_INVALID;
    if (dt != OPS_F64 && dt != OPS_F32) return OPS_INVALID;
    *packed = ops_csr{};
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // Host checks and the instances that own rows.
    std::vector<int> live;
    live.reserve(static_cast<size_t>(count));
    for (int i = 0; i < count; i++) {
        const ops_csr &c = instances[i];
        if (c.M < 0 || c.N < 0 || c.nnz < 0 || (c.M > 0 && !c.rpt) || (c.nnz > 0 && (!c.col || !c.val)))
            return OPS_INVALID;
        if (c.M > 0) live.push_back(i);
    }
    const int nlive = static_cast<int>(live.size());
    std::vector<InstDesc> h(static_cast<size_t>(nlive));
    InstDesc *d_inst = nullptr;
    int *d_ends = nullptr;
    auto cleanup = [&]() {
        if (d_inst) (void)hipFree(d_inst);
        if (d_ends) (void)hipFree(d_ends);
        d_inst = nullptr;
        d_ends = nullptr;
    };
    auto fail = [&](hipError_t e) {
        const int st = (e == hipErrorOutOfMemory) ? OPS_OOM : OPS_HIP_ERROR;
        (void)hipGetLastError();
        cleanup();
        ops_csr_free(packed);
        return st;
    };
    hipError_t e;
    std::vector<int> ends(2 * static_cast<size_t>(nlive));
    if (nlive > 0) {
        for (int j = 0; j < nlive; j++) {
            const ops_csr &c = instances[live[j]];
            h[j] = InstDesc{c.rpt, c.col, c.val, c.M, 0, 0, 0};
        }
        if ((e = hipMalloc(&d_inst, sizeof(InstDesc) * h.size())) != hipSuccess) return fail(e);
        if ((e = hipMalloc(&d_ends, sizeof(int) * ends.size())) != hipSuccess) return fail(e);
        if ((e = hipMemcpyAsync(d_inst, h.data(), sizeof(InstDesc) * h.size(), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return fail(e);
        k_row_ends<<<(nlive + 255) / 256, 256, 0, s>>>(d_inst, nlive, d_ends);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e);
        if ((e = hipMemcpyAsync(ends.data(), d_ends, sizeof(int) * ends.size(), hipMemcpyDeviceToHost, s)) !=
            hipSuccess)
            return fail(e);
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
    }
    long long rows = 0, cols = 0, nnz = 0;
    for (int i = 0, j = 0; i < count; i++) {
        const ops_csr &c = instances[i];
        if (row_base) row_base[i] = rows;
        if (col_base) col_base[i] = cols;
        if (c.M > 0) {
            const long long n = static_cast<long long>(ends[2 * j + 1]) - ends[2 * j];
            if (n < 0 || n != c.nnz) {
                cleanup();
                return OPS_INVALID;  // rpt[M] - rpt[0] must be the descriptor's nnz
            }
            h[j].row_base = static_cast<int>(rows);
            h[j].nnz_base = nnz;
            h[j].col_base = static_cast<int>(cols);
            j++;
            nnz += n;
        }
        rows += c.M;
        cols += c.N;
        if (rows > INT32_MAX || cols > INT32_MAX || nnz > INT32_MAX) {
            cleanup();
            return OPS_INT32_OVERFLOW;
        }
    }
    if (row_base) row_base[count] = rows;
    if (col_base) col_base[count] = cols;
    const size_t vsz = dt == OPS_F64 ? sizeof(double) : sizeof(float);
    if ((e = opsamd::out_malloc(reinterpret_cast<void **>(&packed->rpt),
                                sizeof(int) * (static_cast<size_t>(rows) + 1))) != hipSuccess)
        return fail(e);
    if ((e = opsamd::out_malloc(reinterpret_cast<void **>(&packed->col),
                                sizeof(int) * static_cast<size_t>(nnz > 0 ? nnz : 1))) != hipSuccess)
        return fail(e);
    if ((e = opsamd::out_malloc(&packed->val, vsz * static_cast<size_t>(nnz > 0 ? nnz : 1))) != hipSuccess)
        return fail(e);
    packed->M = static_cast<int32_t>(rows);
    packed->N = static_cast<int32_t>(cols);
    packed->nnz = static_cast<int32_t>(nnz);
    if (nlive > 0) {
        if ((e = hipMemcpyAsync(d_inst, h.data(), sizeof(InstDesc) * h.size(), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return fail(e);
        if (dt == OPS_F64)
            k_pack<double><<<nlive, 256, 0, s>>>(d_inst, packed->rpt, packed->col, static_cast<double *>(packed->val));
        else
            k_pack<float><<<nlive, 256, 0, s>>>(d_inst, packed->rpt, packed->col, static_cast<float *>(packed->val));
        if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    }
    k_set_last<<<1, 64, 0, s>>>(packed->rpt, static_cast<int>(rows), static_cast<int>(nnz));
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
    cleanup();
    return OPS_OK;
}
