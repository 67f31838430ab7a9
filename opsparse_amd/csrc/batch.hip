// Block-diagonal packing of independent CSR matrices (ops_batch_pack, and
// the packing step of a batched executor, ops_exec_create_batch).
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

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "batch.hpp"
#include "opsparse_amd.h"
#include "outmem.hpp"

namespace opsamd {

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

int status_of(hipError_t e)
{
    (void)hipGetLastError();
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? OPS_OOM : OPS_HIP_ERROR;
}

}  // namespace

void PackedSide::release()
{
    if (inst) (void)hipFree(inst);
    inst = nullptr;
    nlive = 0;
}

int pack_side_prepare(int count, const ops_csr *instances, PackedSide &side, int64_t *row_base, int64_t *col_base,
                      hipStream_t s, bool stack)
{
    side.release();
    side.rows = side.cols = side.nnz = 0;
    if (count < 0 || (count > 0 && !instances)) return OPS_INVALID;
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
    std::vector<int> ends(2 * static_cast<size_t>(nlive));
    InstDesc *d_inst = nullptr;
    int *d_ends = nullptr;
    auto fail = [&](int st) {
        if (d_inst) (void)hipFree(d_inst);
        if (d_ends) (void)hipFree(d_ends);
        return st;
    };
    hipError_t e;
    if (nlive > 0) {
        for (int j = 0; j < nlive; j++) {
            const ops_csr &c = instances[live[j]];
            h[j] = InstDesc{c.rpt, c.col, c.val, c.M, 0, 0, 0};
        }
        if ((e = hipMalloc(&d_inst, sizeof(InstDesc) * h.size())) != hipSuccess) return fail(status_of(e));
        if ((e = hipMalloc(&d_ends, sizeof(int) * ends.size())) != hipSuccess) return fail(status_of(e));
        if ((e = hipMemcpyAsync(d_inst, h.data(), sizeof(InstDesc) * h.size(), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return fail(status_of(e));
        k_row_ends<<<(nlive + 255) / 256, 256, 0, s>>>(d_inst, nlive, d_ends);
        if ((e = hipGetLastError()) != hipSuccess) return fail(status_of(e));
        if ((e = hipMemcpyAsync(ends.data(), d_ends, sizeof(int) * ends.size(), hipMemcpyDeviceToHost, s)) !=
            hipSuccess)
            return fail(status_of(e));
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(status_of(e));
        (void)hipFree(d_ends);
        d_ends = nullptr;
    }
    long long rows = 0, cols = 0, nnz = 0;
    for (int i = 0, j = 0; i < count; i++) {
        const ops_csr &c = instances[i];
        if (row_base) row_base[i] = rows;
        if (col_base) col_base[i] = cols;
        if (c.M > 0) {
            const long long n = static_cast<long long>(ends[2 * j + 1]) - ends[2 * j];
            if (n < 0 || n != c.nnz) return fail(OPS_INVALID);  // rpt[M] - rpt[0] must be the descriptor's nnz
            h[j].row_base = static_cast<int>(rows);
            h[j].nnz_base = nnz;
            h[j].col_base = stack ? 0 : static_cast<int>(cols);
            j++;
            nnz += n;
        }
        rows += c.M;
        cols = stack ? std::max<long long>(cols, c.N) : cols + c.N;
        if (rows > INT32_MAX || cols > INT32_MAX || nnz > INT32_MAX) return fail(OPS_INT32_OVERFLOW);
    }
    if (row_base) row_base[count] = rows;
    if (col_base) col_base[count] = cols;
    if (nlive > 0) {
        if ((e = hipMemcpyAsync(d_inst, h.data(), sizeof(InstDesc) * h.size(), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return fail(status_of(e));
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(status_of(e));
    }
    side.inst = d_inst;
    side.nlive = nlive;
    side.rows = static_cast<int>(rows);
    side.cols = static_cast<int>(cols);
    side.nnz = static_cast<int>(nnz);
    return OPS_OK;
}

void pack_side_launch(const PackedSide &side, bool f64, int *prpt, int *pcol, void *pval, hipStream_t s)
{
    const InstDesc *inst = static_cast<const InstDesc *>(side.inst);
    if (side.nlive > 0) {
        if (f64)
            k_pack<double><<<side.nlive, 256, 0, s>>>(inst, prpt, pcol, static_cast<double *>(pval));
        else
            k_pack<float><<<side.nlive, 256, 0, s>>>(inst, prpt, pcol, static_cast<float *>(pval));
    }
    k_set_last<<<1, 64, 0, s>>>(prpt, side.rows, side.nnz);
}

}  // namespace opsamd

extern "C" int ops_batch_pack(int32_t count, const ops_csr *instances, ops_dtype dt, ops_csr *packed,
                              int64_t *row_base, int64_t *col_base, void *stream)
{
    if (!packed || count < 0 || (count > 0 && !instances)) return OPS_INVALID;
    if (dt != OPS_F64 && dt != OPS_F32) return OPS_INVALID;
    *packed = ops_csr{};
    const hipStream_t s = static_cast<hipStream_t>(stream);
    opsamd::PackedSide side;
    int st = opsamd::pack_side_prepare(count, instances, side, row_base, col_base, s);
    if (st != OPS_OK) return st;
    auto fail = [&](hipError_t e) {
        st = opsamd::status_of(e);
        side.release();
        ops_csr_free(packed);
        return st;
    };
    hipError_t e;
    const size_t vsz = dt == OPS_F64 ? sizeof(double) : sizeof(float);
    const size_t nz = static_cast<size_t>(side.nnz > 0 ? side.nnz : 1);
    if ((e = opsamd::out_malloc(reinterpret_cast<void **>(&packed->rpt),
                                sizeof(int) * (static_cast<size_t>(side.rows) + 1))) != hipSuccess)
        return fail(e);
    if ((e = opsamd::out_malloc(reinterpret_cast<void **>(&packed->col), sizeof(int) * nz)) != hipSuccess)
        return fail(e);
    if ((e = opsamd::out_malloc(&packed->val, vsz * nz)) != hipSuccess) return fail(e);
    packed->M = side.rows;
    packed->N = side.cols;
    packed->nnz = side.nnz;
    opsamd::pack_side_launch(side, dt == OPS_F64, packed->rpt, packed->col, packed->val, s);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
    side.release();
    return OPS_OK;
}
