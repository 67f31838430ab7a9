"""The allocation-free entry point (ops_spgemm64_noalloc) against the
allocating one (ops_spgemm64_f64 / _f32) on the same inputs.

Two calls in a row, the second with a smaller input (the shared scratch
keeps the first call's contents), FP64 and FP32, short rows, a wave-rank
row spilled to the block kernel and long rows on the cursor path; the
caller's buffers come from torch before the calls, and the device's free
memory is the same before and after them.  Structure must be bit-exact and
values within the parity tolerance of each other and of the oracle.  Also:
an output capacity below nnz(C) and inputs above the limits are refused
with OPS_INVALID."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from opsparse_amd._lib import OPS_F32, OPS_F64, OPS_INVALID, OPS_OK, OpsCsr64, lib
from opsparse_amd.spgemm import DeviceCSR, Workspace, _stream_ptr, spgemm
from helpers import TOL, random_csr

pytestmark = pytest.mark.gpu


def mem_free():
    free, _ = torch.cuda.mem_get_info()
    return free


def noalloc(dA, dB, rpt, col, val, capacity):
    c = OpsCsr64()
    c.rpt, c.col, c.val = rpt.data_ptr(), col.data_ptr(), val.data_ptr()
    a, b = dA.desc(), dB.desc()
    dt = OPS_F64 if dA.val.dtype == torch.float64 else OPS_F32
    st = lib().ops_spgemm64_noalloc(dt, C.byref(a), C.byref(b), C.byref(c), capacity, _stream_ptr())
    return st, c


def case(seed, M, K, N, long_rows):
    rng = np.random.default_rng(seed)
    la = np.r_[rng.integers(0, 12, M - long_rows), rng.integers(300, 400, long_rows)]
    A = random_csr(M, K, la, rng)
    B = random_csr(K, N, rng.integers(0, 300, K), rng)
    return A, B


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_matches_allocating_entry(gpu, dtype):
    npdt = np.float64 if dtype == torch.float64 else np.float32
    big = case(5, 3000, 4000, 200000, 6)    # long rows: cursor path over two 2^17-column windows
    small = case(6, 700, 1500, 60000, 2)
    ws = Workspace()
    inputs = []
    for A, B in (big, small):
        Ah, Bh = A.astype(npdt), B.astype(npdt)
        dA, dB = DeviceCSR.from_host(Ah, dtype), DeviceCSR.from_host(Bh, dtype)
        ref = spgemm(dA, dB, ws=ws, wide=True)
        rr, rc, rv = ref.to_host()
        assert ref.nnz > 0
        cap = ref.nnz + 100
        rpt = torch.full((A.M + 1,), -7, dtype=torch.int64, device="cuda")
        col = torch.full((cap,), -7, dtype=torch.int32, device="cuda")
        val = torch.full((cap,), 7.0, dtype=dtype, device="cuda")
        inputs.append((A, B, Ah, Bh, dA, dB, rr, rc, rv, rpt, col, val, cap))
        ref.free()
    torch.cuda.synchronize()
    free = [mem_free()]
    for A, B, Ah, Bh, dA, dB, rr, rc, rv, rpt, col, val, cap in inputs:   # big, then small
        st, c = noalloc(dA, dB, rpt, col, val, cap)
        assert st == OPS_OK and c.M == A.M and c.N == B.N and c.nnz == rc.size
        torch.cuda.synchronize()
        hr, hc, hv = rpt.cpu().numpy(), col[:c.nnz].cpu().numpy(), val[:c.nnz].cpu().numpy()
        assert np.array_equal(hr, rr) and np.array_equal(hc, rc)
        scale = np.maximum(np.abs(rv.astype(np.float64)), 1e-300)
        oracle.compare(hr, hc, hv, oracle.spgemm(Ah, Bh), TOL[npdt])
        assert np.max(np.abs(hv.astype(np.float64) - rv) / scale, initial=0.0) <= (1e-12 if npdt == np.float64 else 1e-6)
        assert torch.all(col[c.nnz:] == -7)  # nothing written past nnz(C)
        torch.cuda.synchronize()
        free.append(mem_free())
    # The first call on a device may load the code object that holds the
    # scratch array (about 21 MB); after that no call takes device memory.
    assert free[0] - free[1] <= 64 << 20 and free[2] == free[1], free


def test_noalloc_refusals(gpu):
    A, B = case(7, 200, 500, 5000, 0)
    dA, dB = DeviceCSR.from_host(A), DeviceCSR.from_host(B)
    nnz = oracle.spgemm(A, B)[1].size
    rpt = torch.empty(A.M + 1, dtype=torch.int64, device="cuda")
    col = torch.full((nnz,), -7, dtype=torch.int32, device="cuda")
    val = torch.empty(nnz, dtype=torch.float64, device="cuda")
    st, c = noalloc(dA, dB, rpt, col, val, nnz - 1)  # one entry short
    assert st == OPS_INVALID and c.nnz == nnz
    assert torch.all(col == -7)
    st, c = noalloc(dA, dB, rpt, col, val, nnz)
    assert st == OPS_OK and c.nnz == nnz
    # more rows than the limit (262,144)
    M = (1 << 18) + 1
    tall = DeviceCSR(M, B.M, torch.zeros(M + 1, dtype=torch.int32, device="cuda"), dA.col, dA.val, 0)
    big_rpt = torch.empty(M + 1, dtype=torch.int64, device="cuda")
    st, _ = noalloc(tall, dB, big_rpt, col, val, nnz)
    assert st == OPS_INVALID
    # entry indices beyond the limit (2^20): a valid two-row slice whose
    # entries sit at indices 2^20 and 2^20 + 1 of its column array
    n = (1 << 20) + 2
    far_col = torch.zeros(n, dtype=torch.int32, device="cuda")
    far_col[n - 1] = 1
    far_rpt = torch.tensor([n - 2, n - 1, n], dtype=torch.int32, device="cuda")
    far = DeviceCSR(2, B.M, far_rpt, far_col, torch.ones(n, dtype=torch.float64, device="cuda"), 2)
    st, _ = noalloc(far, dB, rpt, col, val, nnz)
    assert st == OPS_INVALID
