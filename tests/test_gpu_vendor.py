"""The benchmark-only rocSPARSE comparator (libvendor_rocsparse.so, the role
of the reference's cuSPARSE baseline, benchmarks/cusparse_reg.cu): its
product is checked against the oracle and against this library's product
on the same device inputs, so the timing comparison in bench.py compares
two correct results."""
import ctypes as C
import os

import numpy as np
import pytest
import torch

import oracle
import opsparse_amd.gen as G
from opsparse_amd._lib import OPS_F32, OPS_F64, OpsCsr
from opsparse_amd.spgemm import DeviceCSR, LibCSR, _hip, spgemm
from helpers import TOL, random_csr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def vendor():
    path = os.path.join(ROOT, "opsparse_amd", "libvendor_rocsparse.so")
    if not os.path.exists(path):
        pytest.fail("libvendor_rocsparse.so is not built")
    lib = C.CDLL(path)
    f = lib.vendor_rocsparse_spgemm
    f.restype = C.c_int
    f.argtypes = [C.POINTER(OpsCsr), C.POINTER(OpsCsr), C.c_int, C.c_int, C.POINTER(C.c_double),
                  C.POINTER(C.c_int64), C.POINTER(OpsCsr)]
    return f


def vendor_product(A, B, dtype):
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    dA, dB = DeviceCSR.from_host(A, tdt), DeviceCSR.from_host(B, tdt)
    a, b, out = dA.desc(), dB.desc(), OpsCsr()
    sec, nnz = C.c_double(), C.c_int64()
    st = vendor()(C.byref(a), C.byref(b), OPS_F32 if dtype == np.float32 else OPS_F64, 1, C.byref(sec), C.byref(nnz),
                  C.byref(out))
    assert st == 0, st
    assert out.nnz == nnz.value
    host = LibCSR(out, False, tdt, owned=False).to_host()
    hip = _hip()
    hip.hipFree.argtypes = [C.c_void_p]
    for p in (out.rpt, out.col, out.val):
        if p:
            hip.hipFree(p)
    ours = spgemm(dA, dB).to_host()
    return host, ours


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_vendor_product_matches_oracle_and_ours(gpu, dtype):
    rng = np.random.default_rng(21)
    cases = [
        (G.uniform_rows(4096, 4096, 8, 0), None),
        (G.stencil27(12, 1), None),
        (random_csr(300, 500, rng.integers(0, 40, 300), rng), random_csr(500, 700, rng.integers(0, 60, 500), rng)),
    ]
    for A, B in cases:
        B = A if B is None else B
        (rpt, col, val), (orpt, ocol, oval) = vendor_product(A, B, dtype)
        ref = oracle.spgemm(A, B)
        oracle.compare(rpt, col, val, ref, TOL[dtype])
        assert np.array_equal(rpt, orpt) and np.array_equal(col, ocol)
