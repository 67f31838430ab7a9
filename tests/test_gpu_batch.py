"""Batched SpGEMM: instances packed block-diagonally run as one product.

Every instance's rows of the fused result must equal (structure bit-exact,
values within tolerance) both the CPU oracle and the plain call looped over
the instances (row-slice descriptors of the packed matrices).
"""
import numpy as np
import pytest
import torch

import oracle
import opsparse_amd.gen as G
from opsparse_amd.spgemm import DeviceCSR, Plan, spgemm

pytestmark = pytest.mark.gpu


def row_slice(D: DeviceCSR, r0: int, r1: int) -> DeviceCSR:
    # descriptor of rows [r0, r1): rpt keeps absolute offsets into col/val
    return DeviceCSR(r1 - r0, D.N, D.rpt[r0:r1 + 1], D.col, D.val)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_fused_equals_loop_and_oracle(gpu, dtype):
    b = G.batch(48, 7, n_lo=64, n_hi=2048)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    A, B = b.A.astype(npdt), b.B.astype(npdt)
    dA, dB = DeviceCSR.from_host(A, dtype), DeviceCSR.from_host(B, dtype)
    fused = spgemm(dA, dB, wide=True)
    frpt, fcol, fval = fused.to_host()
    ref = oracle.spgemm(A, B)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    oracle.compare(frpt, fcol, fval, ref, tol)
    for i in range(b.count):
        r0, r1 = int(b.row_base[i]), int(b.row_base[i + 1])
        Ci = spgemm(row_slice(dA, r0, r1), dB)
        rpt, col, val = Ci.to_host()
        s0, s1 = frpt[r0], frpt[r1]
        assert np.array_equal(rpt, frpt[r0:r1 + 1] - s0)
        assert np.array_equal(col, fcol[s0:s1])
        assert col.size == 0 or (col.min() >= r0 and col.max() < r1)  # stays in the instance's block
        # same kernels, same order inside a row -> same values
        assert np.allclose(val, fval[s0:s1], rtol=0, atol=0) or np.max(np.abs(val - fval[s0:s1])) <= tol * np.max(np.abs(fval[s0:s1]) + 1)


def test_batch_plan_graph_replays(gpu):
    b = G.batch(32, 9, n_lo=64, n_hi=1024)
    dA, dB = DeviceCSR.from_host(b.A, torch.float64), DeviceCSR.from_host(b.B, torch.float64)
    p = Plan(dA, dB)
    ref = oracle.spgemm(b.A, b.B)
    rpt = p.rowptr64().cpu().numpy()
    col = torch.empty(p.nnz, dtype=torch.int32, device="cuda")
    val = torch.empty(p.nnz, dtype=torch.float64, device="cuda")
    for rep in range(3):
        val.zero_()
        p.numeric(dA, dB, col, val, graph=True)
        torch.cuda.synchronize()
        # eager work between replays
        _ = (dA.val * 2).sum().item()
        oracle.compare(rpt, col.cpu().numpy(), val.cpu().numpy(), ref, 1e-12)
