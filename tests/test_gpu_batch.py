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


def test_batch_pack_matches_generator_packing(gpu):
    b = G.batch(20, 11, n_lo=64, n_hi=512)
    insts_a, insts_b = [], []
    for i in range(b.count):
        r0, r1 = int(b.row_base[i]), int(b.row_base[i + 1])
        for X, out in ((b.A, insts_a), (b.B, insts_b)):
            h = X.row_slice(r0, r1)
            h = G.HostCSR(h.M, r1 - r0, h.rpt, (h.col - r0).astype(np.int32), h.val)  # local columns
            out.append(DeviceCSR.from_host(h, torch.float64))
    from opsparse_amd.spgemm import batch_pack
    pa, rb, cb = batch_pack(insts_a)
    assert np.array_equal(rb, b.row_base) and np.array_equal(cb, b.row_base)
    rpt, col, val = pa.to_host()
    assert np.array_equal(rpt, b.A.rpt) and np.array_equal(col, b.A.col) and np.array_equal(val, b.A.val)
    pb, _, _ = batch_pack(insts_b)
    C = spgemm(pa, pb, wide=True)
    crpt, ccol, cval = C.to_host()
    oracle.compare(crpt, ccol, cval, oracle.spgemm(b.A, b.B), 1e-12)


def test_every_tuning_candidate_is_exact(gpu):
    from opsparse_amd import autotune
    from opsparse_amd.spgemm import Workspace
    A = G.rmat(12, 10, 13)
    dA = DeviceCSR.from_host(A, torch.float64)
    ref = oracle.spgemm(A, A)
    ws = Workspace()
    for t in autotune.candidates():
        ws.set_tuning(t)
        assert ws.get_tuning() == t
        rpt, col, val = spgemm(dA, dA, ws=ws).to_host()
        oracle.compare(rpt, col, val, ref, 1e-12)
    bad = autotune.candidates()[0]
    bad["sym_bound"][0] = 10 ** 6  # above the bin-0 table capacity
    with pytest.raises(Exception):
        ws.set_tuning(bad)
    res = autotune.autotune(dA, dA, ws=ws, use_cache=False, reps=1)
    assert res["tuning"] in autotune.candidates() and len(res["trials"]) == len(autotune.candidates())
