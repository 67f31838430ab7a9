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
    return D.row_slice(r0, r1)


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


def test_batch_pack_row_slices_and_empty_instances(gpu):
    """Packing descriptors that are row slices of larger matrices (rpt not
    starting at 0; each instance's entry count is its own rpt[M] - rpt[0]),
    instances with no rows (null rpt, never read) and a descriptor whose nnz
    disagrees with its row offsets (rejected).  The packed product matches
    the oracle instance by instance."""
    from opsparse_amd._lib import OPS_INVALID, OpsError
    from opsparse_amd.spgemm import batch_pack
    rng = np.random.default_rng(31)
    from helpers import random_csr
    parents_a = [random_csr(400, 300, rng.integers(0, 20, 400), rng) for _ in range(3)]
    parents_b = [random_csr(300, 500, rng.integers(0, 25, 300), rng) for _ in range(3)]
    dpa = [DeviceCSR.from_host(X, torch.float64) for X in parents_a]
    dpb = [DeviceCSR.from_host(X, torch.float64) for X in parents_b]
    sl = [(0, 120), (150, 400), (37, 38)]
    ha, hb, da, db = [], [], [], []
    for j, (r0, r1) in enumerate(sl):
        ha.append(parents_a[j].row_slice(r0, r1))
        da.append(dpa[j].row_slice(r0, r1))
        hb.append(parents_b[j])
        db.append(dpb[j])
    # an instance with no rows: A_1 is 0 x 0 and B_1 is 0 x 500 (rpt null)
    empty_a = DeviceCSR(0, 0, torch.empty(0, dtype=torch.int32, device="cuda"),
                        torch.empty(0, dtype=torch.int32, device="cuda"),
                        torch.empty(0, dtype=torch.float64, device="cuda"), 0)
    empty_b = DeviceCSR(0, 500, empty_a.rpt, empty_a.col, empty_a.val, 0)
    assert empty_a.desc().rpt is None or empty_a.desc().rpt == 0
    # instance order: slice 0, the empty one, slice 1, slice 2
    order_a = [da[0], empty_a, da[1], da[2]]
    order_b = [dpb[0].row_slice(0, 300), empty_b, dpb[1], dpb[2]]
    pa, rba, cba = batch_pack(order_a)
    pb, rbb, cbb = batch_pack(order_b)
    assert list(rba) == [0, 120, 120, 370, 371] and list(cba) == [0, 300, 300, 600, 900]
    assert list(rbb) == [0, 300, 300, 600, 900] and list(cbb) == [0, 500, 1000, 1500, 2000]
    assert pa.nnz == sum(h.nnz for h in ha)
    C = spgemm(pa, pb, wide=True)
    crpt, ccol, cval = C.to_host()
    inst_a = [ha[0], None, ha[1], ha[2]]
    inst_b = [hb[0], None, hb[1], hb[2]]
    for i in range(4):
        r0, r1 = int(rba[i]), int(rba[i + 1])
        s0, s1 = crpt[r0], crpt[r1]
        if inst_a[i] is None:
            assert r0 == r1
            continue
        ref = oracle.spgemm(inst_a[i], inst_b[i])
        oracle.compare(crpt[r0:r1 + 1] - s0, ccol[s0:s1] - int(cbb[i]), cval[s0:s1], ref, 1e-12)
    # nnz disagreeing with rpt[M] - rpt[0] is rejected
    bad = DeviceCSR(da[1].M, da[1].N, da[1].rpt, da[1].col, da[1].val, da[1].nnz + 1)
    with pytest.raises(OpsError) as ei:
        batch_pack([da[0], bad])
    assert ei.value.status == OPS_INVALID


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
    res = autotune.autotune(dA, dA, ws=ws, use_cache=False, reps=1, shapes=False)
    assert res["tuning"] in autotune.candidates() and len(res["trials"]) == len(autotune.candidates())
    # with the kernel-shape descent the pick is still exact
    res = autotune.autotune(dA, dA, ws=ws, use_cache=False, reps=1)
    assert len(res["trials"]) > len(autotune.candidates()) and ws.get_tuning() == res["tuning"]
    rpt, col, val = spgemm(dA, dA, ws=ws).to_host()
    oracle.compare(rpt, col, val, ref, 1e-12)


@pytest.mark.parametrize("shape", [1, 2])
def test_every_kernel_shape_is_exact(gpu, shape):
    """Every alternative kernel shape (table size, lanes per row, block
    size) of bins 0-4, symbolic and numeric, FP64 and FP32, with the
    bitmap counting on and off."""
    from opsparse_amd.spgemm import Workspace, default_tuning
    mats = [G.rmat(12, 10, 13), G.stencil27(14, 2)]
    ws = Workspace()
    for bitmap in (0, 1):
        t = default_tuning()
        t["sym_shape"] = [shape] * 5 + [0, 0]
        t["num_shape"] = [shape] * 5 + [0, 0]
        t["sym_bitmap"] = bitmap
        ws.set_tuning(t)
        assert ws.get_tuning() == t
        for A in mats:
            for dt, npdt, tol in ((torch.float64, np.float64, 1e-12), (torch.float32, np.float32, 1e-5)):
                dA = DeviceCSR.from_host(A, dt)
                rpt, col, val = spgemm(dA, dA, ws=ws).to_host()
                Ar = A.astype(npdt) if npdt == np.float32 else A
                oracle.compare(rpt, col, val, oracle.spgemm(Ar, Ar), tol)
