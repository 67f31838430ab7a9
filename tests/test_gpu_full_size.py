"""GPU parity at the full sizes of the five benchmark workloads.

  1. 4096 x 4096, 8 uniform columns per row, FP64 (test_gpu_spgemm.py)
  2. 27-point stencil on 100^3, FP64 and FP32: the whole product against
     the oracle
  3. R-MAT 2^20 A*A, FP64, 64-bit row offsets (nnz(C) ~ 9.7e9): every row
     offset against the oracle's symbolic count, columns and values of the
     highest-degree rows and of row blocks whose offsets lie beyond 2^31 and
     2^32 against the oracle; the 32-bit entry point refuses the product
     after its symbolic phase; the device product count against the oracle
  4. 4096-instance batch (FP64 and FP32), one fused call: every row offset
     against the oracle, sampled instances' columns and values against the
     oracle, and a sample of instances against the plain call looped over
     them
  5. R-MAT 2^18 A*B, FP32 forward and masked backward: row offsets and
     sampled rows against the oracle; FP64 gradients against central
     differences through the forward on 1024 sampled entries and against
     the oracle's gradient on the highest-degree rows; FP32 gradients
     against FP64 sums

Structure is bit-exact everywhere; values within max |d| / sum |a*b| of
1e-12 (FP64) or 1e-5 (FP32); gradients within 1e-10 (FP64) of the same
normalisation.  The output cache and the workspaces' buffers are released
after every workload, so each starts with the device's memory free.
"""
import numpy as np
import pytest
import torch

import oracle
import opsparse_amd.gen as G
from opsparse_amd._lib import OPS_INT32_OVERFLOW, OpsError
from opsparse_amd.spgemm import (DeviceCSR, Plan, Workspace, backward, compute_flop_device, mem_cache_release,
                                  mem_cache_stats, spgemm)
from helpers import TOL

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def release(ws=None):
    torch.cuda.synchronize()
    if ws is not None:
        ws.trim()
    mem_cache_release()
    torch.cuda.empty_cache()


def check_rows(C, A, B, r0, r1, tol):
    ref = oracle.spgemm_rows(A, B, r0, r1)
    rpt, col, val = C.rows_to_host(r0, r1)
    return oracle.compare(rpt, col, val, ref, tol)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_stencil_100_full(gpu, dtype):
    npdt = np.float64 if dtype == torch.float64 else np.float32
    A = G.stencil27(100, 1)
    Ah = A if npdt == np.float64 else A.astype(npdt)
    ws = Workspace()
    dA = DeviceCSR.from_host(Ah, dtype)
    log(f"stencil 100^3 {dtype}: product")
    C = spgemm(dA, dA, ws=ws)
    assert C.nnz == 120553784
    rpt, col, val = C.to_host()
    log("stencil: oracle")
    e = oracle.compare(rpt, col, val, oracle.spgemm(Ah, Ah), TOL[npdt])
    log(f"stencil {dtype}: {e}")
    C.free()
    release(ws)


def test_rmat20_int64_offsets(gpu):
    A = G.rmat(20, 16, 2)
    ws = Workspace()
    dA = DeviceCSR.from_host(A, torch.float64)
    # device product count against the oracle, per row and in total
    rf, tot = compute_flop_device(dA, dA, ws=ws)
    P, rf_ref = oracle.compute_flop(A, A, per_row=True)
    assert tot == P and np.array_equal(rf.cpu().numpy().astype(np.int64), rf_ref), (tot, P)
    log(f"rmat20: products {P}")
    # the 32-bit entry point refuses nnz(C) > INT32_MAX after its symbolic
    # phase and leaves nothing allocated
    live = mem_cache_stats()["live_bytes"]
    with pytest.raises(OpsError) as ei:
        spgemm(dA, dA, ws=ws)
    assert ei.value.status == OPS_INT32_OVERFLOW
    assert mem_cache_stats()["live_bytes"] == live
    log("rmat20: 64-bit product")
    C = spgemm(dA, dA, ws=ws, wide=True)
    log(f"rmat20: nnz(C) {C.nnz}")
    # every row offset against the oracle's symbolic count
    ref_rpt = oracle.symbolic(A, A)
    full = C.rowptr_to_host()
    assert C.nnz == int(ref_rpt[-1]) > 2 ** 32
    assert np.array_equal(full, ref_rpt)
    # columns and values: the hub rows and blocks past 2^31 and 2^32
    blocks = [(0, 512)]
    for off in (2 ** 31 + 12345, 2 ** 32 + 777, int(ref_rpt[-1]) - 1):
        r = int(np.searchsorted(ref_rpt, off, side="right")) - 1
        r = min(max(r - 8, 0), A.M - 1024)
        blocks.append((r, r + 1024))
    for r0, r1 in blocks:
        e = check_rows(C, A, A, r0, r1, TOL[np.float64])
        log(f"rmat20 rows [{r0},{r1}) offsets from {int(ref_rpt[r0])}: {e}")
    assert any(ref_rpt[r0] > 2 ** 32 for r0, _ in blocks)
    C.free()
    release(ws)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_4096_fused(gpu, dtype):
    npdt = np.float64 if dtype == torch.float64 else np.float32
    b = G.batch(4096, 3, 64, 16384, 4, 32)
    A, B = b.A.astype(npdt), b.B.astype(npdt)
    ws = Workspace()
    dA, dB = DeviceCSR.from_host(A, dtype), DeviceCSR.from_host(B, dtype)
    log(f"batch 4096 {dtype}: fused call, {A.M} rows, {oracle.compute_flop(A, B)} products")
    C = spgemm(dA, dB, ws=ws, wide=True)
    full = C.rowptr_to_host()
    ref_rpt = oracle.symbolic(A, B)
    assert np.array_equal(full, ref_rpt) and C.nnz == int(ref_rpt[-1])
    rng = np.random.default_rng(3)
    sample = sorted(set([0, 1, 2047, 4095] + list(rng.choice(4096, 28, replace=False))))
    for i in sample:
        r0, r1 = int(b.row_base[i]), int(b.row_base[i + 1])
        check_rows(C, A, B, r0, r1, TOL[npdt])
    # fused = looped plain call on a sample (row-slice descriptors of the packed A)
    for i in sample[:12]:
        r0, r1 = int(b.row_base[i]), int(b.row_base[i + 1])
        Ci = spgemm(dA.row_slice(r0, r1), dB, ws=ws)
        rpt, col, val = Ci.to_host()
        frpt, fcol, fval = C.rows_to_host(r0, r1)
        assert np.array_equal(rpt, frpt) and np.array_equal(col, fcol)
        scale = np.maximum(np.abs(fval).astype(np.float64), 1.0)
        assert np.max(np.abs(val.astype(np.float64) - fval) / scale, initial=0.0) <= 10 * TOL[npdt]
        Ci.free()
    log(f"batch {dtype}: {len(sample)} instances checked")
    C.free()
    release(ws)


def test_rmat18_forward_backward_fd(gpu):
    from opsparse_amd.gradcheck import gradient_checks
    A = G.rmat(18, 12, 4, normal=True)
    B = G.rmat(18, 12, 1004, normal=True)
    ws = Workspace()
    A32, B32 = A.astype(np.float32), B.astype(np.float32)
    dA, dB = DeviceCSR.from_host(A32, torch.float32), DeviceCSR.from_host(B32, torch.float32)
    log("rmat18: FP32 forward")
    C = spgemm(dA, dB, ws=ws, wide=True)
    assert np.array_equal(C.rowptr_to_host(), oracle.symbolic(A, B))
    for r0, r1 in ((0, 2048), (100000, 102048), (A.M - 2048, A.M)):
        check_rows(C, A32, B32, r0, r1, TOL[np.float32])
    C.free()
    plan = Plan(dA, dB, ws=ws)
    rpt = plan.rowptr64()
    col, _ = plan.numeric(dA, dB)
    g = torch.Generator(device="cuda").manual_seed(4)
    dc = torch.randn(plan.nnz, dtype=torch.float32, device="cuda", generator=g)
    log("rmat18: FP32 and FP64 backward")
    ga, gb = backward(dA, dB, rpt, col, dc, ws=ws)
    dA64, dB64 = DeviceCSR.from_host(A, torch.float64), DeviceCSR.from_host(B, torch.float64)
    dc64 = dc.double()
    ga64, gb64 = backward(dA64, dB64, rpt, col, dc64, ws=ws)
    torch.cuda.synchronize()
    crpt_h, ccol_h, dc_h = rpt.cpu().numpy(), col.cpu().numpy(), dc64.cpu().numpy()
    log("rmat18: finite differences on 1024 entries")
    chk = gradient_checks(A, B, dA64, dB64, crpt_h, ccol_h, dc_h, ga, gb, ga64, gb64, 1024, ws)
    log(f"rmat18 gradient checks: {chk}")
    assert chk["fd64"] <= 1e-10 and chk["ref64"] <= 1e-12 and chk["ref32"] <= 1e-5
    # entries with no terms (an empty B row, or a B row no A column feeds):
    # the gradient is exactly 0; the difference is 0 up to the rounding of
    # the two perturbed rows' other terms, summed in different orders
    assert chk["zero_max_abs_grad"] == 0.0 and chk["zero_max_abs_fd"] <= 1e-12
    # oracle dA on the 4096 highest-degree rows (those rows of dA depend on
    # those rows of C only)
    s0, s1 = int(crpt_h[0]), int(crpt_h[4096])
    Ar = A.row_slice(0, 4096)
    ra, _ = oracle.backward(Ar, B, crpt_h[:4097] - s0, ccol_h[s0:s1], dc_h[s0:s1])
    sa, _ = oracle.backward(G.HostCSR(Ar.M, Ar.N, Ar.rpt, Ar.col, np.abs(Ar.val)),
                            G.HostCSR(B.M, B.N, B.rpt, B.col, np.abs(B.val)), crpt_h[:4097] - s0, ccol_h[s0:s1],
                            np.abs(dc_h[s0:s1]))
    a1 = int(A.rpt[4096])
    sc = np.maximum(sa, 1e-300)
    e64 = float(np.max(np.abs(ga64[:a1].cpu().numpy() - ra) / sc, initial=0.0))
    e32 = float(np.max(np.abs(ga[:a1].double().cpu().numpy() - ra) / sc, initial=0.0))
    log(f"rmat18 oracle dA rows 0..4095: fp64 {e64:.2e} fp32 {e32:.2e}")
    assert e64 <= 1e-12 and e32 <= 1e-5
    del plan
    release(ws)
