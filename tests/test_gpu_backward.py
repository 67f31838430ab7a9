"""GPU parity of the masked backward (dA on pattern(A), dB on pattern(B)).

Checked against the FP64 CPU oracle gradient with the error normalised by
sum |terms| (the summation-order-independent bound; computed by running the
oracle on absolute values), and against FP64 central differences of the
bilinear loss L = <dC, A*B> (exact up to rounding for any step).
"""
import numpy as np
import pytest
import torch

import oracle
import opsparse_amd.gen as G
from opsparse_amd.spgemm import DeviceCSR, Plan, SpGEMMFunction, backward
from helpers import random_csr

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 1e-5}


def absv(X):
    return G.HostCSR(X.M, X.N, X.rpt, X.col, np.abs(X.val))


def check_grad(A, B, dtype, seed=0):
    rng = np.random.default_rng(seed)
    dA, dB = DeviceCSR.from_host(A, dtype), DeviceCSR.from_host(B, dtype)
    plan = Plan(dA, dB)
    rpt = plan.rowptr64()
    col, _ = plan.numeric(dA, dB)
    dc = rng.normal(size=plan.nnz)
    ga, gb = backward(dA, dB, rpt, col, torch.from_numpy(dc).to(device="cuda", dtype=dtype))
    torch.cuda.synchronize()
    rpt_h, col_h = rpt.cpu().numpy(), col.cpu().numpy()
    if dtype == torch.float32:
        A, B = A.astype(np.float32), B.astype(np.float32)
        dc = dc.astype(np.float32)
    ra, rb = oracle.backward(A, B, rpt_h, col_h, dc)
    sa, sb = oracle.backward(absv(A), absv(B), rpt_h, col_h, np.abs(dc))
    ea = np.abs(ga.double().cpu().numpy() - ra) / np.maximum(sa, 1e-300)
    eb = np.abs(gb.double().cpu().numpy() - rb) / np.maximum(sb, 1e-300)
    err = max(ea.max(initial=0.0), eb.max(initial=0.0))
    assert err <= TOL[dtype], f"gradient error {err:.3e}"
    return err


def test_backward_rmat_fp64_fp32(gpu):
    A = G.rmat(11, 8, 21, normal=True)
    B = G.rmat(11, 8, 22, normal=True)
    check_grad(A, B, torch.float64)
    check_grad(A, B, torch.float32)


def test_backward_stencil_and_long_rows(gpu):
    S = G.stencil27(12, 3)
    check_grad(S, S, torch.float64)
    rng = np.random.default_rng(5)
    N = 20000
    A = random_csr(24, N, np.r_[rng.integers(100, 160, 12), rng.integers(1, 4, 12)], rng)
    B = random_csr(N, N, rng.integers(50, 90, N), rng)
    check_grad(A, B, torch.float64)  # rows of C above 4096 entries use the bitmap-rank windows (k_bwd_rank)


def test_autograd_function_and_finite_differences(gpu):
    A = G.rmat(9, 6, 31, normal=True)
    B = G.rmat(9, 6, 32, normal=True)
    dA, dB = DeviceCSR.from_host(A, torch.float64), DeviceCSR.from_host(B, torch.float64)
    plan = Plan(dA, dB)
    a = dA.val.clone().requires_grad_(True)
    b = dB.val.clone().requires_grad_(True)
    val, col = SpGEMMFunction.apply(a, b, dA, dB, plan)
    w = torch.randn(val.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    loss = (val * w).sum()
    loss.backward()
    ra, rb = oracle.backward(A, B, plan.rowptr64().cpu().numpy(), col.cpu().numpy(), w.cpu().numpy())
    assert np.allclose(a.grad.cpu().numpy(), ra, rtol=1e-11, atol=1e-11)
    assert np.allclose(b.grad.cpu().numpy(), rb, rtol=1e-11, atol=1e-11)

    def L(av, bv):
        _, v = plan.numeric(dA.with_values(av), dB.with_values(bv))
        return float((v * w).sum())

    h = 1e-3
    rng = np.random.default_rng(0)
    for e in rng.choice(A.nnz, 16, replace=False):
        d = torch.zeros_like(dA.val)
        d[e] = h
        fd = (L(dA.val + d, dB.val) - L(dA.val - d, dB.val)) / (2 * h)
        assert abs(fd - a.grad[e].item()) <= 1e-9 * (1 + abs(fd))
    for e in rng.choice(B.nnz, 16, replace=False):
        d = torch.zeros_like(dB.val)
        d[e] = h
        fd = (L(dA.val, dB.val + d) - L(dA.val, dB.val - d)) / (2 * h)
        assert abs(fd - b.grad[e].item()) <= 1e-9 * (1 + abs(fd))


def test_backward_long_rows_skipped_windows(gpu):
    """Rows of C above 4096 entries whose columns sit in two far-apart bands:
    the long-row walk visits only the windows that hold entries."""
    rng = np.random.default_rng(9)
    N = 700000
    A = random_csr(16, N, np.r_[rng.integers(60, 90, 8), rng.integers(1, 3, 8)], rng)
    used = np.unique(A.col)
    lens = np.zeros(N, dtype=np.int64)
    lens[used] = rng.integers(60, 100, used.size)
    B = random_csr(N, 40000, lens, rng)
    # columns [0, 20000) stay, [20000, 40000) move to [600000, 620000): monotone, so rows stay sorted;
    # the 2^18-column windows in between hold no entry of C and are skipped
    col = np.where(B.col < 20000, B.col, B.col + 580000).astype(np.int32)
    B = G.HostCSR(N, N, B.rpt, col, B.val)
    check_grad(A, B, torch.float64)
    check_grad(A, B, torch.float32)


def test_backward_row_slice_descriptors(gpu):
    """A and B given as row-slice descriptors (rpt not starting at 0, col and
    val shared with a larger matrix): dA and dB hold one value per entry of
    each slice, from the slice's first entry, and match the oracle on the
    same slices rebased to 0.  Buffers are exactly the slices' nnz, with
    guard values after them that must stay untouched."""
    rng = np.random.default_rng(41)
    big_a = random_csr(300, 500, np.r_[rng.integers(0, 40, 290), rng.integers(200, 300, 10)], rng)
    big_b = random_csr(900, 60000, rng.integers(0, 90, 900), rng)
    for dtype in (torch.float64, torch.float32):
        npdt = np.float64 if dtype == torch.float64 else np.float32
        Ah_full, Bh_full = big_a.astype(npdt), big_b.astype(npdt)
        dA_full, dB_full = DeviceCSR.from_host(Ah_full, dtype), DeviceCSR.from_host(Bh_full, dtype)
        dA, dB = dA_full.row_slice(40, 300), dB_full.row_slice(200, 700)
        Ah, Bh = Ah_full.row_slice(40, 300), Bh_full.row_slice(200, 700)
        assert dA.nnz == Ah.nnz and dB.nnz == Bh.nnz and int(dA.rpt[0]) > 0
        plan = Plan(dA, dB)
        rpt = plan.rowptr64()
        col, _ = plan.numeric(dA, dB)
        dc = torch.from_numpy(rng.normal(size=plan.nnz)).to(device="cuda", dtype=dtype)
        from opsparse_amd.spgemm import dtype_code, _stream_ptr, default_workspace
        from opsparse_amd._lib import check, lib
        import ctypes as C
        guard = 12345.0
        da = torch.full((dA.nnz + 64,), guard, dtype=dtype, device="cuda")
        db = torch.full((dB.nnz + 64,), guard, dtype=dtype, device="cuda")
        a, b = dA.desc(), dB.desc()
        ws = default_workspace()
        check(lib().ops_spgemm_backward(ws.ptr, dtype_code(dtype), C.byref(a), C.byref(b), rpt.data_ptr(),
                                        col.data_ptr(), dc.data_ptr(), da.data_ptr(), db.data_ptr(),
                                        _stream_ptr()), ws.ptr)
        torch.cuda.synchronize()
        assert torch.all(da[dA.nnz:] == guard) and torch.all(db[dB.nnz:] == guard)
        rpt_h, col_h, dc_h = rpt.cpu().numpy(), col.cpu().numpy(), dc.cpu().numpy()
        ra, rb = oracle.backward(Ah, Bh, rpt_h, col_h, dc_h)
        sa, sb = oracle.backward(absv(Ah), absv(Bh), rpt_h, col_h, np.abs(dc_h))
        ea = np.abs(da[:dA.nnz].double().cpu().numpy() - ra) / np.maximum(sa, 1e-300)
        eb = np.abs(db[:dB.nnz].double().cpu().numpy() - rb) / np.maximum(sb, 1e-300)
        assert max(ea.max(initial=0.0), eb.max(initial=0.0)) <= TOL[dtype]
        # the Python wrapper returns the same slice-sized gradients
        ga, gb = backward(dA, dB, rpt, col, dc)
        assert ga.numel() == dA.nnz and gb.numel() == dB.nnz


def test_random_backward_and_plan_parity(gpu):
    """Seeded random products (short and long rows, narrow and wide spans,
    FP64 and FP32): the plan's numeric phase against the oracle product and
    the masked backward against the oracle gradient."""
    from helpers import check_product
    rng = np.random.default_rng(77)
    for case in range(20):
        K = int(rng.choice([40, 400, 4000, 30000]))
        N = int(rng.choice([64, 3000, 140000, 300000, 1100000]))
        M = int(rng.integers(1, 200))
        la = np.where(rng.random(M) < 0.1, rng.integers(0, min(K, 500) + 1, M), rng.integers(0, min(K, 30) + 1, M))
        lb = np.where(rng.random(K) < 0.05, rng.integers(0, min(N, 2500) + 1, K), rng.integers(0, min(N, 50) + 1, K))
        A, B = random_csr(M, K, la, rng), random_csr(K, N, lb, rng)
        dtype = torch.float64 if rng.random() < 0.5 else torch.float32
        print(f"random backward case {case}: M={M} K={K} N={N} nnz(A)={A.nnz} nnz(B)={B.nnz} {dtype}", flush=True)
        npdt = np.float64 if dtype == torch.float64 else np.float32
        Ah, Bh = (A, B) if npdt == np.float64 else (A.astype(npdt), B.astype(npdt))
        dA, dB = DeviceCSR.from_host(Ah, dtype), DeviceCSR.from_host(Bh, dtype)
        plan = Plan(dA, dB)
        col, val = plan.numeric(dA, dB)
        torch.cuda.synchronize()
        check_product(plan.rowptr64().cpu().numpy(), col.cpu().numpy(), val.cpu().numpy(), A, B, npdt)
        check_grad(A, B, dtype, seed=case)
