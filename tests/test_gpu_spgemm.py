"""GPU parity of the HIP SpGEMM against the CPU oracle.

Structure (row_ptr, sorted columns) must be bit-exact; values must be
within max |d| / sum_k |a_ik b_kj| <= 1e-12 (FP64) or 1e-5 (FP32), the
summation-order-independent error bound.  Every call goes through the C ABI
of libopsparse_amd.so.
"""
import numpy as np
import pytest
import torch

import oracle
import opsparse_amd.gen as G
from opsparse_amd.spgemm import (DeviceCSR, Plan, mem_cache_release, mem_cache_stats, set_mem_cache_limit,
                                  spgemm)
from helpers import TOL, check_product, csr_from_dense, csr_from_rows, random_csr

pytestmark = pytest.mark.gpu


def run(A, B, dtype=np.float64, wide=False):
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    dA = DeviceCSR.from_host(A, tdt)
    dB = DeviceCSR.from_host(B, tdt)
    C = spgemm(dA, dB, wide=wide)
    rpt, col, val = C.to_host()
    return check_product(rpt, col, val, A, B, dtype)


def log(msg):
    print(msg, flush=True)


def test_tiny_dense(gpu):
    rng = np.random.default_rng(11)
    D1 = rng.uniform(-1, 1, (7, 5)) * (rng.random((7, 5)) < 0.6)
    D2 = rng.uniform(-1, 1, (5, 9)) * (rng.random((5, 9)) < 0.6)
    A, B = csr_from_dense(D1), csr_from_dense(D2)
    run(A, B)
    # cross-check the dense triple loop
    ref = oracle.spgemm(A, B)
    dense = np.zeros((7, 9))
    for i in range(7):
        for p in range(ref[0][i], ref[0][i + 1]):
            dense[i, ref[1][p]] = ref[2][p]
    assert np.allclose(dense, D1 @ D2, atol=1e-14)


def test_identity_and_diagonal(gpu):
    rng = np.random.default_rng(1)
    n = 300
    B = random_csr(n, n, rng.integers(0, 40, n), rng)
    I = G.HostCSR(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n))
    run(I, B)
    run(B, I)
    Dg = G.HostCSR(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), rng.uniform(1, 2, n))
    run(Dg, B)


def test_empty_rows_and_empty_product(gpu):
    rng = np.random.default_rng(2)
    n = 200
    lens = rng.integers(0, 6, n)
    lens[::3] = 0
    A = random_csr(n, n, lens, rng)
    B = random_csr(n, n, np.where(np.arange(n) % 2 == 0, 0, 5), rng)
    run(A, B)
    Z = G.HostCSR(n, n, np.zeros(n + 1, dtype=np.int32), np.zeros(0, np.int32), np.zeros(0))
    run(Z, B)
    run(A, Z)


def test_uniform_4096_fp64_fp32(gpu):
    A = G.uniform_rows(4096, 4096, 8, 0)
    log("uniform fp64")
    e = run(A, A, np.float64)
    assert e["max_norm"] <= TOL[np.float64]
    log("uniform fp32")
    run(A, A, np.float32)


@pytest.mark.parametrize("make", ["stencil", "rmat"])
def test_fp32_entries_rounded_once(gpu, make):
    """FP32 inputs are multiplied and summed in FP64 and each entry is
    rounded to FP32 once, so against FP64 sums of the same FP32 inputs the
    error is at most half an FP32 ulp: max |d| / sum |a*b| <= 2^-24 (an FP32
    accumulation would exceed it on rows of many products)."""
    A = G.stencil27(20, 1) if make == "stencil" else G.rmat(12, 16, 3)
    e = run(A, A, np.float32)
    assert e["max_norm"] <= 2.0 ** -24 * 1.0001, e


def test_stencil_small(gpu):
    A = G.stencil27(24, 1)
    run(A, A)


def test_short_bins_narrow_and_wide_rows(gpu):
    """Rows with few products, alternating between a narrow column span
    (banded) and a wide one, in the same launches of bins 0-4; with the
    symbolic bitmap tuning on, the narrow rows are counted in a bitmap and
    the wide ones hashed."""
    rng = np.random.default_rng(15)
    N = 20000
    b_rows = []
    for k in range(N):
        c = np.arange(max(0, k - 40), min(N, k + 41))
        b_rows.append(np.sort(rng.choice(c, size=min(c.size, int(rng.integers(2, 13))), replace=False)))
    B = csr_from_rows(N, N, b_rows, rng)
    a_rows = []
    for i in range(N):
        n = int(rng.integers(1, 31))
        pool = np.arange(max(0, i - 50), min(N, i + 51)) if i % 2 == 0 else np.arange(N)
        a_rows.append(np.sort(rng.choice(pool, size=min(n, pool.size), replace=False)))
    A = csr_from_rows(N, N, a_rows, rng)
    from opsparse_amd.spgemm import Workspace, default_tuning
    dA, dB = DeviceCSR.from_host(A, torch.float64), DeviceCSR.from_host(B, torch.float64)
    for bitmap in (0, 1):
        ws = Workspace()
        ws.set_tuning(dict(default_tuning(), sym_bitmap=bitmap))
        rpt, col, val = spgemm(dA, dB, ws=ws).to_host()
        check_product(rpt, col, val, A, B)
    run(A, B, np.float32)


def test_every_bin_rmat(gpu):
    A = G.rmat(13, 16, 7)
    t = G.compute_flop_host(A, A)
    log(f"rmat13 flop {t}")
    run(A, A)
    run(A, A, np.float32)


def test_long_rows_windows(gpu):
    """Rows with > 16384 products (symbolic long path) and > 4096 entries
    (numeric bitmap-rank windows)."""
    rng = np.random.default_rng(3)
    N = 50000
    A = random_csr(40, N, np.r_[rng.integers(150, 260, 20), rng.integers(1, 4, 20)], rng)
    B = random_csr(N, N, rng.integers(60, 140, N), rng)
    run(A, B)
    run(A, B, np.float32)


def test_symbolic_multi_window(gpu):
    """More than 2^20 columns: the symbolic bitmap runs in several windows."""
    rng = np.random.default_rng(4)
    N = (1 << 21) + 777
    K = 400
    A = random_csr(8, K, rng.integers(100, 200, 8), rng)
    B = random_csr(K, N, rng.integers(150, 300, K), rng)
    run(A, B)


def test_symbolic_long_rows_skip_empty_windows(gpu):
    """B with 2^24 columns (16 bitmap windows) whose columns fall in a few
    windows: the long-row symbolic kernel visits only windows holding a
    product, so its time stays flat as B.N grows from 2^21 to 2^24 (it
    would grow about 8x if every window were walked)."""
    from opsparse_amd._lib import OpsTimings
    rng = np.random.default_rng(23)
    K, M = 400, 96
    A = random_csr(M, K, rng.integers(150, 250, M), rng)
    # B's columns in windows 0 and 1 of 2^20 columns, plus, in the spread
    # case, a few in windows 9 and 15 (hit by a third of B's rows)
    base = random_csr(K, 1 << 21, rng.integers(150, 300, K), rng)
    rows = [base.col[base.rpt[k]:base.rpt[k + 1]] for k in range(K)]
    far = [np.unique(np.r_[r, (9 << 20) + r[:3], (15 << 20) + 5 + r[:2]]) if k % 3 == 0 else r
           for k, r in enumerate(rows)]
    cases = {"narrow": ((1 << 21), rows), "wide": ((1 << 24), rows), "spread": ((1 << 24), far)}
    ms = {}
    for name, (N, rr) in cases.items():
        B = csr_from_rows(K, N, rr, rng)
        dA, dB = DeviceCSR.from_host(A, torch.float64), DeviceCSR.from_host(B, torch.float64)
        best = None
        for _ in range(3):
            t = OpsTimings()
            t.kernel_events = 1
            spgemm(dA, dB, timings=t).free()
            best = t.sym_ms[7] if best is None else min(best, t.sym_ms[7])
        assert t.sym_rows[7] == M, list(t.sym_rows)
        ms[name] = best
        run(A, B)
    log(f"long-row symbolic ms by case: {ms}")
    assert ms["wide"] <= 2.0 * ms["narrow"] + 0.05, ms
    assert ms["spread"] <= 3.0 * ms["narrow"] + 0.05, ms


def test_symbolic_bin6_hash_or_bitmap(gpu):
    """Rows of 4097..16384 products: counted by the bin-6 hash kernel when B
    has more columns than one long-row bitmap window (2^20), by the bitmap
    kernel (bin 7) otherwise."""
    from opsparse_amd._lib import OpsTimings
    rng = np.random.default_rng(17)
    K = 400
    A = random_csr(16, K, rng.integers(30, 40, 16), rng)
    for N in ((1 << 21) + 5, 300000):
        B = random_csr(K, N, rng.integers(150, 300, K), rng)
        dA, dB = DeviceCSR.from_host(A, torch.float64), DeviceCSR.from_host(B, torch.float64)
        t = OpsTimings()
        spgemm(dA, dB, timings=t).free()
        if N > (1 << 20):
            assert t.sym_rows[6] > 0, list(t.sym_rows)
        else:
            assert t.sym_rows[6] == 0 and t.sym_rows[7] > 0, list(t.sym_rows)
        run(A, B)


def test_wide_rowptr(gpu):
    A = G.rmat(11, 8, 5)
    run(A, A, wide=True)


@pytest.mark.parametrize("wide", [False, True])
def test_rows_to_host_matches_full_copy(gpu, wide):
    A = G.rmat(10, 8, 6)
    dA = DeviceCSR.from_host(A, torch.float64)
    C = spgemm(dA, dA, wide=wide)
    rpt, col, val = C.to_host()
    for r0, r1 in ((0, 1), (0, A.M), (17, 300), (A.M - 5, A.M), (40, 40)):
        hr, hc, hv = C.rows_to_host(r0, r1)
        s0, s1 = rpt[r0], rpt[r1]
        assert np.array_equal(hr, rpt[r0:r1 + 1] - s0)
        assert np.array_equal(hc, col[s0:s1]) and np.array_equal(hv, val[s0:s1])
    with pytest.raises(ValueError):
        C.rows_to_host(5, A.M + 1)


def test_plan_and_graph(gpu):
    A = G.rmat(12, 12, 9)
    dA = DeviceCSR.from_host(A, torch.float64)
    p = Plan(dA, dA)
    ref = oracle.spgemm(A, A)
    rpt = p.rowptr64().cpu().numpy()
    col, val = p.numeric(dA, dA)
    oracle.compare(rpt, col.cpu().numpy(), val.cpu().numpy(), ref, TOL[np.float64])
    col_g = torch.empty(p.nnz, dtype=torch.int32, device="cuda")
    val_g = torch.empty(p.nnz, dtype=torch.float64, device="cuda")
    for rep in range(3):
        val_g.fill_(7.0)
        p.numeric(dA, dA, col_g, val_g, graph=True)
        torch.cuda.synchronize()
        oracle.compare(rpt, col_g.cpu().numpy(), val_g.cpu().numpy(), ref, TOL[np.float64])
        # new values, same pattern, replay of the same graph
    A2 = G.HostCSR(A.M, A.N, A.rpt, A.col, A.val * 2.0)
    dA2 = DeviceCSR.from_host(A2, torch.float64)
    dA.val.copy_(dA2.val)
    p.numeric(dA, dA, col_g, val_g, graph=True)
    torch.cuda.synchronize()
    oracle.compare(rpt, col_g.cpu().numpy(), val_g.cpu().numpy(), oracle.spgemm(A2, A2), TOL[np.float64])


def test_plan_graph_long_spill_and_short_rows(gpu):
    """A plan whose numeric phase runs the long-row kernel over several
    windows (with B's window-cut table), the wave kernel of bin 5, its spill
    list into the block kernel, and short bins, eagerly and as a replayed
    hipGraph (values changed between replays)."""
    rng = np.random.default_rng(16)
    W = 1 << 18
    N = 5 * W + 7
    Kw, Kn = 1100, 100
    wide = random_csr(Kw, 3 * W, rng.integers(250, 330, Kw), rng)
    col = wide.col.astype(np.int64)
    col = np.array([0, 2, 5], dtype=np.int64)[col // W] * W + col % W
    rows = [np.minimum(col[wide.rpt[i]:wide.rpt[i + 1]], N - 1) for i in range(Kw)]
    rows += [np.sort(rng.choice(np.arange(1000, 9000), size=int(rng.integers(60, 100)), replace=False))
             for _ in range(Kn)]
    B = csr_from_rows(Kw + Kn, N, rows, rng)
    a_rows = ([np.sort(rng.choice(Kw, size=int(rng.integers(300, 420)), replace=False)) for _ in range(3)]
              + [np.sort(rng.choice(np.arange(Kw, Kw + Kn), size=int(rng.integers(6, 10)), replace=False))
                 for _ in range(3)]
              + [np.sort(rng.choice(Kw, size=int(rng.integers(2, 4)), replace=False)) for _ in range(3)]
              + [np.array([int(rng.integers(0, Kw))]) for _ in range(3)])
    A = csr_from_rows(len(a_rows), Kw + Kn, a_rows, rng)
    dA, dB = DeviceCSR.from_host(A, torch.float64), DeviceCSR.from_host(B, torch.float64)
    from opsparse_amd._lib import OpsTimings
    t = OpsTimings()
    spgemm(dA, dB, timings=t).free()
    assert t.num_rows[7] == 3 and t.num_rows[5] >= 3 and sum(t.num_rows[:5]) >= 3, list(t.num_rows)
    p = Plan(dA, dB)
    rpt = p.rowptr64().cpu().numpy()
    ref = oracle.spgemm(A, B)
    col_e, val_e = p.numeric(dA, dB)
    oracle.compare(rpt, col_e.cpu().numpy(), val_e.cpu().numpy(), ref, TOL[np.float64])
    col_g = torch.empty(p.nnz, dtype=torch.int32, device="cuda")
    val_g = torch.empty(p.nnz, dtype=torch.float64, device="cuda")
    for scale in (1.0, -0.5):
        A2 = G.HostCSR(A.M, A.N, A.rpt, A.col, A.val * scale)
        dA.val.copy_(DeviceCSR.from_host(A2, torch.float64).val)
        for _ in range(2):
            val_g.fill_(7.0)
            p.numeric(dA, dB, col_g, val_g, graph=True)
            torch.cuda.synchronize()
            oracle.compare(rpt, col_g.cpu().numpy(), val_g.cpu().numpy(), oracle.spgemm(A2, B), TOL[np.float64])


def test_output_cache_reuse_and_limit(gpu):
    A = G.rmat(12, 10, 8)
    dA = DeviceCSR.from_host(A, torch.float64)
    ref = oracle.spgemm(A, A)
    mem_cache_release()
    before = mem_cache_stats()
    for wide in (False, False, True, True):
        C = spgemm(dA, dA, wide=wide)
        rpt, col, val = C.to_host()
        oracle.compare(rpt, col, val, ref, TOL[np.float64])
        C.free()
    st = mem_cache_stats()
    assert st["hits"] - before["hits"] >= 4  # the second call of each width reuses col, val (and rpt)
    assert st["cached_bytes"] > 0 and st["live_bytes"] <= before["live_bytes"]
    set_mem_cache_limit(0)  # caching off: cached blocks are freed, releases go to hipFree
    assert mem_cache_stats()["cached_bytes"] == 0
    C = spgemm(dA, dA)
    C.free()
    assert mem_cache_stats()["cached_bytes"] == 0
    set_mem_cache_limit(-1)
    mem_cache_release()


def test_output_cache_keeps_recent_blocks(gpu):
    """With a cache limit that holds either of two products' C but not both,
    releasing one product's arrays evicts the other's (least recently
    released) blocks, so repeating the product in use hits the cache."""
    A1, A2 = G.rmat(12, 10, 8), G.rmat(11, 8, 5)
    d1, d2 = DeviceCSR.from_host(A1, torch.float64), DeviceCSR.from_host(A2, torch.float64)
    set_mem_cache_limit(-1)
    mem_cache_release()
    spgemm(d1, d1).free()
    b1 = mem_cache_stats()["cached_bytes"]
    mem_cache_release()
    spgemm(d2, d2).free()
    b2 = mem_cache_stats()["cached_bytes"]
    assert b1 > b2 > 0
    limit = b1 + b2 // 2
    set_mem_cache_limit(limit)
    spgemm(d2, d2).free()
    spgemm(d1, d1).free()  # misses; releasing it evicts the blocks of the first product
    st = mem_cache_stats()
    assert st["cached_bytes"] <= limit
    C = spgemm(d1, d1)
    rpt, col, val = C.to_host()
    oracle.compare(rpt, col, val, oracle.spgemm(A1, A1), TOL[np.float64])
    C.free()
    assert mem_cache_stats()["hits"] - st["hits"] >= 3  # rpt, col and val
    assert mem_cache_stats()["cached_bytes"] <= limit
    set_mem_cache_limit(-1)
    mem_cache_release()


@pytest.mark.parametrize("unit_cuts", ["1", "0"])
def test_long_rows_many_rank_subwindows(gpu, monkeypatch, unit_cuts):
    """Rows of ~90K entries over two 2^17-column bitmap windows: the
    accumulate pass runs over several value sub-windows per window (12288
    entries each, FP64 accumulators for both value types), cut at the
    column of each rank boundary; each sub-window's end in a B row is found
    between the cuts of its 2^14-column unit, or (unit cut table off) by
    galloping."""
    monkeypatch.setenv("OPSAMD_UNIT_CUTS", unit_cuts)
    rng = np.random.default_rng(12)
    N = 200000
    A = random_csr(6, 1000, np.r_[rng.integers(350, 450, 4), rng.integers(1, 3, 2)], rng)
    B = random_csr(1000, N, rng.integers(250, 350, 1000), rng)
    run(A, B)
    run(A, B, np.float32)
    run(A, B, wide=True)


@pytest.mark.parametrize("N", [90000, 200000])
def test_long_rows_kept_bitmaps(gpu, monkeypatch, N):
    """Long rows whose column bitmaps the symbolic phase keeps for the
    numeric phase (B within 2^20 columns): all rows kept, none kept (every
    window marked again) and only two kept (the rest marked), over one
    bitmap window (whole B rows) and two (window-cut table), with one, two
    and more value sub-windows per window; through the drop-in call and
    through a plan (eager and replayed hipGraph)."""
    rng = np.random.default_rng(18)
    K = 1000
    lens = np.r_[rng.integers(350, 450, 3), rng.integers(120, 160, 3), rng.integers(40, 60, 3),
                 rng.integers(1, 3, 2)]
    A = random_csr(len(lens), K, lens, rng)
    B = random_csr(K, N, rng.integers(250, 350, K), rng)
    ref = oracle.spgemm(A, B)
    from opsparse_amd._lib import OpsTimings
    n_long = None
    for kept in ("", "0", "2"):
        if kept:
            monkeypatch.setenv("OPSAMD_KEPT_BITMAPS", kept)
        log(f"kept bitmaps N={N} limit={kept or 'none'}")
        run(A, B)
        run(A, B, np.float32)
        dA, dB = DeviceCSR.from_host(A, torch.float64), DeviceCSR.from_host(B, torch.float64)
        t = OpsTimings()
        spgemm(dA, dB, timings=t).free()
        if not kept:
            n_long = int(t.num_rows[7])
            assert n_long >= 3 and t.kept_rows == n_long, (list(t.num_rows), t.kept_rows)
        else:
            assert t.kept_rows == min(int(kept), n_long), (kept, t.kept_rows)
        # a plan keeps its own bitmaps for its eager and replayed numeric phases
        p = Plan(dA, dB)
        assert p.kept_rows == t.kept_rows
        rpt = p.rowptr64().cpu().numpy()
        col, val = p.numeric(dA, dB)
        oracle.compare(rpt, col.cpu().numpy(), val.cpu().numpy(), ref, TOL[np.float64])
        for _ in range(2):
            val.fill_(7.0)
            p.numeric(dA, dB, col, val, graph=True)
            torch.cuda.synchronize()
            oracle.compare(rpt, col.cpu().numpy(), val.cpu().numpy(), ref, TOL[np.float64])
        del p


def test_long_rows_multi_window_cut_table_and_gallop(gpu, monkeypatch):
    """Long rows whose columns lie in blocks 0, 2 and 5 of six 2^18-column
    blocks: the bitmap windows (2^17 columns) in between are empty and
    skipped, and the filled windows need several value sub-windows.  Run with
    the window-cut table of B and its unit cuts (default), with the table
    alone, and with the table turned off (window boundaries found by
    galloping in each B row)."""
    rng = np.random.default_rng(14)
    W = 1 << 18
    N = 5 * W + 7
    K = 1200
    # columns drawn over 3 windows' worth, then placed in windows 0, 2 and 5
    lens = np.r_[rng.integers(250, 330, K - 8), rng.integers(1, 4, 8)]
    B3 = random_csr(K, 3 * W, lens, rng)
    col = B3.col.astype(np.int64)
    wmap = np.array([0, 2, 5], dtype=np.int64)
    col = wmap[col // W] * W + col % W
    col = np.minimum(col, N - 1)
    B = G.HostCSR(K, N, B3.rpt.copy(), col.astype(np.int32), B3.val.copy())
    A = random_csr(5, K, np.r_[rng.integers(300, 420, 3), 1, 0], rng)
    # Rows with a dense first window: B rows put half their entries in
    # columns [0, 16384) and the rest anywhere in 3 * 2^17 columns, so C's
    # long rows hold about 20K columns in window 0 (two value passes) and
    # sparse windows beyond.
    Kd, Nd = 900, 3 * (1 << 17)
    dl = rng.integers(180, 240, Kd)
    rows = []
    for i in range(Kd):
        h = int(dl[i]) // 2
        c = np.union1d(rng.choice(16384, h, replace=False), rng.choice(Nd, int(dl[i]) - h, replace=False))
        rows.append(c)
    drpt = np.zeros(Kd + 1, dtype=np.int32)
    drpt[1:] = np.cumsum([len(c) for c in rows])
    Bd = G.HostCSR(Kd, Nd, drpt, np.concatenate(rows).astype(np.int32), rng.uniform(-1, 1, int(drpt[-1])))
    Ad = random_csr(4, Kd, np.r_[rng.integers(200, 260, 2), 30, 1], rng)
    # cut table with its unit cuts (windows of more than one value pass
    # staged per sub-window), the cut table alone (two-pass windows walked
    # twice over the same ranges) with every row in the 1024-thread shape,
    # no table (galloping)
    for mode, units, light in (("1", "1", "1"), ("1", "0", "0"), ("0", "1", "1")):
        monkeypatch.setenv("OPSAMD_CUT_TABLE", mode)
        monkeypatch.setenv("OPSAMD_UNIT_CUTS", units)
        monkeypatch.setenv("OPSAMD_LIGHT_ROWS", light)
        log(f"multi-window cut table={mode} unit cuts={units} light rows={light}")
        run(A, B)
        run(A, B, np.float32)
        run(A, B, wide=True)
        run(Ad, Bd)
        run(Ad, Bd, np.float32)


def test_block_bins_rank_and_hash_paths(gpu):
    """Rows of 256..4096 entries: a narrow column span takes the bitmap-rank
    path of the block bins, a span over 2^21 columns the hash + sort path."""
    rng = np.random.default_rng(13)
    for N in (40000, (1 << 21) + 5):
        K = 3000
        A = random_csr(96, K, np.r_[rng.integers(2, 7, 48), rng.integers(15, 26, 48)], rng)
        B = random_csr(K, N, rng.integers(80, 150, K), rng)
        log(f"block bins N={N}")
        run(A, B)
        run(A, B, np.float32)


def test_random_products_parity(gpu):
    """Seeded random products mixing short and long A and B rows, narrow and
    wide column ranges (up to 2.2M columns), FP64 and FP32, 32- and 64-bit
    row offsets and all symbolic tunings, each against the oracle."""
    from opsparse_amd.spgemm import Workspace, default_tuning
    rng = np.random.default_rng(2024)
    for case in range(40):
        K = int(rng.choice([50, 500, 5000, 40000]))
        N = int(rng.choice([64, 3000, 70000, 140000, 300000, 1100000, 2200000]))
        M = int(rng.integers(1, 300))
        la = np.where(rng.random(M) < 0.1, rng.integers(0, min(K, 600) + 1, M), rng.integers(0, min(K, 40) + 1, M))
        lb = np.where(rng.random(K) < 0.05, rng.integers(0, min(N, 3000) + 1, K), rng.integers(0, min(N, 60) + 1, K))
        A, B = random_csr(M, K, la, rng), random_csr(K, N, lb, rng)
        dt = np.float64 if rng.random() < 0.5 else np.float32
        tdt = torch.float64 if dt == np.float64 else torch.float32
        wide = bool(rng.random() < 0.3)
        tun = dict(default_tuning(), sym_bitmap=int(rng.random() < 0.5), short_hash=int(rng.random() < 0.5))
        log(f"random case {case}: M={M} K={K} N={N} nnz(A)={A.nnz} nnz(B)={B.nnz} {dt.__name__} wide={wide} "
            f"bitmap={tun['sym_bitmap']} hash={tun['short_hash']}")
        Ah, Bh = (A, B) if dt == np.float64 else (A.astype(dt), B.astype(dt))
        ws = Workspace()
        ws.set_tuning(tun)
        C = spgemm(DeviceCSR.from_host(Ah, tdt), DeviceCSR.from_host(Bh, tdt), ws=ws, wide=wide)
        rpt, col, val = C.to_host()
        check_product(rpt, col, val, A, B, dt)
        C.free()
