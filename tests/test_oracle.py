"""Pins the CPU oracle (oracle/gustavson.c) before anything is checked against it.

The reference ships no tests, golden vectors or fixtures for this path, so
the oracle is pinned by independent computations:
  * scipy.sparse products on positive values (no cancellation, so scipy's
    dropping of numerical zeros cannot hide entries): structure and values;
  * a dense triple loop (numpy matmul) on tiny inputs with signed values;
  * special cases: identity, diagonal scaling, permutations;
  * closed forms of the two headline workloads (products and nnz(C)) and the
    exact entries of the noise-free 27-point stencil squared;
  * the masked backward against dense matmuls and FP64 central differences
    (the loss <dC, A*B> is bilinear, so the difference is exact up to rounding);
  * planted mistakes, to show the comparison rejects them.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import opsparse_amd.gen as G
from helpers import csr_from_dense, random_csr


def to_scipy(A, absval=False):
    v = np.abs(A.val) if absval else A.val
    return sp.csr_matrix((v, A.col, A.rpt), shape=(A.M, A.N))


def dense_of(rpt, col, val, M, N):
    D = np.zeros((M, N))
    for i in range(M):
        for p in range(rpt[i], rpt[i + 1]):
            D[i, col[p]] += val[p]
    return D


@pytest.mark.parametrize("shape", [(50, 70, 40), (300, 300, 300), (1, 1000, 5), (257, 31, 1029)])
def test_matches_scipy_positive(shape):
    rng = np.random.default_rng(sum(shape))
    M, K, N = shape
    A = random_csr(M, K, rng.integers(0, min(K, 20) + 1, M), rng)
    B = random_csr(K, N, rng.integers(0, min(N, 30) + 1, K), rng)
    A.val, B.val = np.abs(A.val) + 0.1, np.abs(B.val) + 0.1
    rpt, col, val, ab = oracle.spgemm(A, B)
    P = (to_scipy(A) @ to_scipy(B)).tocsr()
    P.sort_indices()
    assert np.array_equal(rpt, P.indptr)
    assert np.array_equal(col, P.indices)
    assert np.allclose(val, P.data, rtol=1e-13, atol=0)
    assert np.allclose(ab, val, rtol=1e-13)  # all products positive


def test_dense_triple_loop_signed_and_structural_zero():
    rng = np.random.default_rng(5)
    for _ in range(20):
        D1 = rng.uniform(-1, 1, (9, 6)) * (rng.random((9, 6)) < 0.5)
        D2 = rng.uniform(-1, 1, (6, 8)) * (rng.random((6, 8)) < 0.5)
        A, B = csr_from_dense(D1), csr_from_dense(D2)
        rpt, col, val, _ = oracle.spgemm(A, B)
        assert np.allclose(dense_of(rpt, col, val, 9, 8), D1 @ D2, atol=1e-15)
        mask = (np.abs(D1) > 0).astype(float) @ (np.abs(D2) > 0).astype(float) > 0
        assert int(rpt[-1]) == int(mask.sum())
    # exact cancellation stays a structural entry with value 0.0
    A = csr_from_dense(np.array([[1.0, 1.0]]))
    B = csr_from_dense(np.array([[2.0], [-2.0]]))
    rpt, col, val, ab = oracle.spgemm(A, B)
    assert list(rpt) == [0, 1] and list(col) == [0] and val[0] == 0.0 and ab[0] == 4.0


def test_identity_diagonal_permutation():
    rng = np.random.default_rng(6)
    n = 120
    B = random_csr(n, n, rng.integers(0, 15, n), rng)
    I = G.HostCSR(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n))
    for X, Y in ((I, B), (B, I)):
        rpt, col, val, _ = oracle.spgemm(X, Y)
        assert np.array_equal(rpt, B.rpt) and np.array_equal(col, B.col) and np.array_equal(val, B.val)
    d = rng.uniform(1, 3, n)
    Dg = G.HostCSR(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d)
    rpt, col, val, _ = oracle.spgemm(Dg, B)
    assert np.array_equal(col, B.col)
    assert np.array_equal(val, B.val * np.repeat(d, np.diff(B.rpt)))
    perm = rng.permutation(n)
    Pm = G.HostCSR(n, n, np.arange(n + 1, dtype=np.int32), perm.astype(np.int32), np.ones(n))
    rpt, col, val, _ = oracle.spgemm(Pm, B)  # row i of P*B is row perm[i] of B
    ref = to_scipy(B)[perm]
    assert np.array_equal(rpt, ref.indptr) and np.array_equal(col, ref.indices) and np.array_equal(val, ref.data)


def test_compute_flop_brute_force():
    rng = np.random.default_rng(7)
    A = random_csr(60, 80, rng.integers(0, 12, 60), rng)
    B = random_csr(80, 50, rng.integers(0, 9, 80), rng)
    tot, per = oracle.compute_flop(A, B, per_row=True)
    brute = [sum(B.rpt[k + 1] - B.rpt[k] for k in A.col[A.rpt[i]:A.rpt[i + 1]]) for i in range(A.M)]
    assert list(per) == brute and tot == sum(brute)
    assert tot == G.compute_flop_host(A, B)


def test_closed_form_uniform_4096():
    A = G.uniform_rows(4096, 4096, 8, 0)
    assert oracle.compute_flop(A, A) == 4096 * 8 * 8 == 262144


def test_closed_form_stencil_100():
    A = G.stencil27(100, 1)
    assert A.nnz == (2 * 2 + 98 * 3) ** 3 == 26463592
    assert oracle.compute_flop(A, A) == (2 * 2 ** 2 + 98 * 3 ** 2) ** 3 == 704969000
    rpt = np.zeros(A.M + 1, dtype=np.int64)
    nnz = oracle.lib().oracle_symbolic(A.M, A.N, A.rpt.ctypes.data, A.col.ctypes.data, A.rpt.ctypes.data,
                                       A.col.ctypes.data, rpt.ctypes.data)
    assert nnz == (2 * 3 + 2 * 4 + 96 * 5) ** 3 == 120553784


def stencil_square_entry(i, j, n):
    """Exact (A*A)[i,j] for the noise-free 27-point stencil (diag 26, off -1)
    on an n^3 grid, from per-axis common-neighbour counts."""
    xi, yi, zi = i % n, (i // n) % n, i // (n * n)
    xj, yj, zj = j % n, (j // n) % n, j // (n * n)
    common = 1
    for a, b in ((xi, xj), (yi, yj), (zi, zj)):
        lo, hi = max(a, b) - 1, min(a, b) + 1
        common *= max(0, min(hi, n - 1) - max(lo, 0) + 1)
    if i == j:
        return 26.0 * 26.0 + (common - 1)
    adjacent = max(abs(xi - xj), abs(yi - yj), abs(zi - zj)) <= 1
    if adjacent:
        return -26.0 - 26.0 + (common - 2)
    return float(common)


def test_stencil_square_known_answer():
    n = 7
    A = G.stencil27(n, 1, noise=0.0)
    rpt, col, val, _ = oracle.spgemm(A, A)
    for i in range(A.M):
        for p in range(rpt[i], rpt[i + 1]):
            assert val[p] == stencil_square_entry(i, int(col[p]), n)
    # row nnz is the product of per-axis reach-2 counts
    for i in range(A.M):
        x, y, z = i % n, (i // n) % n, i // (n * n)
        cnt = 1
        for c in (x, y, z):
            cnt *= min(c + 2, n - 1) - max(c - 2, 0) + 1
        assert rpt[i + 1] - rpt[i] == cnt


def test_symbolic_matches_scipy_pattern_and_closed_form():
    """oracle.symbolic (row offsets only) against scipy on all-ones values
    (no cancellation: the pattern is the structural one) and against the
    stencil's per-row closed form."""
    rng = np.random.default_rng(19)
    A = random_csr(400, 300, rng.integers(0, 30, 400), rng)
    B = random_csr(300, 5000, rng.integers(0, 80, 300), rng)
    ones = lambda X: sp.csr_matrix((np.ones(X.nnz), X.col, X.rpt), shape=(X.M, X.N))
    P = ones(A) @ ones(B)
    assert np.array_equal(oracle.symbolic(A, B), P.indptr.astype(np.int64))
    n = 9
    S = G.stencil27(n, 2)
    rpt = oracle.symbolic(S, S)
    for i in range(S.M):
        x, y, z = i % n, (i // n) % n, i // (n * n)
        cnt = 1
        for c in (x, y, z):
            cnt *= min(c + 2, n - 1) - max(c - 2, 0) + 1
        assert rpt[i + 1] - rpt[i] == cnt


def test_rows_matches_full():
    A = G.rmat(10, 8, 3)
    full = oracle.spgemm(A, A)
    r0, r1 = 100, 700
    rpt, col, val, ab = oracle.spgemm_rows(A, A, r0, r1)
    b0, b1 = full[0][r0], full[0][r1]
    assert np.array_equal(rpt, full[0][r0:r1 + 1] - b0)
    assert np.array_equal(col, full[1][b0:b1]) and np.array_equal(val, full[2][b0:b1])


def test_backward_dense_and_finite_differences():
    rng = np.random.default_rng(8)
    A = random_csr(30, 25, rng.integers(0, 8, 30), rng)
    B = random_csr(25, 35, rng.integers(0, 8, 25), rng)
    rpt, col, val, _ = oracle.spgemm(A, B)
    dc = rng.normal(size=col.size)
    da, db = oracle.backward(A, B, rpt, col, dc)
    DA, DB = dense_of(A.rpt, A.col, A.val, 30, 25), dense_of(B.rpt, B.col, B.val, 25, 35)
    DC = dense_of(rpt, col, dc, 30, 35)
    gA = DC @ DB.T
    gB = DA.T @ DC
    for i in range(30):
        for p in range(A.rpt[i], A.rpt[i + 1]):
            assert abs(da[p] - gA[i, A.col[p]]) <= 1e-12 * (1 + abs(gA[i, A.col[p]]))
    for k in range(25):
        for p in range(B.rpt[k], B.rpt[k + 1]):
            assert abs(db[p] - gB[k, B.col[p]]) <= 1e-12 * (1 + abs(gB[k, B.col[p]]))

    def loss(Aval, Bval):
        AA = G.HostCSR(A.M, A.N, A.rpt, A.col, Aval)
        BB = G.HostCSR(B.M, B.N, B.rpt, B.col, Bval)
        return float(np.dot(oracle.spgemm(AA, BB, with_abs=False)[2], dc))

    h = 1e-3
    for p in rng.choice(A.col.size, 10, replace=False):
        e = np.zeros_like(A.val)
        e[p] = h
        fd = (loss(A.val + e, B.val) - loss(A.val - e, B.val)) / (2 * h)
        assert abs(fd - da[p]) <= 1e-9 * (1 + abs(da[p]))
    for p in rng.choice(B.col.size, 10, replace=False):
        e = np.zeros_like(B.val)
        e[p] = h
        fd = (loss(A.val, B.val + e) - loss(A.val, B.val - e)) / (2 * h)
        assert abs(fd - db[p]) <= 1e-9 * (1 + abs(db[p]))


def test_compare_rejects_planted_mistakes():
    A = G.rmat(9, 8, 2)
    ref = oracle.spgemm(A, A)
    rpt, col, val, _ = ref
    oracle.compare(rpt, col, val, ref, 1e-12)
    bad_col = col.copy()
    i = int(np.nonzero(np.diff(rpt) >= 2)[0][0])
    bad_col[rpt[i]], bad_col[rpt[i] + 1] = bad_col[rpt[i] + 1], bad_col[rpt[i]]
    with pytest.raises(AssertionError):
        oracle.compare(rpt, bad_col, val, ref, 1e-12)
    bad_rpt = rpt.copy()
    bad_rpt[5] += 1
    with pytest.raises(AssertionError):
        oracle.compare(bad_rpt, col, val, ref, 1e-12)
    bad_val = val.copy()
    j = int(np.argmax(ref[3]))
    bad_val[j] += 1e-10 * ref[3][j]
    with pytest.raises(AssertionError):
        oracle.compare(rpt, col, bad_val, ref, 1e-12)
    # a dropped term (one product missing from one entry) is caught
    k = int(np.argmax(ref[3]))
    with pytest.raises(AssertionError):
        oracle.compare(rpt, col, val - np.where(np.arange(val.size) == k, ref[3][k] * 1e-3, 0), ref, 1e-12)
