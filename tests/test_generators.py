"""Invariants of the deterministic synthetic generators (CPU)."""
import numpy as np

import opsparse_amd.gen as G


def rows_sorted_unique(A):
    for i in range(A.M):
        c = A.col[A.rpt[i]:A.rpt[i + 1]]
        if c.size and (np.any(np.diff(c) <= 0) or c[0] < 0 or c[-1] >= A.N):
            return False
    return True


def test_uniform_rows():
    A = G.uniform_rows(2000, 3000, 8, 0)
    assert A.M == 2000 and A.nnz == 16000 and np.all(np.diff(A.rpt) == 8)
    assert rows_sorted_unique(A)
    assert np.all(np.abs(A.val) <= 1.0) and A.val.min() < -0.9 and A.val.max() > 0.9
    B = G.uniform_rows(2000, 3000, 8, 0)
    assert np.array_equal(A.col, B.col) and np.array_equal(A.val, B.val)
    C = G.uniform_rows(2000, 3000, 8, 1)
    assert not np.array_equal(A.col, C.col)
    # columns cover the range roughly uniformly
    h = np.bincount(A.col, minlength=3000)
    assert h.max() < 30
    # dense rows (d == N) are the full column range
    D = G.uniform_rows(5, 40, 40, 3)
    assert all(np.array_equal(D.col[D.rpt[i]:D.rpt[i + 1]], np.arange(40)) for i in range(5))


def test_stencil():
    n = 9
    A = G.stencil27(n, 1)
    assert rows_sorted_unique(A)
    for i in range(A.M):
        x, y, z = i % n, (i // n) % n, i // (n * n)
        cnt = 1
        for c in (x, y, z):
            cnt *= min(c + 1, n - 1) - max(c - 1, 0) + 1
        assert A.rpt[i + 1] - A.rpt[i] == cnt
        cols = A.col[A.rpt[i]:A.rpt[i + 1]]
        vals = A.val[A.rpt[i]:A.rpt[i + 1]]
        assert vals[cols == i][0] == 26.0
        off = vals[cols != i]
        assert np.all(np.abs(off + 1.0) <= 0.01 + 1e-15)
    Z = G.stencil27(n, 1, noise=0.0)
    assert np.all((Z.val == 26.0) | (Z.val == -1.0))


def test_rmat():
    A = G.rmat(12, 16, 2)
    assert A.M == A.N == 4096
    assert rows_sorted_unique(A)
    assert 0.5 * 16 * 4096 < A.nnz <= 16 * 4096
    deg = np.diff(A.rpt)
    assert deg[0] == deg.max()  # no permutation: vertex 0 is the hub for a > d
    B = G.rmat(12, 16, 2)
    assert np.array_equal(A.col, B.col) and np.array_equal(A.val, B.val)
    Nn = G.rmat(12, 16, 2, normal=True)
    assert np.array_equal(Nn.col, A.col)
    assert abs(Nn.val.mean()) < 0.05 and abs(Nn.val.std() - 1.0) < 0.05


def test_batch():
    b = G.batch(64, 3)
    assert b.count == 64 and b.row_base[0] == 0 and b.row_base[-1] == b.A.M == b.B.M
    n = np.diff(b.row_base)
    assert n.min() >= 64 and n.max() <= 16384
    assert b.d.min() >= 4 and b.d.max() <= 32
    for X in (b.A, b.B):
        assert rows_sorted_unique(X)
        for i in range(b.count):
            r0, r1 = b.row_base[i], b.row_base[i + 1]
            seg = X.col[X.rpt[r0]:X.rpt[r1]]
            assert seg.min() >= r0 and seg.max() < r1
            assert np.all(np.diff(X.rpt[r0:r1 + 1]) == b.d[i])
    assert not np.array_equal(b.A.col, b.B.col)
