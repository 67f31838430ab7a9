"""Shared helpers for the GPU parity tests (host-side, numpy)."""
from __future__ import annotations

import numpy as np

import oracle
from opsparse_amd.gen import HostCSR

# Value tolerances from the specification, on max |d| / sum|a*b|.
TOL = {np.float64: 1e-12, np.float32: 1e-5}


def csr_from_dense(D: np.ndarray) -> HostCSR:
    M, N = D.shape
    rpt = np.zeros(M + 1, dtype=np.int32)
    cols, vals = [], []
    for i in range(M):
        nz = np.nonzero(D[i])[0]
        rpt[i + 1] = rpt[i] + nz.size
        cols.append(nz.astype(np.int32))
        vals.append(D[i, nz])
    col = np.concatenate(cols) if cols else np.zeros(0, np.int32)
    val = np.concatenate(vals) if vals else np.zeros(0)
    return HostCSR(M, N, rpt, col.astype(np.int32), val.astype(np.float64))


def csr_from_rows(M: int, N: int, rows, rng) -> HostCSR:
    """rows[i] = sorted unique column array of row i; values U[-1,1]."""
    rpt = np.zeros(M + 1, dtype=np.int32)
    for i, r in enumerate(rows):
        rpt[i + 1] = rpt[i] + len(r)
    col = np.concatenate([np.asarray(r, dtype=np.int32) for r in rows]) if M else np.zeros(0, np.int32)
    val = rng.uniform(-1, 1, size=col.size)
    return HostCSR(M, N, rpt, col.astype(np.int32), val)


def random_csr(M: int, N: int, lens, rng) -> HostCSR:
    """Row i gets lens[i] distinct uniform columns."""
    rows = []
    for l in lens:
        l = int(min(l, N))
        c = np.unique(rng.integers(0, N, size=2 * l + 8))
        while c.size < l:
            c = np.unique(np.r_[c, rng.integers(0, N, size=2 * l + 8)])
        rows.append(np.sort(rng.choice(c, size=l, replace=False)))
    return csr_from_rows(M, N, rows, rng)


def check_product(rpt, col, val, A: HostCSR, B: HostCSR, dtype=np.float64):
    ref = oracle.spgemm(A.astype(dtype) if dtype == np.float32 else A, B.astype(dtype) if dtype == np.float32 else B)
    return oracle.compare(rpt, col, val, ref, TOL[dtype])
