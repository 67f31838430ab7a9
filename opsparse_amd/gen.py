"""Synthetic CSR inputs (host numpy arrays) from the library's generators.

Every draw is a pure function of (seed, index), so the matrices are
bit-identical on every machine and thread count.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import OpsHostCsr, check, lib


@dataclass
class HostCSR:
    """Zero-based CSR on the host: int32 rpt/col, float64 val."""
    M: int
    N: int
    rpt: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.rpt[-1] - self.rpt[0]) if self.M > 0 else 0

    def astype(self, dtype) -> "HostCSR":
        return HostCSR(self.M, self.N, self.rpt, self.col, self.val.astype(dtype))

    def row_slice(self, r0: int, r1: int) -> "HostCSR":
        """Rows [r0, r1) with rpt rebased to 0."""
        base = int(self.rpt[r0])
        end = int(self.rpt[r1])
        return HostCSR(r1 - r0, self.N, (self.rpt[r0:r1 + 1] - base).astype(np.int32), self.col[base:end].copy(),
                       self.val[base:end].copy())


def _take(h: OpsHostCsr) -> HostCSR:
    try:
        M, N, nnz = int(h.M), int(h.N), int(h.nnz)
        rpt = np.ctypeslib.as_array(h.rpt, shape=(M + 1,)).copy()
        col = np.ctypeslib.as_array(h.col, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(h.val, shape=(max(nnz, 1),))[:nnz].copy()
    finally:
        lib().ops_host_csr_free(C.byref(h))
    return HostCSR(M, N, rpt, col, val)


def uniform_rows(M: int, N: int, d: int, seed: int) -> HostCSR:
    """M x N, exactly d distinct uniform columns per row, values U[-1,1]."""
    h = OpsHostCsr()
    check(lib().ops_gen_uniform_rows(M, N, d, seed, C.byref(h)))
    return _take(h)


def stencil27(n: int, seed: int, diag: float = 26.0, off: float = -1.0, noise: float = 0.01,
              ny: int | None = None, nz: int | None = None) -> HostCSR:
    """27-point stencil on an n x ny x nz grid (x fastest)."""
    h = OpsHostCsr()
    check(lib().ops_gen_stencil27(n, ny or n, nz or n, diag, off, noise, seed, C.byref(h)))
    return _take(h)


def rmat(scale: int, edge_factor: int, seed: int, abc=(0.57, 0.19, 0.19), normal: bool = False) -> HostCSR:
    """R-MAT 2^scale x 2^scale, duplicates merged, no vertex permutation."""
    h = OpsHostCsr()
    check(lib().ops_gen_rmat(scale, edge_factor, abc[0], abc[1], abc[2], seed, 1 if normal else 0, C.byref(h)))
    return _take(h)


@dataclass
class Batch:
    """Instances packed block-diagonally: instance i owns rows/columns
    [row_base[i], row_base[i+1]) of A and B."""
    A: HostCSR
    B: HostCSR
    row_base: np.ndarray
    d: np.ndarray

    @property
    def count(self) -> int:
        return len(self.d)


def batch(count: int, seed: int, n_lo: int = 64, n_hi: int = 16384, d_lo: int = 4, d_hi: int = 32) -> Batch:
    ha, hb = OpsHostCsr(), OpsHostCsr()
    rb = np.zeros(count + 1, dtype=np.int64)
    d = np.zeros(max(count, 1), dtype=np.int32)
    check(lib().ops_gen_batch(count, n_lo, n_hi, d_lo, d_hi, seed, C.byref(ha), C.byref(hb),
                              rb.ctypes.data_as(C.POINTER(C.c_int64)), d.ctypes.data_as(C.POINTER(C.c_int32))))
    return Batch(_take(ha), _take(hb), rb, d[:count])


def compute_flop_host(A: HostCSR, B: HostCSR) -> int:
    """Total products of A*B (library's host OpenMP count)."""
    return int(lib().ops_compute_flop_host(A.M, A.rpt.ctypes.data, A.col.ctypes.data, B.rpt.ctypes.data, None))
