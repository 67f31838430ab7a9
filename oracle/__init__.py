"""CPU oracle for C = A*B in CSR: test infrastructure only.

Loads oracle/_build/liboracle.so (built from oracle/gustavson.c by
oracle/Makefile).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, as the checker; the library never calls
it.  See gustavson.c for what each function restates from the reference and
tests/test_oracle.py for what pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_DIR, "_build", "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        h = C.CDLL(_SO)
        v, i64 = C.c_void_p, C.c_int64
        h.oracle_num_threads.restype = C.c_int
        h.oracle_compute_flop.restype = i64
        h.oracle_compute_flop.argtypes = [C.c_int, v, v, v, v]
        h.oracle_symbolic.restype = i64
        h.oracle_symbolic.argtypes = [C.c_int, C.c_int, v, v, v, v, v]
        h.oracle_numeric.restype = None
        h.oracle_numeric.argtypes = [C.c_int, C.c_int, v, v, v, v, v, v, C.c_int, v, v, v, v]
        h.oracle_rows.restype = i64
        h.oracle_rows.argtypes = [C.c_int, C.c_int, C.c_int, v, v, v, v, v, v, C.c_int, v, v, v, v, i64]
        h.oracle_backward.restype = None
        h.oracle_backward.argtypes = [C.c_int, v, v, v, v, v, v, C.c_int, v, v, v, v, v, C.c_int]
        _lib = h
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


def _vals(a):
    if a.dtype == np.float32:
        return np.ascontiguousarray(a), 1
    return np.ascontiguousarray(a, dtype=np.float64), 0


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def compute_flop(A, B, per_row: bool = False):
    """Total products (and per-row counts) of A*B; A, B have rpt/col/M."""
    rf = np.empty(A.M, dtype=np.int64) if per_row else None
    tot = lib().oracle_compute_flop(A.M, _p(A.rpt), _p(A.col), _p(B.rpt), _p(rf))
    return (int(tot), rf) if per_row else int(tot)


def symbolic(A, B):
    """Row offsets of A*B alone: rpt int64[M+1] (exclusive sum of each row's
    number of distinct columns), without computing columns or values."""
    rpt = np.zeros(A.M + 1, dtype=np.int64)
    lib().oracle_symbolic(A.M, B.N, _p(A.rpt), _p(A.col), _p(B.rpt), _p(B.col), _p(rpt))
    return rpt


def spgemm(A, B, with_abs: bool = True):
    """Returns (rpt int64[M+1], col int32, val float64, absum float64 or None)."""
    M, N = A.M, B.N
    rpt = np.zeros(M + 1, dtype=np.int64)
    nnz = lib().oracle_symbolic(M, N, _p(A.rpt), _p(A.col), _p(B.rpt), _p(B.col), _p(rpt))
    col = np.empty(max(nnz, 1), dtype=np.int32)
    val = np.empty(max(nnz, 1), dtype=np.float64)
    ab = np.empty(max(nnz, 1), dtype=np.float64) if with_abs else None
    av, f32 = _vals(A.val)
    bv, _ = _vals(B.val.astype(av.dtype, copy=False))
    lib().oracle_numeric(M, N, _p(A.rpt), _p(A.col), _p(av), _p(B.rpt), _p(B.col), _p(bv), f32, _p(rpt),
                         _p(col), _p(val), _p(ab))
    return rpt, col[:nnz], val[:nnz], (ab[:nnz] if with_abs else None)


def spgemm_rows(A, B, r0: int, r1: int, with_abs: bool = True):
    """Rows [r0, r1) of A*B: (rpt local int64, col, val, absum)."""
    v = lib()
    av, f32 = _vals(A.val)
    bv, _ = _vals(B.val.astype(av.dtype, copy=False))
    rpt = np.zeros(r1 - r0 + 1, dtype=np.int64)
    nnz = v.oracle_rows(r0, r1, B.N, _p(A.rpt), _p(A.col), _p(av), _p(B.rpt), _p(B.col), _p(bv), f32, _p(rpt),
                        None, None, None, 0)
    col = np.empty(max(nnz, 1), dtype=np.int32)
    val = np.empty(max(nnz, 1), dtype=np.float64)
    ab = np.empty(max(nnz, 1), dtype=np.float64) if with_abs else None
    got = v.oracle_rows(r0, r1, B.N, _p(A.rpt), _p(A.col), _p(av), _p(B.rpt), _p(B.col), _p(bv), f32, _p(rpt),
                        _p(col), _p(val), _p(ab), nnz)
    assert got == nnz
    return rpt, col[:nnz], val[:nnz], (ab[:nnz] if with_abs else None)


def backward(A, B, crpt, ccol, dc):
    """FP64 (dA on pattern(A), dB on pattern(B)) for dL/dC = dc."""
    av, f32 = _vals(A.val)
    bv, _ = _vals(B.val.astype(av.dtype, copy=False))
    dcv = np.ascontiguousarray(dc, dtype=av.dtype)
    crpt = np.ascontiguousarray(crpt, dtype=np.int64)
    ccol = np.ascontiguousarray(ccol, dtype=np.int32)
    nA, nB = int(A.rpt[-1]), int(B.rpt[-1])
    da = np.empty(max(nA, 1), dtype=np.float64)
    db = np.empty(max(nB, 1), dtype=np.float64)
    lib().oracle_backward(A.M, _p(A.rpt), _p(A.col), _p(av), _p(B.rpt), _p(B.col), _p(bv), f32, _p(crpt),
                          _p(ccol), _p(dcv), _p(da), _p(db), nB)
    return da[:nA], db[:nB]


def compare(rpt, col, val, ref, tol: float):
    """Structure must be bit-exact; value error is max |d| / sum|a*b|.

    ref = (rpt, col, val, absum) from spgemm().  Returns a dict with the
    error figures; raises AssertionError on a structure mismatch or when
    the normalised error exceeds tol."""
    rr, rc, rv, ra = ref
    rpt = np.asarray(rpt, dtype=np.int64)
    if rpt.shape != rr.shape or not np.array_equal(rpt - rpt[0], rr):
        bad = np.nonzero((rpt - rpt[0]) != rr)[0] if rpt.shape == rr.shape else []
        raise AssertionError(f"row_ptr mismatch (first rows {list(bad[:5])})")
    if not np.array_equal(np.asarray(col), rc):
        bad = np.nonzero(np.asarray(col) != rc)[0]
        raise AssertionError(f"column mismatch at {len(bad)} entries (first {list(bad[:5])})")
    val = np.asarray(val, dtype=np.float64)
    d = np.abs(val - rv)
    scale = np.maximum(ra, np.finfo(np.float64).tiny) if ra is not None else np.maximum(np.abs(rv), 1e-300)
    err = float((d / scale).max()) if d.size else 0.0
    nz = np.abs(rv) >= 1e-300
    rel = float((d[nz] / np.abs(rv[nz])).max()) if nz.any() else 0.0
    out = dict(max_abs=float(d.max()) if d.size else 0.0, max_rel=rel, max_norm=err, nnz=int(rc.size))
    if err > tol:
        raise AssertionError(f"value error {err:.3e} above tolerance {tol:.1e} ({out})")
    return out
