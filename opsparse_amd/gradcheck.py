"""Gradient checks of the masked SpGEMM backward through the GPU forward.

L = <dC, A*B> is linear in every single entry of A and of B, so a central
difference of L in one entry is exact up to rounding for any step.  These
helpers sample entries of A and B, compute that difference with the
library's forward call (only the rows of C an entry feeds are recomputed)
and compare it, and a direct FP64 sum of the gradient's terms, with the
backward's values.  Errors are normalised by the sum of |terms| of the
entry's gradient, which bounds the rounding of any summation order.

Used by the GPU tests and by bench.py's forward+backward workload; it
needs a GPU and the built library, and calls nothing but the library's
forward.
"""
from __future__ import annotations

import numpy as np
import torch

from .gen import HostCSR
from .spgemm import DeviceCSR, spgemm


def log(msg):
    import sys
    print(f"[gradcheck] {msg}", file=sys.stderr, flush=True)


def _gather_rows(rpt, rows):
    """Indices of the entries of `rows` (in that order) of a CSR with row offsets rpt."""
    starts, ends = rpt[rows], rpt[rows + 1]
    lens = ends - starts
    tot = int(lens.sum())
    first = np.cumsum(lens) - lens
    return np.repeat(starts - first, lens) + np.arange(tot, dtype=np.int64)


def gradient_checks(A, B, dA64, dB64, crpt_h, ccol_h, dc_h, ga32, gb32, ga64, gb64, n_entries, ws, seed=4):
    """FP64 checks of sampled gradient entries: direct sums on the host
    (dA[i,k] = sum_j dC[i,j] B[k,j], dB[k,j] = sum_i A[i,k] dC[i,j]) and
    central differences of L = <dC, A*B> through the GPU forward, restricted
    to the rows of C an entry feeds.  Errors are normalised by the sum of
    |terms| of the entry's gradient."""
    rng = np.random.default_rng(seed)
    a_rows = np.repeat(np.arange(A.M, dtype=np.int64), np.diff(A.rpt.astype(np.int64)))
    b_rows = np.repeat(np.arange(B.M, dtype=np.int64), np.diff(B.rpt.astype(np.int64)))
    order = np.argsort(A.col, kind="stable")  # A by column: rows feeding each row of B
    cstart = np.zeros(A.N + 1, dtype=np.int64)
    np.cumsum(np.bincount(A.col, minlength=A.N), out=cstart[1:])
    ga32_h, gb32_h = ga32.double().cpu().numpy(), gb32.double().cpu().numpy()
    ga64_h, gb64_h = ga64.cpu().numpy(), gb64.cpu().numpy()
    # L is linear in each entry, so the central difference is exact up to
    # rounding for any step; a large step keeps the rounding of the other
    # terms of the affected rows (which do not scale with it) negligible.
    h = 1024.0

    def sub_csr(rows, k=None, j_local=None, delta=0.0):
        idx = _gather_rows(A.rpt.astype(np.int64), rows)
        rpt = np.zeros(len(rows) + 1, dtype=np.int32)
        np.cumsum(np.diff(A.rpt.astype(np.int64))[rows], out=rpt[1:])
        val = A.val[idx].astype(np.float64)
        if j_local is not None:
            val[j_local] += delta
        return HostCSR(len(rows), A.N, rpt, A.col[idx].astype(np.int32), val)

    def c_rows(Ahost, Bdev):
        dAh = DeviceCSR.from_host(Ahost, torch.float64)
        c = spgemm(dAh, Bdev, ws=ws, wide=True)
        rpt, col, val = c.to_host()
        c.free()
        return rpt, col, val

    out = {"fd64": 0.0, "ref64": 0.0, "ref32": 0.0, "zero_entries": 0, "zero_max_abs_fd": 0.0,
           "zero_max_abs_grad": 0.0}

    def record(fd, ref, scale, g64, g32):
        if scale == 0.0:  # no terms: the gradient is structurally zero
            out["zero_entries"] += 1
            out["zero_max_abs_fd"] = max(out["zero_max_abs_fd"], abs(fd))
            out["zero_max_abs_grad"] = max(out["zero_max_abs_grad"], abs(g64), abs(g32))
            return
        out["fd64"] = max(out["fd64"], abs(fd - g64) / scale)
        out["ref64"] = max(out["ref64"], abs(ref - g64) / scale)
        out["ref32"] = max(out["ref32"], abs(ref - g32) / scale)
    n_half = n_entries // 2
    for cnt, e in enumerate(rng.choice(A.nnz, n_half, replace=False)):
        if cnt % 128 == 0:
            log(f"gradient check: dA entry {cnt}/{n_half}")
        i, k = int(a_rows[e]), int(A.col[e])
        crow = ccol_h[crpt_h[i]:crpt_h[i + 1]]
        bj = B.col[B.rpt[k]:B.rpt[k + 1]]
        bv = B.val[B.rpt[k]:B.rpt[k + 1]].astype(np.float64)
        pos = np.searchsorted(crow, bj)
        assert np.array_equal(crow[pos], bj), "pattern(C) misses a product of a row of A"
        dcr = dc_h[crpt_h[i] + pos]
        ref, scale = float(np.dot(dcr, bv)), float(np.dot(np.abs(dcr), np.abs(bv)))
        # FD: row i with A[i,k] +- h, both in one call
        local = e - int(A.rpt[i])
        Hp = sub_csr(np.array([i]), j_local=local, delta=h)
        Hm = sub_csr(np.array([i]), j_local=local, delta=-h)
        step = float(Hp.val[local]) - float(Hm.val[local])
        two = HostCSR(2, A.N, np.concatenate([Hp.rpt, Hp.rpt[1:] + Hp.rpt[-1]]).astype(np.int32),
                      np.concatenate([Hp.col, Hm.col]), np.concatenate([Hp.val, Hm.val]))
        rpt, col, val = c_rows(two, dB64)
        n = int(rpt[1])
        assert np.array_equal(col[:n], crow) and np.array_equal(col[n:], crow)
        fd = float(np.dot(dc_h[crpt_h[i]:crpt_h[i + 1]], val[:n] - val[n:])) / step
        record(fd, ref, scale, ga64_h[e], ga32_h[e])
    for cnt, e in enumerate(rng.choice(B.nnz, n_entries - n_half, replace=False)):
        if cnt % 128 == 0:
            log(f"gradient check: dB entry {cnt}/{n_entries - n_half}")
        k, j = int(b_rows[e]), int(B.col[e])
        feed = order[cstart[k]:cstart[k + 1]]
        rows = a_rows[feed]
        av = A.val[feed].astype(np.float64)
        dcv = np.empty(len(rows))
        for t, r in enumerate(rows):
            crow = ccol_h[crpt_h[r]:crpt_h[r + 1]]
            p = int(np.searchsorted(crow, j))
            assert p < crow.size and crow[p] == j
            dcv[t] = dc_h[crpt_h[r] + p]
        ref, scale = float(np.dot(av, dcv)), float(np.dot(np.abs(av), np.abs(dcv)))
        fd = 0.0
        if len(rows):
            sub = sub_csr(rows)
            res, pert = [], []
            for s in (h, -h):
                bval = dB64.val.clone()
                bval[int(e)] += s
                pert.append(float(bval[int(e)]))
                res.append(c_rows(sub, dB64.with_values(bval)))
            (rp, cp, vp), (_, cm, vm) = res
            idx = _gather_rows(crpt_h, rows)
            assert np.array_equal(cp, ccol_h[idx]) and np.array_equal(cm, cp)
            fd = float(np.dot(dc_h[idx], vp - vm)) / (pert[0] - pert[1])
        record(fd, ref, scale, gb64_h[e], gb32_h[e])
    return out
