"""PyTorch-ROCm host interface over the C ABI (include/opsparse_amd.h).

Device memory and streams come from torch; every multiply runs the
library's HIP kernels through ctypes.  There is no CPU fallback: without
the built library or a GPU these functions raise.

Mirrors the reference's operator surface (include/opsparse/core_algo/spgemm/spgemm.hpp:12-22):
``spgemm`` is opsparse_spgemm (the library allocates C), ``spgemm_safe``
is opsparse_spgemm_safe (status instead of exceptions), plus the
structure-reuse ``Plan``, the batched call and the differentiable
``SpGEMMFunction`` that have no reference counterpart.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import (OPS_F32, OPS_F64, OPS_INVALID, OPS_OK, OpsCsr, OpsCsr64, OpsError, OpsMemStats, OpsTimings,
                   OpsTuning, check, lib)
from .gen import HostCSR

_HIP = None


def _hip():
    """The HIP runtime already loaded by torch (same SONAME as the library's)."""
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so.7")
        _HIP.hipMemcpy.restype = C.c_int
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _HIP


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("opsparse_amd needs a HIP GPU; none is visible")


def _stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float64:
        return OPS_F64
    if dtype == torch.float32:
        return OPS_F32
    raise TypeError(f"unsupported value dtype {dtype}")


@dataclass
class DeviceCSR:
    """CSR on the GPU as torch tensors: rpt int32 (M+1), col int32, val f32/f64.

    rpt need not start at 0 (a row slice shares col/val with its parent);
    nnz is then rpt[M] - rpt[0], given as `entries` (row_slice sets it), and
    col.numel() otherwise."""
    M: int
    N: int
    rpt: torch.Tensor
    col: torch.Tensor
    val: torch.Tensor
    entries: int | None = None

    @property
    def nnz(self) -> int:
        return int(self.col.numel()) if self.entries is None else int(self.entries)

    def row_slice(self, r0: int, r1: int) -> "DeviceCSR":
        """Rows [r0, r1) as a descriptor over the same col/val storage."""
        if not 0 <= r0 <= r1 <= self.M:
            raise ValueError(f"row range [{r0}, {r1}) outside [0, {self.M}]")
        ends = self.rpt[[r0, r1]].cpu()
        return DeviceCSR(r1 - r0, self.N, self.rpt[r0:r1 + 1], self.col, self.val,
                         int(ends[1]) - int(ends[0]))

    @staticmethod
    def from_host(h: HostCSR, dtype=torch.float64, device="cuda") -> "DeviceCSR":
        return DeviceCSR(h.M, h.N, torch.from_numpy(np.ascontiguousarray(h.rpt)).to(device),
                         torch.from_numpy(np.ascontiguousarray(h.col)).to(device),
                         torch.from_numpy(np.ascontiguousarray(h.val)).to(device=device, dtype=dtype))

    def desc(self) -> OpsCsr:
        d = OpsCsr()
        d.M, d.N, d.nnz = self.M, self.N, self.nnz
        d.rpt, d.col, d.val = self.rpt.data_ptr(), self.col.data_ptr(), self.val.data_ptr()
        return d

    def with_values(self, val: torch.Tensor) -> "DeviceCSR":
        return DeviceCSR(self.M, self.N, self.rpt, self.col, val, self.entries)


class Workspace:
    """Persistent per-device workspace (stream pool, events, buffers)."""

    def __init__(self):
        _require_gpu()
        self.ptr = C.c_void_p()
        check(lib().ops_workspace_create(C.byref(self.ptr)))

    def set_tuning(self, t) -> None:
        """Installs bin bounds / hash family (an OpsTuning or its dict form)."""
        if isinstance(t, dict):
            t = OpsTuning.from_dict(t)
        check(lib().ops_workspace_set_tuning(self.ptr, C.byref(t)), self.ptr)

    def get_tuning(self) -> dict:
        t = OpsTuning()
        check(lib().ops_workspace_get_tuning(self.ptr, C.byref(t)))
        return t.as_dict()

    def trim(self) -> None:
        """Frees the workspace's device buffers (ops_workspace_trim)."""
        check(lib().ops_workspace_trim(self.ptr), self.ptr)

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and self.ptr.value:
                lib().ops_workspace_destroy(self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass


def mem_cache_stats() -> dict:
    """Output-memory cache counters (ops_mem_cache_stats)."""
    st = OpsMemStats()
    check(lib().ops_mem_cache_stats(C.byref(st)))
    return st.as_dict()


def set_mem_cache_limit(nbytes: int) -> None:
    """Caps the bytes the output cache keeps per device (< 0: default, 0: off)."""
    check(lib().ops_mem_cache_set_limit(int(nbytes)))


def mem_cache_release() -> None:
    """Frees every cached output block."""
    check(lib().ops_mem_cache_release())


def set_kept_bitmap_limit(nbytes: int) -> None:
    """Caps the long-row column bitmaps kept per device (< 0: default, 0: none)."""
    check(lib().ops_kept_bitmap_limit(int(nbytes)))


def default_tuning() -> dict:
    t = OpsTuning()
    check(lib().ops_tuning_default(C.byref(t)))
    return t.as_dict()


def tuning_caps() -> dict:
    t = OpsTuning()
    check(lib().ops_tuning_caps(C.byref(t)))
    return t.as_dict()


_default_ws = None


def default_workspace() -> Workspace:
    global _default_ws
    if _default_ws is None:
        _default_ws = Workspace()
    return _default_ws


class LibCSR:
    """C allocated by the library (opsparse_spgemm semantics); freed with
    ops_csr_free / ops_csr64_free when this object dies."""

    def __init__(self, desc, wide: bool, dtype: torch.dtype, owned: bool = True):
        self.d = desc
        self.wide = wide
        self.dtype = dtype
        self.owned = owned  # False: a view of arrays another object owns (never freed here)

    @property
    def M(self):
        return int(self.d.M)

    @property
    def N(self):
        return int(self.d.N)

    @property
    def nnz(self):
        return int(self.d.nnz)

    def to_host(self):
        """(rpt int64, col int32, val) numpy copies."""
        hip = _hip()
        M, nnz = self.M, self.nnz
        ridx = np.int64 if self.wide else np.int32
        rpt = np.empty(M + 1, dtype=ridx)
        col = np.empty(nnz, dtype=np.int32)
        val = np.empty(nnz, dtype=np.float64 if self.dtype == torch.float64 else np.float32)
        torch.cuda.synchronize()
        for dst, src, n in ((rpt, self.d.rpt, rpt.nbytes), (col, self.d.col, col.nbytes), (val, self.d.val, val.nbytes)):
            if n:
                st = hip.hipMemcpy(dst.ctypes.data, src, n, 2)
                if st != 0:
                    raise RuntimeError(f"hipMemcpy failed: {st}")
        return rpt.astype(np.int64), col, val

    def rowptr_to_host(self):
        """All M+1 row offsets as an int64 numpy array (columns and values stay on the device)."""
        ridx = np.int64 if self.wide else np.int32
        rpt = np.empty(self.M + 1, dtype=ridx)
        torch.cuda.synchronize()
        if _hip().hipMemcpy(rpt.ctypes.data, self.d.rpt, rpt.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy failed")
        return rpt.astype(np.int64)

    def rows_to_host(self, r0: int, r1: int):
        """Rows [r0, r1) as (rpt rebased to 0, int64; col; val) numpy copies,
        reading only that slice of the device arrays."""
        if not 0 <= r0 <= r1 <= self.M:
            raise ValueError(f"row range [{r0}, {r1}) outside [0, {self.M}]")
        hip = _hip()
        ridx = np.int64 if self.wide else np.int32
        rpt = np.empty(r1 - r0 + 1, dtype=ridx)
        torch.cuda.synchronize()

        def copy(dst, src, nbytes):
            if nbytes and hip.hipMemcpy(dst.ctypes.data, src, nbytes, 2) != 0:
                raise RuntimeError("hipMemcpy failed")

        copy(rpt, self.d.rpt + r0 * rpt.itemsize, rpt.nbytes)
        rpt = rpt.astype(np.int64)
        s0, s1 = int(rpt[0]), int(rpt[-1])
        col = np.empty(s1 - s0, dtype=np.int32)
        val = np.empty(s1 - s0, dtype=np.float64 if self.dtype == torch.float64 else np.float32)
        copy(col, self.d.col + s0 * 4, col.nbytes)
        copy(val, self.d.val + s0 * val.itemsize, val.nbytes)
        return rpt - s0, col, val

    def desc(self) -> OpsCsr:
        """Input descriptor of this (int32-offset) result, e.g. a packed batch."""
        if self.wide:
            raise TypeError("a 64-bit-offset result cannot be an input descriptor")
        return self.d

    @property
    def val(self):
        # dtype carrier for spgemm(); the data stay in library memory
        return torch.empty(0, dtype=self.dtype)

    def free(self):
        if self.d is not None:
            if self.owned:
                (lib().ops_csr64_free if self.wide else lib().ops_csr_free)(C.byref(self.d))
            self.d = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def spgemm(A: DeviceCSR, B: DeviceCSR, ws: Workspace | None = None, timings: OpsTimings | None = None,
           wide: bool = False, stream=None) -> LibCSR:
    """C = A*B through ops_spgemm{,64}_{f64,f32}: the library allocates C."""
    _require_gpu()
    if A.val.dtype != B.val.dtype:
        raise TypeError("A and B must have the same value dtype")
    ws = ws or default_workspace()
    a, b = A.desc(), B.desc()
    f64 = A.val.dtype == torch.float64
    t = C.byref(timings) if timings is not None else None
    if wide:
        c = OpsCsr64()
        fn = lib().ops_spgemm64_f64 if f64 else lib().ops_spgemm64_f32
    else:
        c = OpsCsr()
        fn = lib().ops_spgemm_f64 if f64 else lib().ops_spgemm_f32
    check(fn(ws.ptr, C.byref(a), C.byref(b), C.byref(c), _stream_ptr(stream), t), ws.ptr)
    return LibCSR(c, wide, A.val.dtype)


def compute_flop_device(A: DeviceCSR, B: DeviceCSR, ws: Workspace | None = None, stream=None):
    """Products per row of A*B on the device (ops_compute_flop_device):
    (row_flop int32 tensor of A.M entries, saturated at INT32_MAX; 64-bit total)."""
    _require_gpu()
    ws = ws or default_workspace()
    rf = torch.empty(max(A.M, 1), dtype=torch.int32, device="cuda")
    tot = C.c_int64()
    a, b = A.desc(), B.desc()
    check(lib().ops_compute_flop_device(ws.ptr, C.byref(a), C.byref(b), rf.data_ptr(), C.byref(tot),
                                        _stream_ptr(stream)), ws.ptr)
    return rf[:A.M], int(tot.value)


def spgemm_safe(A: DeviceCSR, B: DeviceCSR, A_host: HostCSR, B_host: HostCSR, ws: Workspace | None = None):
    """opsparse_spgemm_safe: returns (status, LibCSR or None, info dict)."""
    ws = ws or default_workspace()
    a, b, c = A.desc(), B.desc(), OpsCsr()
    flop, cnnz = C.c_int64(), C.c_int64()
    fb, fa, tot = C.c_size_t(), C.c_size_t(), C.c_size_t()
    st = lib().ops_spgemm_safe_f64(ws.ptr, C.byref(a), C.byref(b), A_host.rpt.ctypes.data, A_host.col.ctypes.data,
                                   B_host.rpt.ctypes.data, C.byref(c), _stream_ptr(None), None, C.byref(flop),
                                   C.byref(cnnz), C.byref(fb), C.byref(fa), C.byref(tot))
    info = dict(total_flop=flop.value, c_nnz=cnnz.value, free_before=fb.value, free_after=fa.value,
                total_bytes=tot.value)
    return st, (LibCSR(c, False, torch.float64) if st == OPS_OK else None), info


def batch_pack(instances, dtype=torch.float64, stream=None):
    """Packs device CSR instances block-diagonally (ops_batch_pack).

    Returns (packed LibCSR, row_base, col_base): instance i owns rows
    [row_base[i], row_base[i+1]) and columns [col_base[i], col_base[i+1])."""
    n = len(instances)
    arr = (OpsCsr * max(n, 1))()
    for i, X in enumerate(instances):
        arr[i] = X.desc()
    packed = OpsCsr()
    rb = np.zeros(n + 1, dtype=np.int64)
    cb = np.zeros(n + 1, dtype=np.int64)
    check(lib().ops_batch_pack(n, arr, dtype_code(dtype), C.byref(packed), rb.ctypes.data_as(C.POINTER(C.c_int64)),
                               cb.ctypes.data_as(C.POINTER(C.c_int64)), _stream_ptr(stream)))
    return LibCSR(packed, False, dtype), rb, cb


class Executor:
    """The whole call C = A*B replayed from a hipGraph (ops_exec_*): setup,
    binning, symbolic, scan and numeric with bin sizes and nnz(C) on the
    device, no host synchronisation per launch.  The operands' descriptors
    are fixed; their patterns and values may change between launches (same
    entry counts).  ``Executor.batch`` packs independent instances A_i, B_i
    at every launch (A_i block-diagonally, B_i stacked) so that C_i, rows
    ``row_base[i]:row_base[i+1]`` of the result, is in B_i's own columns."""

    def __init__(self, A: DeviceCSR | None = None, B: DeviceCSR | None = None, dtype=torch.float64,
                 ws: Workspace | None = None, capacity: int = 0, graph: bool = True, _batch=None):
        _require_gpu()
        self.ws = ws or default_workspace()
        self.dtype = dtype
        self.ptr = C.c_void_p()
        self.row_base = None
        if _batch is None:
            self._keep = (A, B)  # the graph holds their device addresses
            a, b = A.desc(), B.desc()
            check(lib().ops_exec_create(self.ws.ptr, dtype_code(dtype), C.byref(a), C.byref(b), int(capacity),
                                        int(graph), C.byref(self.ptr)), self.ws.ptr)
        else:
            As, Bs = _batch
            self._keep = (list(As), list(Bs))
            n = len(As)
            arr_a, arr_b = (OpsCsr * max(n, 1))(), (OpsCsr * max(n, 1))()
            for i in range(n):
                arr_a[i], arr_b[i] = As[i].desc(), Bs[i].desc()
            rb = np.zeros(n + 1, dtype=np.int64)
            check(lib().ops_exec_create_batch(self.ws.ptr, dtype_code(dtype), n, arr_a, arr_b, int(capacity),
                                              int(graph), rb.ctypes.data_as(C.POINTER(C.c_int64)),
                                              C.byref(self.ptr)), self.ws.ptr)
            self.row_base = rb

    @classmethod
    def batch(cls, As, Bs, dtype=torch.float64, ws: Workspace | None = None, capacity: int = 0, graph: bool = True):
        return cls(dtype=dtype, ws=ws, capacity=capacity, graph=graph, _batch=(As, Bs))

    def launch(self, stream=None) -> None:
        check(lib().ops_exec_launch(self.ptr, _stream_ptr(stream)), self.ws.ptr)

    def result(self, stream=None) -> tuple[int, LibCSR]:
        """(status, C view) of the last launch (waits for it).  status is
        OPS_OK, or OPS_INVALID when nnz(C) exceeded the capacity (C.nnz is
        then the size needed and no entry was written)."""
        c, nnz = OpsCsr64(), C.c_int64()
        st = lib().ops_exec_result(self.ptr, C.byref(c), C.byref(nnz), _stream_ptr(stream))
        if st not in (OPS_OK, OPS_INVALID):
            check(st, self.ws.ptr)
        return st, LibCSR(c, True, self.dtype, owned=False)

    def destroy(self) -> None:
        if self.ptr:
            lib().ops_exec_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Plan:
    """Structure of C = A*B for fixed patterns: C's row offsets and numeric
    bins, computed once; the numeric phase can then run many times (also
    as a replayed hipGraph)."""

    def __init__(self, A: DeviceCSR, B: DeviceCSR, ws: Workspace | None = None, stream=None):
        _require_gpu()
        self.ws = ws or default_workspace()
        self.M, self.N = A.M, B.N
        self.ptr = C.c_void_p()
        nnz = C.c_int64()
        a, b = A.desc(), B.desc()
        check(lib().ops_plan_create(self.ws.ptr, C.byref(a), C.byref(b), C.byref(self.ptr), C.byref(nnz),
                                    _stream_ptr(stream)), self.ws.ptr)
        self.nnz = int(nnz.value)
        self._rpt64 = None

    @property
    def kept_rows(self) -> int:
        """Long rows whose symbolic column bitmap the plan keeps."""
        n = C.c_int64()
        check(lib().ops_plan_kept_rows(self.ptr, C.byref(n)))
        return int(n.value)

    def rowptr64(self, stream=None) -> torch.Tensor:
        if self._rpt64 is None:
            t = torch.empty(self.M + 1, dtype=torch.int64, device="cuda")
            check(lib().ops_plan_rowptr64(self.ptr, t.data_ptr(), _stream_ptr(stream)), self.ws.ptr)
            self._rpt64 = t
        return self._rpt64

    def rowptr32(self, stream=None) -> torch.Tensor:
        t = torch.empty(self.M + 1, dtype=torch.int32, device="cuda")
        check(lib().ops_plan_rowptr32(self.ptr, t.data_ptr(), _stream_ptr(stream)), self.ws.ptr)
        return t

    def numeric(self, A: DeviceCSR, B: DeviceCSR, col: torch.Tensor | None = None, val: torch.Tensor | None = None,
                graph: bool = False, stream=None):
        dt = A.val.dtype
        if col is None:
            col = torch.empty(max(self.nnz, 1), dtype=torch.int32, device="cuda")
        if val is None:
            val = torch.empty(max(self.nnz, 1), dtype=dt, device="cuda")
        a, b = A.desc(), B.desc()
        fn = lib().ops_plan_numeric_graph if graph else lib().ops_plan_numeric
        check(fn(self.ptr, dtype_code(dt), C.byref(a), C.byref(b), col.data_ptr(), val.data_ptr(),
                 _stream_ptr(stream)), self.ws.ptr)
        return col[:self.nnz], val[:self.nnz]

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib().ops_plan_destroy(self.ptr)
            self.ptr = C.c_void_p()


def backward(A: DeviceCSR, B: DeviceCSR, crpt64: torch.Tensor, ccol: torch.Tensor, dc: torch.Tensor,
             ws: Workspace | None = None, stream=None):
    """(dA values on pattern(A), dB values on pattern(B)) for L with dL/dC = dc.

    Each has one value per entry of its descriptor, from rpt[0] on (a row
    slice's gradients are the slice's nnz values)."""
    ws = ws or default_workspace()
    dt = A.val.dtype
    da = torch.empty(max(A.nnz, 1), dtype=dt, device="cuda")
    db = torch.empty(max(B.nnz, 1), dtype=dt, device="cuda")
    a, b = A.desc(), B.desc()
    dc = dc.contiguous()
    check(lib().ops_spgemm_backward(ws.ptr, dtype_code(dt), C.byref(a), C.byref(b), crpt64.data_ptr(),
                                    ccol.data_ptr(), dc.data_ptr(), da.data_ptr(), db.data_ptr(),
                                    _stream_ptr(stream)), ws.ptr)
    return da[:A.nnz], db[:B.nnz]


class SpGEMMFunction(torch.autograd.Function):
    """C_values = spgemm(A, B) on fixed patterns, differentiable in the
    values of A and B.  Backward is the masked pair
    dA = (dC B^T) on pattern(A), dB = (A^T dC) on pattern(B), computed by
    the library's HIP backward kernels.  The backward reads C's structure
    from the plan (row offsets) and the forward's column list; no hash
    tables are stored between the passes.  The plan itself holds device
    memory until it dies, including kept long-row column bitmaps within
    the process-wide limit (set_kept_bitmap_limit)."""

    @staticmethod
    def forward(ctx, a_val, b_val, A: DeviceCSR, B: DeviceCSR, plan: Plan):
        Aw, Bw = A.with_values(a_val.contiguous()), B.with_values(b_val.contiguous())
        col, val = plan.numeric(Aw, Bw)
        ctx.save_for_backward(a_val, b_val)
        ctx.A, ctx.B, ctx.plan, ctx.col = A, B, plan, col
        ctx.mark_non_differentiable(col)
        return val, col

    @staticmethod
    def backward(ctx, dval, dcol):
        a_val, b_val = ctx.saved_tensors
        Aw, Bw = ctx.A.with_values(a_val.contiguous()), ctx.B.with_values(b_val.contiguous())
        da, db = backward(Aw, Bw, ctx.plan.rowptr64(), ctx.col, dval.contiguous())
        return da, db, None, None, None


def spgemm_autograd(a_val: torch.Tensor, b_val: torch.Tensor, A: DeviceCSR, B: DeviceCSR, plan: Plan | None = None):
    """Differentiable C = A*B; returns (plan, C values, C columns)."""
    plan = plan or Plan(A, B)
    val, col = SpGEMMFunction.apply(a_val, b_val, A, B, plan)
    return plan, val, col


__all__ = ["DeviceCSR", "Workspace", "LibCSR", "Executor", "Plan", "spgemm", "spgemm_safe", "backward", "SpGEMMFunction",
           "spgemm_autograd", "default_workspace", "OpsTimings", "OpsError", "dtype_code", "batch_pack", "compute_flop_device",
           "default_tuning", "tuning_caps", "mem_cache_stats", "set_mem_cache_limit", "mem_cache_release",
           "set_kept_bitmap_limit"]
