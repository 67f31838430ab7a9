"""ctypes binding of the C ABI in include/opsparse_amd.h.

The shared library is built in-tree (``opsparse_amd/libopsparse_amd.so``,
see ``__graft_entry__.build``).  torch is imported first when it is
available so that the library binds to the same HIP runtime instance as
torch (both carry the SONAME libamdhip64.so.7); device pointers and streams
are then interchangeable between the two.
"""
from __future__ import annotations

import ctypes as C
import os

try:  # share torch's HIP runtime when torch is present
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is part of the image
    torch = None

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libopsparse_amd.so")

OPS_OK = 0
OPS_NONSQUARE = 1
OPS_INT32_OVERFLOW = 2
OPS_OOM = 3
OPS_HIP_ERROR = 4
OPS_EXCEPTION = 5
OPS_INVALID = 6
STATUS_NAMES = {0: "ok", 1: "nonsquare", 2: "int32_overflow", 3: "oom", 4: "hip_error", 5: "exception", 6: "invalid"}

OPS_F64 = 0
OPS_F32 = 1
NUM_BINS = 8


class OpsCsr(C.Structure):
    _fields_ = [("M", C.c_int32), ("N", C.c_int32), ("nnz", C.c_int32), ("rpt", C.c_void_p), ("col", C.c_void_p),
                ("val", C.c_void_p)]


class OpsCsr64(C.Structure):
    _fields_ = [("M", C.c_int64), ("N", C.c_int64), ("nnz", C.c_int64), ("rpt", C.c_void_p), ("col", C.c_void_p),
                ("val", C.c_void_p)]


class OpsTimings(C.Structure):
    _fields_ = [("measure_separate", C.c_int), ("kernel_events", C.c_int),
                ("setup", C.c_double), ("symbolic_binning", C.c_double), ("symbolic", C.c_double),
                ("reduce", C.c_double), ("numeric_binning", C.c_double), ("prefix", C.c_double),
                ("allocate", C.c_double), ("numeric", C.c_double), ("cleanup", C.c_double), ("total", C.c_double),
                ("total_flop", C.c_int64), ("nnz_c", C.c_int64), ("max_row_flop", C.c_int32),
                ("max_row_nnz", C.c_int32),
                ("sym_rows", C.c_int64 * NUM_BINS), ("sym_flop", C.c_int64 * NUM_BINS), ("sym_arows_nnz", C.c_int64 * NUM_BINS),
                ("sym_ms", C.c_double * NUM_BINS),
                ("num_rows", C.c_int64 * NUM_BINS), ("num_flop", C.c_int64 * NUM_BINS),
                ("num_nnz", C.c_int64 * NUM_BINS), ("num_arows_nnz", C.c_int64 * NUM_BINS),
                ("num_ms", C.c_double * NUM_BINS), ("kept_rows", C.c_int64)]


class OpsTuning(C.Structure):
    _fields_ = [("sym_bound", C.c_int32 * (NUM_BINS - 1)), ("num_bound", C.c_int32 * (NUM_BINS - 1)),
                ("short_hash", C.c_int32), ("sym_bitmap", C.c_int32),
                ("sym_shape", C.c_int32 * (NUM_BINS - 1)), ("num_shape", C.c_int32 * (NUM_BINS - 1))]

    def as_dict(self):
        return {"sym_bound": list(self.sym_bound), "num_bound": list(self.num_bound), "short_hash": self.short_hash,
                "sym_bitmap": self.sym_bitmap, "sym_shape": list(self.sym_shape), "num_shape": list(self.num_shape)}

    @staticmethod
    def from_dict(d):
        t = OpsTuning()
        for i, v in enumerate(d["sym_bound"]):
            t.sym_bound[i] = v
        for i, v in enumerate(d["num_bound"]):
            t.num_bound[i] = v
        t.short_hash = d["short_hash"]
        t.sym_bitmap = d.get("sym_bitmap", 0)
        for i, v in enumerate(d.get("sym_shape", [0] * (NUM_BINS - 1))):
            t.sym_shape[i] = v
        for i, v in enumerate(d.get("num_shape", [0] * (NUM_BINS - 1))):
            t.num_shape[i] = v
        return t


class OpsHostCsr(C.Structure):
    _fields_ = [("M", C.c_int32), ("N", C.c_int32), ("nnz", C.c_int64), ("rpt", C.POINTER(C.c_int32)),
                ("col", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double))]


class OpsMemStats(C.Structure):
    _fields_ = [("cached_bytes", C.c_int64), ("cached_blocks", C.c_int64), ("live_bytes", C.c_int64),
                ("live_blocks", C.c_int64), ("hits", C.c_int64), ("misses", C.c_int64), ("limit", C.c_int64)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


# name -> (restype, argtypes); every function declared in include/opsparse_amd.h
_vp = C.c_void_p
_P = C.POINTER
SIGNATURES = {
    "ops_workspace_create": (C.c_int, [_P(_vp)]),
    "ops_workspace_destroy": (C.c_int, [_vp]),
    "ops_workspace_last_error": (C.c_char_p, [_vp]),
    "ops_workspace_trim": (C.c_int, [_vp]),
    "ops_kept_bitmap_limit": (C.c_int, [C.c_int64]),
    "ops_spgemm_f64": (C.c_int, [_vp, _P(OpsCsr), _P(OpsCsr), _P(OpsCsr), _vp, _P(OpsTimings)]),
    "ops_spgemm_f32": (C.c_int, [_vp, _P(OpsCsr), _P(OpsCsr), _P(OpsCsr), _vp, _P(OpsTimings)]),
    "ops_spgemm64_f64": (C.c_int, [_vp, _P(OpsCsr), _P(OpsCsr), _P(OpsCsr64), _vp, _P(OpsTimings)]),
    "ops_spgemm64_f32": (C.c_int, [_vp, _P(OpsCsr), _P(OpsCsr), _P(OpsCsr64), _vp, _P(OpsTimings)]),
    "ops_spgemm64_noalloc": (C.c_int, [C.c_int, _P(OpsCsr), _P(OpsCsr), _P(OpsCsr64), C.c_int64, _vp]),
    "ops_csr_free": (C.c_int, [_P(OpsCsr)]),
    "ops_csr64_free": (C.c_int, [_P(OpsCsr64)]),
    "ops_free": (C.c_int, [_vp]),
    "ops_mem_cache_set_limit": (C.c_int, [C.c_int64]),
    "ops_mem_cache_release": (C.c_int, []),
    "ops_mem_cache_stats": (C.c_int, [_P(OpsMemStats)]),
    "ops_spgemm_safe_f64": (C.c_int, [_vp, _P(OpsCsr), _P(OpsCsr), _vp, _vp, _vp, _P(OpsCsr), _vp, _P(OpsTimings),
                                      _P(C.c_int64), _P(C.c_int64), _P(C.c_size_t), _P(C.c_size_t),
                                      _P(C.c_size_t)]),
    "ops_compute_flop_host": (C.c_int64, [C.c_int32, _vp, _vp, _vp, _vp]),
    "ops_compute_flop_device": (C.c_int, [_vp, _P(OpsCsr), _P(OpsCsr), _vp, _P(C.c_int64), _vp]),
    "ops_plan_create": (C.c_int, [_vp, _P(OpsCsr), _P(OpsCsr), _P(_vp), _P(C.c_int64), _vp]),
    "ops_plan_destroy": (C.c_int, [_vp]),
    "ops_plan_kept_rows": (C.c_int, [_vp, _P(C.c_int64)]),
    "ops_plan_rowptr32": (C.c_int, [_vp, _vp, _vp]),
    "ops_plan_rowptr64": (C.c_int, [_vp, _vp, _vp]),
    "ops_plan_numeric": (C.c_int, [_vp, C.c_int, _P(OpsCsr), _P(OpsCsr), _vp, _vp, _vp]),
    "ops_plan_numeric_graph": (C.c_int, [_vp, C.c_int, _P(OpsCsr), _P(OpsCsr), _vp, _vp, _vp]),
    "ops_spgemm_backward": (C.c_int, [_vp, C.c_int, _P(OpsCsr), _P(OpsCsr), _vp, _vp, _vp, _vp, _vp, _vp]),
    "ops_tuning_default": (C.c_int, [_P(OpsTuning)]),
    "ops_tuning_caps": (C.c_int, [_P(OpsTuning)]),
    "ops_workspace_set_tuning": (C.c_int, [_vp, _P(OpsTuning)]),
    "ops_workspace_get_tuning": (C.c_int, [_vp, _P(OpsTuning)]),
    "ops_batch_pack": (C.c_int, [C.c_int32, _P(OpsCsr), C.c_int, _P(OpsCsr), _P(C.c_int64), _P(C.c_int64), _vp]),
    "ops_exec_create": (C.c_int, [_vp, C.c_int, _P(OpsCsr), _P(OpsCsr), C.c_int64, C.c_int, _P(_vp)]),
    "ops_exec_create_batch": (C.c_int, [_vp, C.c_int, C.c_int32, _P(OpsCsr), _P(OpsCsr), C.c_int64, C.c_int,
                                        _P(C.c_int64), _P(_vp)]),
    "ops_exec_launch": (C.c_int, [_vp, _vp]),
    "ops_exec_result": (C.c_int, [_vp, _P(OpsCsr64), _P(C.c_int64), _vp]),
    "ops_exec_destroy": (C.c_int, [_vp]),
    "ops_gen_uniform_rows": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _P(OpsHostCsr)]),
    "ops_gen_stencil27": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                    C.c_uint64, _P(OpsHostCsr)]),
    "ops_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                               _P(OpsHostCsr)]),
    "ops_gen_batch": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                _P(OpsHostCsr), _P(OpsHostCsr), _P(C.c_int64), _P(C.c_int32)]),
    "ops_host_csr_free": (C.c_int, [_P(OpsHostCsr)]),
    "ops_version": (C.c_char_p, []),
}

_lib = None


class OpsError(RuntimeError):
    def __init__(self, status: int, msg: str = ""):
        self.status = status
        super().__init__(f"opsparse_amd status {status} ({STATUS_NAMES.get(status, '?')}): {msg}")


def lib() -> C.CDLL:
    """Loads the built library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                              f"g.build()'` or `make -C opsparse_amd/csrc`")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(status: int, ws=None) -> None:
    if status != OPS_OK:
        msg = ""
        if ws:
            raw = lib().ops_workspace_last_error(ws)
            msg = raw.decode() if raw else ""
        raise OpsError(status, msg)
