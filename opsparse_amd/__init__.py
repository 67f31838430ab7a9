"""opsparse_amd: SpGEMM (C = A*B in CSR) for AMD Instinct MI355X (gfx950).

The compute path is the HIP library ``libopsparse_amd.so`` (C ABI in
include/opsparse_amd.h); this package is its Python/PyTorch host side.
"""
from ._lib import LIB_PATH, OpsError, OpsTimings, lib  # noqa: F401
from . import gen  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # torch-facing API is loaded on first use
    if name in ("spgemm", "spgemm_safe", "Plan", "DeviceCSR", "Workspace", "backward", "SpGEMMFunction",
                "spgemm_autograd", "default_workspace", "LibCSR"):
        from . import spgemm as _s
        return getattr(_s, name)
    raise AttributeError(name)
