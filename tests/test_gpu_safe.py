"""opsparse_spgemm_safe on the GPU (src/opsparse/core_algo/spgemm/spgemm.cu:146-219):
status ok against the oracle, oom when the device cannot hold the output,
and the same workspace producing a correct product once memory is back.

The oom case is forced from the host: the test holds all but 192 MiB of
device memory in one tensor, so the output's hipMalloc (660 MiB) reports
out-of-memory, which the call maps to OPS_OOM as the reference maps
cudaErrorMemoryAllocation (spgemm.cu:187-199).

nnz(C) never exceeds the product count, so once the host pre-check has
passed (products <= INT32_MAX) the reference's post-run test for a wrapped
nnz(C) (spgemm.cu:208-213) cannot fire; the int32_overflow status is
covered through the pre-check (test_capi, test_drivers).
"""
import numpy as np
import pytest
import torch

import oracle
import opsparse_amd.gen as G
from opsparse_amd._lib import OPS_OK, OPS_OOM
from opsparse_amd.spgemm import DeviceCSR, Workspace, mem_cache_release, spgemm_safe

from helpers import TOL

pytestmark = pytest.mark.gpu


def test_safe_ok_oom_and_recovery(gpu):
    A = G.rmat(15, 16, 7)
    ref = oracle.spgemm(A, A)
    dA = DeviceCSR.from_host(A, torch.float64)
    ws = Workspace()
    st, C, info = spgemm_safe(dA, dA, A, A, ws)
    assert st == OPS_OK, st
    assert info["total_flop"] == oracle.compute_flop(A, A) and info["c_nnz"] == len(ref[1])
    oracle.compare(*C.to_host(), ref, TOL[np.float64])
    c_bytes = info["c_nnz"] * 12
    del C
    mem_cache_release()
    torch.cuda.synchronize()

    free, _ = torch.cuda.mem_get_info()
    keep = 192 << 20
    assert c_bytes > 3 * keep
    filler = torch.empty(free - keep, dtype=torch.uint8, device="cuda")
    try:
        ws2 = Workspace()
        st, C, info = spgemm_safe(dA, dA, A, A, ws2)
        assert st == OPS_OOM and C is None, st
    finally:
        del filler
        torch.cuda.empty_cache()
    st, C, info = spgemm_safe(dA, dA, A, A, ws2)
    assert st == OPS_OK, st
    oracle.compare(*C.to_host(), ref, TOL[np.float64])
    del C
    mem_cache_release()
