"""Autotuner for the SpGEMM bin boundaries, hash family and kernel shapes.

The reference fixes its bin bounds and hash at compile time
(kernels/binning.cuh:21, 51, 105, 143; utils/config.hpp:32).  Here they are
workspace settings (ops_workspace_set_tuning); this module times a small set
of candidates on the caller's own matrices with the full drop-in call and
keeps the fastest.  Choices are cached per device and matrix signature in a
JSON file (default ~/.cache/opsparse_amd/tuning.json; OPSPARSE_AMD_TUNE_CACHE
overrides it), so a workload is tuned once per machine.

Candidate space (every candidate respects the kernels' capacities):
  * short-row hash: linear (col*107) or Fibonacci;
  * symbolic bounds: default, or each bound halved (rows move to wider
    kernels with larger tables earlier);
  * numeric bounds: default, or each bound halved;
  * symbolic bitmap for rows with a narrow column span: off or on;
then, from the best of those, one bin at a time (numeric bins 0-4, then
symbolic bins 0-4, only bins that hold rows): the kernel shape, i.e. hash
table size, lanes per row and block size (ops_tuning's shape table), kept
when it is more than 2% faster.
"""
from __future__ import annotations

import copy
import json
import os
import time

import torch

from .spgemm import Workspace, default_tuning, spgemm, tuning_caps


def _cache_path() -> str:
    return os.environ.get("OPSPARSE_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "opsparse_amd", "tuning.json"))


def signature(A, B) -> str:
    name = torch.cuda.get_device_properties(0).name if torch.cuda.is_available() else "cpu"
    return f"{name}|{A.M}x{A.N}x{B.N}|nnzA={A.nnz}|nnzB={B.nnz}|{str(A.val.dtype).replace('torch.', '')}"


def candidates() -> list[dict]:
    base = default_tuning()
    caps = tuning_caps()
    out = []
    for short_hash in (0, 1):
        for sym_half in (False, True):
            for num_half in (False, True):
                for sym_bitmap in (0, 1):
                    t = {"sym_bound": list(base["sym_bound"]), "num_bound": list(base["num_bound"]),
                         "short_hash": short_hash, "sym_bitmap": sym_bitmap,
                         "sym_shape": list(base["sym_shape"]), "num_shape": list(base["num_shape"])}
                    if sym_half:
                        t["sym_bound"] = [max(1, b // 2) for b in t["sym_bound"]]
                    if num_half:
                        t["num_bound"] = [max(1, b // 2) for b in t["num_bound"]]
                    assert all(v <= c for v, c in zip(t["sym_bound"], caps["sym_bound"]))
                    assert all(v <= c for v, c in zip(t["num_bound"], caps["num_bound"]))
                    out.append(t)
    return out


def _bins_in_use(A, B, ws: Workspace, wide: bool):
    from ._lib import OpsTimings
    t = OpsTimings()
    spgemm(A, B, ws=ws, wide=wide, timings=t).free()
    return [t.sym_rows[b] > 0 for b in range(5)], [t.num_rows[b] > 0 for b in range(5)]


def _time_call(A, B, ws: Workspace, wide: bool, reps: int) -> float:
    spgemm(A, B, ws=ws, wide=wide).free()  # warm-up
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        c = spgemm(A, B, ws=ws, wide=wide)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
        c.free()
    return best


def autotune(A, B, ws: Workspace | None = None, wide: bool = True, reps: int = 2, use_cache: bool = True,
             verbose: bool = False, shapes: bool = True) -> dict:
    """Picks and installs the fastest tuning for A*B on `ws`; returns
    {"tuning": ..., "seconds": ..., "from_cache": bool, "trials": [...]}."""
    from .spgemm import default_workspace
    ws = ws or default_workspace()
    key = signature(A, B)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            cache = json.load(open(path))
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        ws.set_tuning(cache[key]["tuning"])
        return {"tuning": cache[key]["tuning"], "seconds": cache[key]["seconds"], "from_cache": True, "trials": []}
    trials = []
    for t in candidates():
        ws.set_tuning(t)
        sec = _time_call(A, B, ws, wide, reps)
        trials.append({"tuning": t, "seconds": sec})
        if verbose:
            print(f"  autotune {t['short_hash']} sym0={t['sym_bound'][0]} num0={t['num_bound'][0]}: {sec * 1e3:.3f} ms",
                  flush=True)
    best = min(trials, key=lambda r: r["seconds"])
    # Timing noise should not move a workload off the default bounds: the
    # default is kept unless another candidate is more than 2% faster.
    default = next(r for r in trials if r["tuning"] == default_tuning())
    if default["seconds"] <= best["seconds"] * 1.02:
        best = default
    if shapes:
        ws.set_tuning(best["tuning"])
        sym_used, num_used = _bins_in_use(A, B, ws, wide)
        for kind, used in (("num_shape", num_used), ("sym_shape", sym_used)):
            for b in range(5):
                if not used[b]:
                    continue
                for sh in range(1, 3):
                    t = copy.deepcopy(best["tuning"])
                    t[kind][b] = sh
                    ws.set_tuning(t)
                    sec = _time_call(A, B, ws, wide, reps)
                    trials.append({"tuning": t, "seconds": sec})
                    if verbose:
                        print(f"  autotune {kind}[{b}]={sh}: {sec * 1e3:.3f} ms", flush=True)
                    if sec < best["seconds"] * 0.98:
                        best = {"tuning": t, "seconds": sec}
    ws.set_tuning(best["tuning"])
    if use_cache:
        cache[key] = {"tuning": best["tuning"], "seconds": best["seconds"]}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            json.dump(cache, open(path, "w"), indent=1)
        except OSError:
            pass
    return {"tuning": best["tuning"], "seconds": best["seconds"], "from_cache": False, "trials": trials}
