"""SpGEMM benchmark on one MI355X.

A step is one drop-in call C = A*B (ops_spgemm_f64: setup, binning,
symbolic, scan, allocation of C, numeric; inputs resident in HBM) followed
by freeing C, the reference's convention (benchmarks/opsparse_profile.cu:24-39:
warm-up, then timed iterations with C released between them).

Prints ONE JSON line.  value = 2 * products / time in GFLOP/s.  roofline
= the kernel with the largest HIP-event time in the timed region (events
recorded on the pool stream the kernel runs on), with its algorithmic bytes
per launch:
  numeric bin  : (4+w)*P + (4+w)*Z + (12+w)*nnzA + 28*rows
  symbolic bin : 4*P + 12*nnzA + 16*rows
(P products, Z output entries, nnzA A-nonzeros and rows of that bin; w the
value size).  cpu_baseline = the CPU oracle (OpenMP Gustavson) timed on this
host's cores over the same matrix.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SpGEMM GFLOP/s (2×intermediate products/s), % HBM roofline; batched & fwd+bwd"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="stencil", choices=["stencil", "uniform"])
    p.add_argument("--cpu-iters", type=int, default=6)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-check", action="store_true")
    p.add_argument("--copy-calibration", action="store_true",
                   help="after the run, copy exactly 1 GiB device-to-device (the known-size kernel that "
                        "FETCH_SIZE / WRITE_SIZE counter runs are scaled by)")
    return p.parse_args()


def workload(name):
    import opsparse_amd.gen as G
    if name == "stencil":
        A = G.stencil27(100, 1)
        cfg = {"workload": "C=A*A, A = 27-point stencil on 100^3 (1,000,000 rows), diag 26, off -1+U[-0.01,0.01], "
                           "FP64, seed 1", "rows": A.M, "nnz_a": A.nnz}
    else:
        A = G.uniform_rows(4096, 4096, 8, 0)
        cfg = {"workload": "C=A*A, A = 4096x4096, 8 uniform distinct columns per row, U[-1,1], FP64, seed 0",
               "rows": A.M, "nnz_a": A.nnz}
    return A, cfg


def kernel_bytes(kind, b, t, w):
    if kind == "numeric":
        return ((4 + w) * t.num_flop[b] + (4 + w) * t.num_nnz[b] + (12 + w) * t.num_arows_nnz[b]
                + 28 * t.num_rows[b])
    return 4 * t.sym_flop[b] + 12 * t.sym_arows_nnz[b] + 16 * t.sym_rows[b]


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding is "
              f"design-only in this build", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import opsparse_amd.gen as G
    from opsparse_amd._lib import OpsTimings, lib
    from opsparse_amd.spgemm import DeviceCSR, Workspace, spgemm

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    t_gen = time.time()
    A, cfg = workload(args.workload)
    P = G.compute_flop_host(A, A)
    t_gen = time.time() - t_gen
    w = 8
    dA = DeviceCSR.from_host(A, torch.float64)
    ws = Workspace()
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        spgemm(dA, dA, ws=ws).free()
    torch.cuda.synchronize()

    agg = {}
    per_call = []
    nnz_c = 0
    outs = []
    timings = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        t = OpsTimings()
        t.kernel_events = 1
        # Each call allocates and computes its own C; releasing C happens after
        # the timed region, as the reference releases it outside the timed call.
        outs.append(spgemm(dA, dA, ws=ws, timings=t))
        timings.append(t)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for t in timings:
        per_call.append(t.total)
        for kind, ms_arr in (("numeric", t.num_ms), ("symbolic", t.sym_ms)):
            for b in range(8):
                if ms_arr[b] > 0:
                    agg.setdefault((kind, b), []).append(ms_arr[b])
    c_last, t_last = outs[-1], timings[-1]
    nnz_c = c_last.nnz
    for c in outs[:-1]:
        c.free()
    ms_per_step = elapsed / args.steps * 1e3
    gflops = 2.0 * P / (elapsed / args.steps) / 1e9

    # dominant kernel and its roofline
    (kind, b), ms_list = max(agg.items(), key=lambda kv: float(np.mean(kv[1])))
    ms_k = float(np.mean(ms_list))
    bytes_k = kernel_bytes(kind, b, t_last, w)
    achieved = bytes_k / (ms_k * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"{kind} bin {b}", "kernel_ms": round(ms_k, 4), "bytes_per_launch": int(bytes_k)}
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{kind}_bin{b}_{args.workload}.txt")
    if os.path.exists(traffic_file):
        for line in open(traffic_file):
            if line.startswith("traffic_bytes_per_launch:"):
                roofline["traffic"] = int(float(line.split(":")[1]))

    # whole-pipeline algorithmic bytes (setup + symbolic + numeric)
    bytes_alg = (8 + w) * P + (4 + w) * nnz_c + (36 + w) * A.nnz + 32 * A.M
    pipe_frac = bytes_alg / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS

    check = None
    if not args.no_check:
        import oracle
        rpt, col, val = c_last.to_host()
        ref = oracle.spgemm(A, A)
        check = oracle.compare(rpt, col, val, ref, 1e-12)
    c_last.free()

    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        times = []
        oracle.spgemm(A, A, with_abs=False)  # warm-up
        for _ in range(args.cpu_iters):
            s = time.perf_counter()
            oracle.spgemm(A, A, with_abs=False)
            times.append(time.perf_counter() - s)
        tc = float(np.mean(times))
        cpu = {"value": round(2.0 * P / tc / 1e9, 3), "unit": "GFLOP/s", "cores": oracle.num_threads(),
               "kind": "port",
               "sample": f"full workload, OpenMP Gustavson (oracle), 1 warm-up + mean of {args.cpu_iters} runs, "
                         f"{tc * 1e3:.1f} ms/run, threads={oracle.num_threads()}, host cpus={os.cpu_count()}"}

    cfg.update({"products": int(P), "nnz_c": int(nnz_c), "gen_s": round(t_gen, 2),
                "ref_convention_ms": round(float(np.mean(per_call)) * 1e3, 4),
                "pipeline_roofline_frac": round(pipe_frac, 4), "bytes_alg": int(bytes_alg),
                "parallelism": "single GPU"})
    if check is not None:
        cfg["parity"] = {"structure": "exact", "max_norm_err": check["max_norm"], "max_rel_err": check["max_rel"]}
    out = {"metric": METRIC, "value": round(gflops, 3), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": cfg,
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    if args.copy_calibration:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
