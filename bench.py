"""SpGEMM benchmark on one MI355X.

Default workload: C = A*A for the R-MAT 2^20 matrix (1,048,576 rows,
(a,b,c,d) = (.57,.19,.19,.05), 16 * 2^20 edges merged, FP64, seed 2), the
largest configuration that fits one GPU: 2.09e10 products and 9.7e9 output
entries, so C takes 64-bit row offsets (ops_spgemm64_f64).  A step is one
call (setup, binning, symbolic, scan, allocation of C, numeric) with inputs
resident in HBM; C is released outside the timed region, as the reference
releases it outside its timed call (benchmarks/opsparse_profile.cu:24-39).
The 27-point stencil on 100^3 (FP64, seed 1) and the 4096-instance batch
(FP64: fused call, packing included, the whole call replayed from one
hipGraph, the looped plain call) are measured after it and kept as nested
records (config.stencil_100, config.batch_4096).

Other workloads (--workload):
  stencil  : the stencil alone
  uniform  : 4096^2, 8 distinct uniform columns per row, FP64, seed 0
  batch    : 4096 independent C_i = A_i*B_i, one fused autotuned call
             (also timed with the packing of the instances inside), the
             whole call (packing, product, shift of each C_i to its own
             columns) replayed from one hipGraph with bin sizes and nnz(C)
             on the device, against the plain call looped over the
             instances; FP64 line, FP32 in the config
  backward : R-MAT 2^18 A*B in FP32, forward + masked backward, gradients
             checked against FP64 direct sums and FP64 finite differences

Prints ONE JSON line.  value = 2 * products / time in GFLOP/s, time = the
mean of the timed calls, each bracketed by synchronisations.  roofline = the
kernel with the largest mean HIP-event time over three diagnostic calls made
after the timed ones (events recorded on the pool stream it runs on; the
timed calls record no events), with its algorithmic bytes per launch
(frac = those bytes / event time / 8 TB/s; traffic_frac = the measured HBM
bytes of the committed counter summary / event time / 8 TB/s):
  numeric bin  : (4+w)*P + (4+w)*Z + (12+w)*nnzA + 28*rows
  symbolic bin : 4*P + 12*nnzA + 16*rows
(P products, Z output entries, nnzA A-nonzeros and rows of that bin; w the
value size).  traffic = measured HBM bytes per launch of that kernel when a
counter summary for it is committed under profiles/.  cpu_baseline = the CPU
oracle (OpenMP Gustavson, threads bound with OMP_PROC_BIND=spread) timed on
this host's cores over the workload or a stated row sample of it; the same
oracle output checks the GPU rows.  cold_* = one call with the output cache
emptied (and with a new workspace), the reference's per-call allocation.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The CPU baseline's OpenMP threads are spread over the host's cores; set
# before any OpenMP runtime in this process starts.
os.environ.setdefault("OMP_PROC_BIND", "spread")

HBM_PEAK_GBS = 8000.0
METRIC = "SpGEMM GFLOP/s (2×intermediate products/s), % HBM roofline; batched & fwd+bwd"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="rmat20", choices=["stencil", "uniform", "rmat20", "batch", "backward"])
    p.add_argument("--no-stencil", action="store_true", help="rmat20: skip the nested stencil and batch records")
    p.add_argument("--cpu-iters", type=int, default=6)
    p.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle (timing and row checks)")
    p.add_argument("--no-check", action="store_true", help="skip output checks")
    p.add_argument("--batch-count", type=int, default=4096)
    p.add_argument("--fd-entries", type=int, default=1024)
    p.add_argument("--vendor", action="store_true", help="also time rocSPARSE SpGEMM (default for stencil/uniform)")
    p.add_argument("--no-vendor", action="store_true", help="skip the rocSPARSE comparison")
    p.add_argument("--copy-calibration", action="store_true",
                   help="after the run, copy exactly 1 GiB device-to-device (the known-size kernel that "
                        "FETCH_SIZE / WRITE_SIZE counter runs are scaled by)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ----------------------------------------------------------------------------- timing helpers

def kernel_bytes(kind, b, t, w):
    if kind == "numeric":
        return ((4 + w) * t.num_flop[b] + (4 + w) * t.num_nnz[b] + (12 + w) * t.num_arows_nnz[b]
                + 28 * t.num_rows[b])
    return 4 * t.sym_flop[b] + 12 * t.sym_arows_nnz[b] + 16 * t.sym_rows[b]


def roofline_of(timings, w, workload_tag):
    """Dominant kernel over the diagnostic calls: the bin with the most
    algorithmic bytes (bins run on concurrent streams, so a bin's event
    span also holds the time its blocks waited for CUs the long-row bin
    held), timed by its per-bin event span."""
    import numpy as np
    agg = {}
    for t in timings:
        for kind, ms_arr in (("numeric", t.num_ms), ("symbolic", t.sym_ms)):
            for b in range(8):
                if ms_arr[b] > 0:
                    agg.setdefault((kind, b), []).append(ms_arr[b])
    (kind, b), ms_list = max(agg.items(), key=lambda kv: (kernel_bytes(kv[0][0], kv[0][1], timings[-1], w),
                                                          float(np.mean(kv[1]))))
    ms_k = float(np.mean(ms_list))
    bytes_k = kernel_bytes(kind, b, timings[-1], w)
    achieved = bytes_k / (ms_k * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"{kind} bin {b}", "kernel_ms": round(ms_k, 4), "bytes_per_launch": int(bytes_k)}
    tf = os.path.join(ROOT, "profiles", f"traffic_{kind}_bin{b}_{workload_tag}.txt")
    if os.path.exists(tf):
        for ln in open(tf):
            if ln.startswith("traffic_bytes_per_launch:"):
                roof["traffic"] = int(float(ln.split(":")[1]))
    if roof["traffic"]:
        # bandwidth the kernel actually drew from HBM (measured bytes, not the
        # no-reuse model): below `frac` when B's rows are served from caches
        roof["traffic_frac"] = round(roof["traffic"] / (ms_k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return roof


def timed_calls(A, B, ws, steps, warmup, wide, diag_calls=3):
    """warmup + steps drop-in calls; returns (timed seconds, timings of the diagnostic
    calls, last C, host timings of the timed calls).

    As in the reference's profile loop (benchmarks/opsparse_profile.cu:24-39),
    each timed call is bracketed by synchronisations and its output is
    released between calls, outside the timed spans; the seconds are their
    sum.  The timed calls record host phase times only.  Per-bin kernel
    event times (a diagnostic whose event records and waits would add to
    the call) come from `diag_calls` further untimed calls, whose timings
    are the ones returned (phases and the roofline kernel read them)."""
    import torch
    from opsparse_amd._lib import OpsTimings
    from opsparse_amd.spgemm import spgemm
    for _ in range(max(warmup, 1)):
        spgemm(A, B, ws=ws, wide=wide).free()
    torch.cuda.synchronize()
    elapsed, last = 0.0, None
    timed = []
    for _ in range(steps):
        if last is not None:
            last.free()
        t = OpsTimings()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        last = spgemm(A, B, ws=ws, timings=t, wide=wide)
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        timed.append(t)
    tims = []
    for _ in range(diag_calls):
        last.free()
        t = OpsTimings()
        t.kernel_events = 1
        last = spgemm(A, B, ws=ws, timings=t, wide=wide)
        torch.cuda.synchronize()
        tims.append(t)
    return elapsed, tims, last, timed


def cpu_leg(A, B, r0, r1, iters, desc, gpu_c, tol, check):
    """Times the CPU oracle on rows [r0, r1) of A*B and, unless check is
    False, compares the GPU result's rows with the oracle's."""
    import numpy as np
    import oracle
    import opsparse_amd.gen as G
    full = r0 == 0 and r1 == A.M

    def run(with_abs):
        return oracle.spgemm(A, B, with_abs) if full else oracle.spgemm_rows(A, B, r0, r1, with_abs)

    Ps = G.compute_flop_host(A if full else A.row_slice(r0, r1), B)
    ref = run(True)  # warm-up run; its |a*b| sums normalise the check
    times = []
    for _ in range(iters):
        s = time.perf_counter()
        run(False)
        times.append(time.perf_counter() - s)
    tc = float(np.mean(times))
    cpu = {"value": round(2.0 * Ps / tc / 1e9, 3), "unit": "GFLOP/s", "cores": oracle.num_threads(), "kind": "port",
           "sample": f"{desc} ({Ps} products); OpenMP Gustavson (oracle), 1 warm-up + mean of {iters} runs, "
                     f"{tc * 1e3:.1f} ms/run, threads={oracle.num_threads()}, host cpus={os.cpu_count()}"}
    err = None
    if check:
        rpt, col, val = gpu_c.rows_to_host(r0, r1)
        err = oracle.compare(rpt, col, val, ref, tol)
    return cpu, err


def phase_ms(tims, host=None):
    """Mean host-side phase times (ms) of `host` (default: tims) and the last of `tims`' per-bin
    kernel event times.  The host phases are read in calls that do not synchronise between
    phases, so "symbolic_and_scan" runs from the end of setup to the end of the scan (the
    symbolic kernels and the scan, waited for together)."""
    import numpy as np
    host = host or tims
    mean = lambda f: round(float(np.mean([getattr(t, f) for t in host])) * 1e3, 3)
    return {"setup": mean("setup"), "symbolic_and_scan": mean("prefix"), "allocate_c": mean("allocate"),
            "numeric": mean("numeric"), "total": mean("total"),
            "symbolic_bin_ms": [round(x, 3) for x in tims[-1].sym_ms],
            "numeric_bin_ms": [round(x, 3) for x in tims[-1].num_ms],
            "symbolic_bin_rows": list(tims[-1].sym_rows), "numeric_bin_rows": list(tims[-1].num_rows)}


def line(value, steps, warmup, ms_per_step, dtype, cfg, roof, cpu):
    return {"metric": METRIC, "value": round(value, 3), "unit": "GFLOP/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": cfg, "roofline": roof,
            "cpu_baseline": cpu}


# ----------------------------------------------------------------------------- plain products

def vendor_leg(dA, dB, dtype_f32, reps, ours_ms, ours_nnz):
    """rocSPARSE SpGEMM on the same device inputs (benchmark-only comparator
    in its own library, libvendor_rocsparse.so): mean ms per complete product
    over `reps` calls after one warm-up, with C allocated inside the call as
    for ours.  Also checks that its nnz(C) equals ours."""
    import ctypes as C
    import os
    from opsparse_amd._lib import OPS_F32, OPS_F64, OpsCsr
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "opsparse_amd", "libvendor_rocsparse.so")
    if not os.path.exists(path):
        return {"available": False, "reason": "libvendor_rocsparse.so not built"}
    lib = C.CDLL(path)
    f = lib.vendor_rocsparse_spgemm
    f.restype = C.c_int
    f.argtypes = [C.POINTER(OpsCsr), C.POINTER(OpsCsr), C.c_int, C.c_int, C.POINTER(C.c_double),
                  C.POINTER(C.c_int64), C.POINTER(OpsCsr)]
    a, b = dA.desc(), dB.desc()
    sec, nnz = C.c_double(), C.c_int64()
    st = f(C.byref(a), C.byref(b), OPS_F32 if dtype_f32 else OPS_F64, reps, C.byref(sec), C.byref(nnz), None)
    if st != 0:
        return {"available": False, "reason": f"rocsparse_spgemm failed (status {st})"}
    ms = sec.value * 1e3
    return {"available": True, "library": "rocSPARSE generic SpGEMM (rocsparse_spgemm, alg default)",
            "ms_per_call": round(ms, 4), "nnz_c": int(nnz.value), "nnz_matches": int(nnz.value) == int(ours_nnz),
            "speedup_ours_vs_rocsparse": round(ms / ours_ms, 2), "calls": reps}


def cold_calls(dA, ws, wide):
    """The first call's cost when its memory is not yet held, in ms: with
    the output cache emptied (C's arrays come from hipMalloc, whose fresh
    VRAM the driver clears, as the reference's per-call cudaMalloc of C does,
    src/opsparse/core_algo/spgemm/spgemm.cu:44-45), and additionally with a
    new workspace (its buffers allocated again, as the reference's
    SpgemmMeta is in every call, meta.cu:16-48)."""
    import torch
    from opsparse_amd.spgemm import Workspace, mem_cache_release, spgemm
    out = {}
    for key, fresh_ws in (("cold_output_ms", False), ("cold_output_and_workspace_ms", True)):
        torch.cuda.synchronize()
        mem_cache_release()
        w = Workspace() if fresh_ws else ws
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c = spgemm(dA, dA, ws=w, wide=wide)
        torch.cuda.synchronize()
        out[key] = round((time.perf_counter() - t0) * 1e3, 3)
        c.free()
        if fresh_ws:
            del w
    mem_cache_release()
    return out


def run_single(args, name):
    """stencil / uniform / rmat20: one drop-in call per step.  Returns
    (GFLOP/s, ms per call, config, roofline, cpu_baseline)."""
    import numpy as np
    import torch
    import opsparse_amd.gen as G
    from opsparse_amd.spgemm import DeviceCSR, Workspace
    t_gen = time.time()
    if name == "stencil":
        A = G.stencil27(100, 1)
        desc = ("C=A*A, A = 27-point stencil on 100^3 (1,000,000 rows), diag 26, off -1+U[-0.01,0.01], FP64, "
                "seed 1")
    elif name == "uniform":
        A = G.uniform_rows(4096, 4096, 8, 0)
        desc = "C=A*A, A = 4096x4096, 8 uniform distinct columns per row, U[-1,1], FP64, seed 0"
    else:
        A = G.rmat(20, 16, 2)
        desc = ("C=A*A, A = R-MAT 2^20 x 2^20, (a,b,c,d)=(.57,.19,.19,.05), 16*2^20 edges merged, no vertex "
                "permutation, U[-1,1], FP64, seed 2, 64-bit row offsets")
    P = G.compute_flop_host(A, A)
    t_gen = time.time() - t_gen
    wide = name == "rmat20"
    w = 8
    dA = DeviceCSR.from_host(A, torch.float64)
    ws = Workspace()
    torch.cuda.synchronize()
    log(f"{name}: rows {A.M} nnz {A.nnz} products {P}")
    elapsed, tims, c_last, timed = timed_calls(dA, dA, ws, args.steps, args.warmup, wide)
    nnz_c = c_last.nnz
    ms_per_step = elapsed / args.steps * 1e3
    gflops = 2.0 * P / (elapsed / args.steps) / 1e9
    log(f"{name}: {ms_per_step:.3f} ms/call, {gflops:.1f} GFLOP/s, nnz(C) {nnz_c}")
    roof = roofline_of(tims, w, name)
    bytes_alg = (8 + w) * P + (4 + w) * nnz_c + (36 + w) * A.nnz + 32 * A.M
    # Compulsory bytes: read A and B once, write C once (a lower bound no
    # SpGEMM can go below).
    bytes_comp = (4 + w) * (2 * A.nnz + nnz_c) + 4 * (A.M + 1) * 3
    t_step = elapsed / args.steps
    cfg = {"workload": desc, "rows": A.M, "nnz_a": A.nnz, "products": int(P), "nnz_c": int(nnz_c),
           "gen_s": round(t_gen, 2), "ref_convention_ms": round(float(np.mean([t.total for t in timed])) * 1e3, 4),
           "pipeline_roofline_frac": round(bytes_alg / t_step / 1e9 / HBM_PEAK_GBS, 4),
           "bytes_alg": int(bytes_alg), "compulsory_bytes": int(bytes_comp),
           "compulsory_roofline_frac": round(bytes_comp / t_step / 1e9 / HBM_PEAK_GBS, 4),
           "kept_bitmap_rows": int(tims[-1].kept_rows),
           "phases": phase_ms(tims, timed), "parallelism": "single GPU"}
    if args.vendor or (name != "rmat20" and not args.no_vendor):
        torch.cuda.synchronize()
        cfg["vendor_rocsparse"] = vendor_leg(dA, dA, False, max(1, min(args.steps, 5)), ms_per_step, nnz_c)
        log(f"{name}: rocSPARSE {cfg['vendor_rocsparse']}")
    cpu = None
    if not args.no_cpu_baseline:
        if name == "rmat20":
            r1 = 1024
            cpu, err = cpu_leg(A, A, 0, r1, max(1, args.cpu_iters // 3),
                               "rows 0..1023 of A times A (the highest-degree rows)", c_last, 1e-12, not args.no_check)
            if err is not None:
                import oracle
                worst = err["max_norm"]
                for s0 in (4096, 262144, 700000, A.M - 4096):  # untimed blocks across the matrix
                    ref = oracle.spgemm_rows(A, A, s0, s0 + 4096)
                    rpt, col, val = c_last.rows_to_host(s0, s0 + 4096)
                    worst = max(worst, oracle.compare(rpt, col, val, ref, 1e-12)["max_norm"])
                # every row offset against the oracle's symbolic count
                t0 = time.time()
                rows_exact = bool(np.array_equal(c_last.rowptr_to_host(), oracle.symbolic(A, A)))
                cfg["parity"] = {"structure": "row offsets of all rows " + ("exact" if rows_exact else "DIFFER")
                                              + "; columns exact on rows 0..1023 and 4 blocks of 4096 rows",
                                 "max_norm_err": worst, "row_offsets_check_s": round(time.time() - t0, 1)}
                assert rows_exact, "row offsets differ from the oracle"
        else:
            cpu, err = cpu_leg(A, A, 0, A.M, args.cpu_iters, "full workload", c_last, 1e-12, not args.no_check)
            if err is not None:
                cfg["parity"] = {"structure": "exact", "max_norm_err": err["max_norm"], "max_rel_err": err["max_rel"]}
    c_last.free()
    cfg.update(cold_calls(dA, ws, wide))
    log(f"{name}: cold calls {cfg['cold_output_ms']} / {cfg['cold_output_and_workspace_ms']} ms")
    ws.trim()
    return gflops, ms_per_step, cfg, roof, cpu


def nested_record(gflops, ms, cfg, roof, cpu):
    """A secondary workload's figures, kept inside the headline's config."""
    keep = ("workload", "products", "nnz_c", "pipeline_roofline_frac", "compulsory_roofline_frac",
            "cold_output_ms", "cold_output_and_workspace_ms", "parity", "vendor_rocsparse", "phases")
    rec = {"gflops": round(gflops, 3), "ms_per_call": round(ms, 4)}
    rec.update({k: cfg[k] for k in keep if k in cfg})
    rec["roofline"] = roof
    rec["cpu_baseline"] = cpu
    return rec


# ----------------------------------------------------------------------------- batched

def batch_instances(dX, row_base, dtype):
    """Descriptors of the instances of a packed square batch in their own
    coordinates: instance i is rows [row_base[i], row_base[i+1]) of dX (a
    row slice: its rpt starts at the instance's first entry) with columns
    shifted back by row_base[i] (square instances: column bases = row bases)."""
    import numpy as np
    import torch
    from opsparse_amd.spgemm import DeviceCSR
    rb = torch.from_numpy(np.asarray(row_base, dtype=np.int64)).to(dX.rpt.device)
    rows_inst = torch.repeat_interleave(torch.arange(len(row_base) - 1, device=rb.device), rb[1:] - rb[:-1])
    ent_inst = torch.repeat_interleave(rows_inst, (dX.rpt[1:] - dX.rpt[:-1]).long())
    local_col = (dX.col.long() - rb[ent_inst]).to(torch.int32)
    ends = dX.rpt.cpu().numpy()
    out = []
    for i in range(len(row_base) - 1):
        r0, r1 = int(row_base[i]), int(row_base[i + 1])
        out.append(DeviceCSR(r1 - r0, r1 - r0, dX.rpt[r0:r1 + 1], local_col, dX.val, int(ends[r1] - ends[r0])))
    return out, local_col


def run_batch(args, nested=False):
    """Independent C_i = A_i*B_i packed block-diagonally: one fused call
    (also with the packing of the instances inside the timed span), the
    whole call replayed from a hipGraph with device-side bin sizes
    (packing, product and the shift of each C_i to its own columns), and
    the replayed numeric hipGraph of a plan, against the plain call looped
    over the instances (row-slice descriptors of the packed A).

    nested: FP64 only, no CPU leg, and a compact record (for the default
    run's config) instead of a JSON line."""
    import numpy as np
    import torch
    import opsparse_amd.gen as G
    from opsparse_amd import autotune
    from opsparse_amd.spgemm import DeviceCSR, Executor, Plan, Workspace, batch_pack, spgemm
    t_gen = time.time()
    b = G.batch(args.batch_count, 3, 64, 16384, 4, 32)
    P = G.compute_flop_host(b.A, b.B)
    t_gen = time.time() - t_gen
    n_cpu = min(256, args.batch_count)
    log(f"batch: {args.batch_count} instances, rows {b.A.M}, nnz(A) {b.A.nnz}, products {P}")
    ws = Workspace()
    res = {}
    kinds = ((torch.float64, 8, "f64"),) if nested else ((torch.float64, 8, "f64"), (torch.float32, 4, "f32"))
    for dt, w, tag in kinds:
        npdt = np.float64 if dt == torch.float64 else np.float32
        Ah, Bh = b.A.astype(npdt), b.B.astype(npdt)
        dA = DeviceCSR.from_host(Ah, dt)
        dB = DeviceCSR.from_host(Bh, dt)
        tune = autotune.autotune(dA, dB, ws=ws, wide=True, reps=3)
        log(f"batch {tag}: tuning {tune['tuning']} (cached: {tune['from_cache']})")
        elapsed, tims, c_last, _ = timed_calls(dA, dB, ws, args.steps, args.warmup, True)
        t_fused = elapsed / args.steps
        nnz_c = c_last.nnz
        log(f"batch {tag}: fused {t_fused * 1e3:.3f} ms, nnz(C) {nnz_c}")
        cpu, err = None, None
        if not args.no_cpu_baseline and not nested:
            r1 = int(b.row_base[n_cpu])
            cpu, err = cpu_leg(Ah, Bh, 0, r1, max(1, args.cpu_iters // 2), f"instances 0..{n_cpu - 1}", c_last,
                               1e-12 if w == 8 else 1e-5, not args.no_check)
            if err is not None:
                import oracle
                for i in (args.batch_count // 2, args.batch_count - 1):  # untimed instances further in
                    s0, s1 = int(b.row_base[i]), int(b.row_base[i + 1])
                    rpt, col, val = c_last.rows_to_host(s0, s1)
                    e = oracle.compare(rpt, col, val, oracle.spgemm_rows(Ah, Bh, s0, s1), 1e-12 if w == 8 else 1e-5)
                    err["max_norm"] = max(err["max_norm"], e["max_norm"])
        c_last.free()
        # the instances as separate descriptors, packed inside the timed span
        iA, colA = batch_instances(dA, b.row_base, dt)
        iB, colB = batch_instances(dB, b.row_base, dt)
        t_pack, t_packed_call = 0.0, 0.0
        for rep in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pa, _, _ = batch_pack(iA, dt)
            pb, _, _ = batch_pack(iB, dt)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            c = spgemm(pa, pb, ws=ws, wide=True)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if rep >= args.warmup:
                t_pack += t1 - t0
                t_packed_call += t2 - t0
            c.free()
            pa.free()
            pb.free()
        t_pack /= args.steps
        t_packed_call /= args.steps
        # the whole call (pack, product, column shift) replayed from one hipGraph
        ex = Executor.batch(iA, iB, dt, ws=ws)
        for _ in range(args.warmup):
            ex.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ex.launch()
        torch.cuda.synchronize()
        t_exec = (time.perf_counter() - t0) / args.steps
        st, cx = ex.result()
        exec_err = None
        if st != 0:
            raise RuntimeError(f"batch executor status {st}")
        if not args.no_check:
            import oracle
            exec_err = 0.0
            for i in (0, args.batch_count // 2, args.batch_count - 1):
                s0, s1 = int(b.row_base[i]), int(b.row_base[i + 1])
                rpt, col, val = cx.rows_to_host(s0, s1)
                ref = oracle.spgemm_rows(Ah, Bh, s0, s1)
                # the oracle's columns are packed ones; the executor's are the instance's own
                ref = (ref[0], ref[1] - s0, ref[2], ref[3])
                e = oracle.compare(rpt, col, val, ref, 1e-12 if w == 8 else 1e-5)
                exec_err = max(exec_err, e["max_norm"])
        log(f"batch {tag}: pack {t_pack * 1e3:.3f} ms, pack + fused call {t_packed_call * 1e3:.3f} ms, "
            f"whole-call graph {t_exec * 1e3:.3f} ms")
        ex.destroy()
        del iA, iB, colA, colB, cx
        # structure reuse: the numeric phase of the fused plan replayed as a hipGraph
        plan = Plan(dA, dB, ws=ws)
        col = torch.empty(max(plan.nnz, 1), dtype=torch.int32, device="cuda")
        val = torch.empty(max(plan.nnz, 1), dtype=dt, device="cuda")
        plan.numeric(dA, dB, col, val, graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.numeric(dA, dB, col, val, graph=True)
        torch.cuda.synchronize()
        t_graph = (time.perf_counter() - t0) / args.steps
        del col, val, plan
        torch.cuda.empty_cache()
        # the plain call looped over the instances
        for i in range(2):  # warm-up
            spgemm(DeviceCSR(int(b.row_base[i + 1] - b.row_base[i]), dA.N,
                             dA.rpt[int(b.row_base[i]):int(b.row_base[i + 1]) + 1], dA.col, dA.val), dB, ws=ws).free()
        t_loop = 0.0
        for i in range(args.batch_count):
            r0, r1 = int(b.row_base[i]), int(b.row_base[i + 1])
            Ai = DeviceCSR(r1 - r0, dA.N, dA.rpt[r0:r1 + 1], dA.col, dA.val)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ci = spgemm(Ai, dB, ws=ws)
            torch.cuda.synchronize()
            t_loop += time.perf_counter() - t0
            ci.free()
        log(f"batch {tag}: graph numeric {t_graph * 1e3:.3f} ms, looped plain call {t_loop * 1e3:.1f} ms")
        res[tag] = {"t_fused": t_fused, "t_graph": t_graph, "t_loop": t_loop, "tims": tims, "nnz_c": nnz_c,
                    "t_pack": t_pack, "t_packed_call": t_packed_call, "t_exec": t_exec, "exec_err": exec_err,
                    "phases": phase_ms(tims),
                    "tuning": tune["tuning"], "cpu": cpu, "err": err}
        del dA, dB
        torch.cuda.empty_cache()
    r = res["f64"]
    roof = roofline_of(r["tims"], 8, "batch")
    if nested:
        return {"workload": f"{args.batch_count} independent C_i=A_i*B_i, n_i log-uniform in [64,16384], nnz/row "
                            f"uniform in [4,32] per matrix, FP64, seed 3",
                "products": int(P), "nnz_c": int(r["nnz_c"]), "gflops_fused": round(2.0 * P / r["t_fused"] / 1e9, 3),
                "fused_ms": round(r["t_fused"] * 1e3, 3), "pack_and_fused_ms": round(r["t_packed_call"] * 1e3, 3),
                "whole_call_graph_ms": round(r["t_exec"] * 1e3, 3),
                "graph_numeric_replay_ms": round(r["t_graph"] * 1e3, 3),
                "looped_plain_call_ms": round(r["t_loop"] * 1e3, 3),
                "speedup_whole_call_graph_vs_loop": round(r["t_loop"] / r["t_exec"], 2),
                "speedup_pack_and_fused_vs_loop": round(r["t_loop"] / r["t_packed_call"], 2),
                "speedup_fused_vs_loop": round(r["t_loop"] / r["t_fused"], 2),
                "tuning": r["tuning"], "roofline": roof,
                "parity": {"whole_call_graph": "structure exact (own columns) on instances 0, middle and last",
                           "max_norm_err": r["exec_err"]}}
    r32 = res["f32"]
    cfg = {"workload": f"{args.batch_count} independent C_i=A_i*B_i, n_i log-uniform in [64,16384], nnz/row uniform "
                       f"in [4,32] per matrix, uniform distinct columns, U[-1,1], seed 3; packed block-diagonally "
                       f"and run as one fused, autotuned call (FP64 line; FP32 under fp32_*)",
           "instances": args.batch_count, "rows": int(b.A.M), "nnz_a": int(b.A.nnz), "nnz_b": int(b.B.nnz),
           "products": int(P), "nnz_c": int(r["nnz_c"]), "gen_s": round(t_gen, 1),
           "fused_ms": round(r["t_fused"] * 1e3, 3), "pack_ms": round(r["t_pack"] * 1e3, 3),
           "pack_and_fused_ms": round(r["t_packed_call"] * 1e3, 3),
           "whole_call_graph_ms": round(r["t_exec"] * 1e3, 3),
           "speedup_whole_call_graph_vs_loop": round(r["t_loop"] / r["t_exec"], 2),
           "speedup_pack_and_fused_vs_loop": round(r["t_loop"] / r["t_packed_call"], 2),
           "graph_numeric_replay_ms": round(r["t_graph"] * 1e3, 3),
           "looped_plain_call_ms": round(r["t_loop"] * 1e3, 3),
           "speedup_fused_vs_loop": round(r["t_loop"] / r["t_fused"], 2),
           "speedup_graph_replay_vs_loop": round(r["t_loop"] / r["t_graph"], 2),
           "fp32_fused_ms": round(r32["t_fused"] * 1e3, 3),
           "fp32_pack_and_fused_ms": round(r32["t_packed_call"] * 1e3, 3),
           "fp32_whole_call_graph_ms": round(r32["t_exec"] * 1e3, 3),
           "fp32_speedup_whole_call_graph_vs_loop": round(r32["t_loop"] / r32["t_exec"], 2),
           "fp32_gflops": round(2.0 * P / r32["t_fused"] / 1e9, 3),
           "fp32_graph_numeric_replay_ms": round(r32["t_graph"] * 1e3, 3),
           "fp32_looped_plain_call_ms": round(r32["t_loop"] * 1e3, 3),
           "fp32_speedup_fused_vs_loop": round(r32["t_loop"] / r32["t_fused"], 2),
           "tuning_f64": r["tuning"], "tuning_f32": r32["tuning"], "phases_f64": r["phases"],
           "phases_f32": r32["phases"], "parallelism": "single GPU"}
    if r["err"] is not None:
        cfg["parity"] = {"structure": f"exact on instances 0..{n_cpu - 1}, middle and last",
                         "max_norm_err_f64": r["err"]["max_norm"], "max_norm_err_f32": r32["err"]["max_norm"],
                         "whole_call_graph": "structure exact (own columns) on instances 0, middle and last",
                         "whole_call_graph_max_norm_err_f64": r["exec_err"],
                         "whole_call_graph_max_norm_err_f32": r32["exec_err"]}
    return line(2.0 * P / r["t_fused"] / 1e9, args.steps, args.warmup, r["t_fused"] * 1e3, "f64", cfg, roof, r["cpu"])


# ----------------------------------------------------------------------------- forward + backward

def run_backward(args):
    """R-MAT 2^18 A*B in FP32: forward (drop-in call) + masked backward."""
    import numpy as np
    import torch
    import opsparse_amd.gen as G
    from opsparse_amd.spgemm import DeviceCSR, Plan, Workspace, backward
    t_gen = time.time()
    A = G.rmat(18, 12, 4, normal=True)
    B = G.rmat(18, 12, 1004, normal=True)
    P = G.compute_flop_host(A, B)
    t_gen = time.time() - t_gen
    log(f"backward: nnz(A) {A.nnz} nnz(B) {B.nnz} products {P}")
    ws = Workspace()
    A32, B32 = A.astype(np.float32), B.astype(np.float32)
    dA = DeviceCSR.from_host(A32, torch.float32)
    dB = DeviceCSR.from_host(B32, torch.float32)
    elapsed, tims, c_last, _ = timed_calls(dA, dB, ws, args.steps, args.warmup, True)
    t_fwd = elapsed / args.steps
    plan = Plan(dA, dB, ws=ws)
    rpt = plan.rowptr64()
    col, _ = plan.numeric(dA, dB)
    for _ in range(args.warmup):
        plan.numeric(dA, dB, col)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.numeric(dA, dB, col)
    torch.cuda.synchronize()
    t_num = (time.perf_counter() - t0) / args.steps
    g = torch.Generator(device="cuda").manual_seed(4)
    dc = torch.randn(plan.nnz, dtype=torch.float32, device="cuda", generator=g)
    for _ in range(args.warmup):
        backward(dA, dB, rpt, col, dc, ws=ws)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ga, gb = backward(dA, dB, rpt, col, dc, ws=ws)
    torch.cuda.synchronize()
    t_bwd = (time.perf_counter() - t0) / args.steps
    log(f"backward: forward call {t_fwd * 1e3:.3f} ms, plan numeric {t_num * 1e3:.3f} ms, backward "
        f"{t_bwd * 1e3:.3f} ms")
    cfg = {"workload": "C=A*B, A and B independent R-MAT 2^18 (.57,.19,.19,.05), 12*2^18 edges merged, N(0,1) "
                       "values in FP32, seeds 4 and 1004; dC ~ N(0,1) on pattern(C); forward = drop-in call, "
                       "backward = masked dA on pattern(A), dB on pattern(B)",
           "rows": A.M, "nnz_a": A.nnz, "nnz_b": B.nnz, "products": int(P), "nnz_c": int(plan.nnz),
           "gen_s": round(t_gen, 1), "forward_call_ms": round(t_fwd * 1e3, 3),
           "plan_numeric_ms": round(t_num * 1e3, 3), "backward_ms": round(t_bwd * 1e3, 3),
           "fwd_plus_bwd_over_fwd": round((t_fwd + t_bwd) / t_fwd, 3),
           "numeric_plus_bwd_over_numeric": round((t_num + t_bwd) / t_num, 3), "phases": phase_ms(tims),
           "parallelism": "single GPU"}
    if not args.no_vendor:
        torch.cuda.synchronize()
        cfg["vendor_rocsparse_forward"] = vendor_leg(dA, dB, True, 3, t_fwd * 1e3, plan.nnz)
        log(f"backward: rocSPARSE forward {cfg['vendor_rocsparse_forward']}")
    if not args.no_check:
        crpt_h, ccol_h = rpt.cpu().numpy(), col.cpu().numpy()
        dA64, dB64 = DeviceCSR.from_host(A, torch.float64), DeviceCSR.from_host(B, torch.float64)
        dc64 = dc.double()
        ga64, gb64 = backward(dA64, dB64, rpt, col, dc64, ws=ws)
        torch.cuda.synchronize()
        t0 = time.time()
        from opsparse_amd.gradcheck import gradient_checks
        chk = gradient_checks(A, B, dA64, dB64, crpt_h, ccol_h, dc64.cpu().numpy(), ga, gb, ga64, gb64,
                              args.fd_entries, ws)
        cfg["gradient_check"] = {
            "entries": args.fd_entries, "normalisation": "sum of |terms| of the entry's gradient",
            "fp64_vs_central_difference_max_err": chk["fd64"], "fp64_vs_fp64_direct_sum_max_err": chk["ref64"],
            "fp32_vs_fp64_direct_sum_max_err": chk["ref32"], "structurally_zero_entries": chk["zero_entries"],
            "zero_entries_max_abs_fd": chk["zero_max_abs_fd"], "zero_entries_max_abs_grad": chk["zero_max_abs_grad"],
            "seconds": round(time.time() - t0, 1)}
        log(f"backward: gradient checks {cfg['gradient_check']}")
    roof = roofline_of(tims, 4, "backward")
    cpu = None
    if not args.no_cpu_baseline:
        cpu, err = cpu_leg(A32, B32, 0, 4096, max(1, args.cpu_iters // 3),
                           "forward, rows 0..4095 of A times B (the highest-degree rows)", c_last, 1e-5,
                           not args.no_check)
        if err is not None:
            cfg["parity"] = {"structure": "exact on rows 0..4095", "max_norm_err": err["max_norm"]}
            import oracle
            # FP64 CPU reference gradient dA on rows 0..4095 (those rows of dA depend on those rows of C only)
            s0, s1 = int(crpt_h[0]), int(crpt_h[4096])
            Ar = A.row_slice(0, 4096)
            ra, _ = oracle.backward(Ar, B, crpt_h[:4097] - s0, ccol_h[s0:s1], dc64[s0:s1].cpu().numpy())
            Aabs = G.HostCSR(Ar.M, Ar.N, Ar.rpt, Ar.col, np.abs(Ar.val))
            Babs = G.HostCSR(B.M, B.N, B.rpt, B.col, np.abs(B.val))
            sa, _ = oracle.backward(Aabs, Babs, crpt_h[:4097] - s0, ccol_h[s0:s1], np.abs(dc64[s0:s1].cpu().numpy()))
            a1 = int(A.rpt[4096])
            sc = np.maximum(sa, 1e-300)
            cfg["gradient_check"]["oracle_dA_rows_0_4095"] = {
                "fp64_max_err": float(np.max(np.abs(ga64[:a1].cpu().numpy() - ra) / sc, initial=0.0)),
                "fp32_max_err": float(np.max(np.abs(ga[:a1].double().cpu().numpy() - ra) / sc, initial=0.0))}
    c_last.free()
    return line(2.0 * P / t_fwd / 1e9, args.steps, args.warmup, t_fwd * 1e3, "f32", cfg, roof, cpu)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sharding a product "
              f"over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload in ("stencil", "uniform", "rmat20"):
        gflops, ms, cfg, roof, cpu = run_single(args, args.workload)
        if args.workload == "rmat20" and not args.no_stencil:
            # the 27-point stencil (1M rows, LDS-hash bins only), as a nested record
            from opsparse_amd.spgemm import mem_cache_release
            mem_cache_release()
            cfg["stencil_100"] = nested_record(*run_single(args, "stencil"))
            # the 4096-instance batch (FP64): fused, packing included, whole call from one hipGraph, looped
            mem_cache_release()
            cfg["batch_4096"] = run_batch(args, nested=True)
        out = line(gflops, args.steps, args.warmup, ms, "f64", cfg, roof, cpu)
    elif args.workload == "batch":
        out = run_batch(args)
    else:
        out = run_backward(args)
    print(json.dumps(out), flush=True)
    if args.copy_calibration:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
