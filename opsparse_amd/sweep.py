"""Sweep of the safe SpGEMM over many MatrixMarket files.

Runs ``opsparse_amd/bin/opsparse_reg`` (the sweep driver) once per matrix
under a per-matrix time limit, classifies each run from its tagged STATUS /
PERF / META / TIMING_CSV lines and writes two CSV files, as the reference's
sweep does (scripts/run_suitesparse_debug.sh:84-160):

  summary.csv   idx,name,status,rc,wall_s,peak_mb,c_nnz
  timings.csv   name,total_ms,gflops          (runs with status ok)

Status is the driver's STATUS value (ok, nonsquare, int32_overflow, oom,
cuda_error, exception), or ``timeout`` (time limit), ``error`` (nonzero
exit), ``no_status`` (exit 0 without a STATUS line), ``unknown`` (a STATUS
value outside that list) or ``missing`` (no such file).

usage: python -m opsparse_amd.sweep LIST_OR_FILES... [--root DIR] [--out DIR] [--timeout S]
A list argument is a text file of matrix names (one per line, resolved as
ROOT/name/name.mtx); any argument ending in .mtx is taken as a file.
"""
from __future__ import annotations

import argparse
import collections
import os
import subprocess
import sys
import time

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "opsparse_reg")
KNOWN = ("ok", "nonsquare", "int32_overflow", "oom", "cuda_error", "exception")


def parse_tagged(text: str) -> dict:
    """First STATUS / PERF / META / TIMING_CSV / ERROR line of a driver log,
    split on commas (without the tag)."""
    out = {}
    for ln in text.splitlines():
        tag, _, rest = ln.partition(",")
        if tag in ("STATUS", "PERF", "META", "TIMING_CSV", "ERROR") and tag not in out:
            out[tag] = rest.split(",") if tag != "ERROR" else [rest]
    return out


def classify(rc: int | None, tagged: dict) -> str:
    """Outcome of one run: rc None means it hit the time limit."""
    if rc is None:
        return "timeout"
    if rc != 0:
        return "error"
    if "STATUS" not in tagged:
        return "no_status"
    st = tagged["STATUS"][0]
    return st if st in KNOWN else "unknown"


def run_one(path: str, timeout: float, driver: str = DRIVER) -> dict:
    """Runs the driver on one file; returns the summary record."""
    name = os.path.basename(path)
    name = name[:-4] if name.endswith(".mtx") else name
    if not os.path.isfile(path):
        return {"name": name, "status": "missing", "rc": "-", "wall_s": 0.0, "peak_mb": 0.0, "c_nnz": -1,
                "timing": None, "log": ""}
    t0 = time.time()
    try:
        p = subprocess.run([driver, path], capture_output=True, text=True, timeout=timeout)
        rc, log = p.returncode, p.stdout + p.stderr
    except subprocess.TimeoutExpired as e:
        rc = None
        log = (e.stdout or b"").decode(errors="replace") if isinstance(e.stdout, bytes) else (e.stdout or "")
    wall = time.time() - t0
    tagged = parse_tagged(log)
    status = classify(rc, tagged)
    peak = int(tagged["PERF"][0]) if "PERF" in tagged else 0
    c_nnz = int(tagged["META"][3]) if "META" in tagged and len(tagged["META"]) > 3 else -1
    timing = tagged.get("TIMING_CSV") if status == "ok" else None
    return {"name": name, "status": status, "rc": "timeout" if rc is None else rc, "wall_s": round(wall, 2),
            "peak_mb": round(peak / (1 << 20), 1), "c_nnz": c_nnz, "timing": timing, "log": log}


def expand(args, root: str) -> list:
    paths = []
    for a in args:
        if a.endswith(".mtx"):
            paths.append(a)
            continue
        with open(a) as f:
            for ln in f:
                n = ln.strip()
                if n:
                    paths.append(os.path.join(root, n, n + ".mtx"))
    return paths


def sweep(paths, out_dir: str, timeout: float = 120.0, driver: str = DRIVER, echo=print) -> collections.Counter:
    """Runs every path, writes summary.csv / timings.csv (and per-matrix
    logs under out_dir/logs); returns the count of each status."""
    os.makedirs(os.path.join(out_dir, "logs"), exist_ok=True)
    counts = collections.Counter()
    with open(os.path.join(out_dir, "summary.csv"), "w") as fs, open(os.path.join(out_dir, "timings.csv"), "w") as ft:
        fs.write("idx,name,status,rc,wall_s,peak_mb,c_nnz\n")
        ft.write("name,total_ms,gflops\n")
        for i, path in enumerate(paths, 1):
            r = run_one(path, timeout, driver)
            counts[r["status"]] += 1
            with open(os.path.join(out_dir, "logs", r["name"] + ".log"), "w") as fl:
                fl.write(r["log"])
            fs.write(f"{i},{r['name']},{r['status']},{r['rc']},{r['wall_s']},{r['peak_mb']},{r['c_nnz']}\n")
            if r["timing"]:
                ft.write(",".join(r["timing"]) + "\n")
            fs.flush()
            ft.flush()
            echo(f"[{i}/{len(paths)}] {r['status']} rc={r['rc']} {r['wall_s']}s peak={r['peak_mb']}MB {r['name']}")
    return counts


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("inputs", nargs="+")
    p.add_argument("--root", default="../matrix/suite_sparse")
    p.add_argument("--out", default="sweep_out")
    p.add_argument("--timeout", type=float, default=120.0)
    a = p.parse_args(argv)
    if not os.path.exists(DRIVER):
        print(f"{DRIVER} is missing: build with make -C opsparse_amd/csrc", file=sys.stderr)
        return 2
    counts = sweep(expand(a.inputs, a.root), a.out, a.timeout)
    print(f"\nDone: total={sum(counts.values())}")
    for k in KNOWN + ("timeout", "error", "no_status", "unknown", "missing"):
        print(f"  {k}={counts.get(k, 0)}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
