"""Benchmark drivers and the sweep, used the way the reference's are.

* opsparse_amd/bin/opsparse_reg (benchmarks/opsparse_reg.cu:46-117): tagged
  STATUS / PERF / META / TIMING_CSV lines, exit code 0 for every classified
  outcome;
* opsparse_amd/bin/opsparse_profile (benchmarks/opsparse_profile.cu:10-48);
* opsparse_amd/sweep.py (scripts/run_suitesparse_debug.sh:84-160);
* bench::load_pair (benchmarks/bench_common.hpp:17-34) and the CSR
  sub-matrix constructor (src/opsparse/core_ds/csr.cu:110-147), host only.

CPU tests cover every path that makes no HIP call (argument checks,
nonsquare and unreadable inputs, the sweep's parsing and classification,
load_pair and the sub-matrix constructor against scipy); GPU tests run the
drivers on square inputs (ok, checked against the oracle) and on an input
whose products exceed INT32_MAX (int32_overflow from the host pre-check).
"""
import os
import subprocess

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import oracle
from opsparse_amd import sweep
from opsparse_amd._lib import LIB_PATH
from opsparse_amd.gen import HostCSR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "opsparse_amd", "bin")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
REG = os.path.join(BIN, "opsparse_reg")
PROFILE = os.path.join(BIN, "opsparse_profile")

needs_drivers = pytest.mark.skipif(not (os.path.exists(REG) and os.path.exists(PROFILE)),
                                   reason="drivers not built (make -C opsparse_amd/csrc)")


def write_mtx(path, S, symmetry="general", field="real"):
    scipy.io.mmwrite(path, S, symmetry=symmetry, field=field)
    return path


def random_sparse(m, n, per_row, seed):
    rng = np.random.default_rng(seed)
    return sp.random(m, n, density=min(1.0, per_row / n), format="csr", random_state=rng,
                     data_rvs=lambda k: rng.uniform(-1, 1, k))


def tagged(stdout):
    return sweep.parse_tagged(stdout)


@needs_drivers
def test_reg_usage_and_host_only_outcomes(tmp_path):
    r = subprocess.run([REG], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    # nonsquare: classified before any device work, META carries the shape
    ns = write_mtx(str(tmp_path / "rect.mtx"), random_sparse(30, 40, 3, 1))
    r = subprocess.run([REG, ns], capture_output=True, text=True, timeout=60)
    t = tagged(r.stdout)
    assert r.returncode == 0 and t["STATUS"] == ["nonsquare"], r.stdout
    S = scipy.io.mmread(ns)
    assert t["META"] == ["30", "40", str(S.nnz), "-1", "0"]
    # unreadable input: exception with the loader's message
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n")
    r = subprocess.run([REG, str(bad)], capture_output=True, text=True, timeout=60)
    t = tagged(r.stdout)
    assert r.returncode == 0 and t["STATUS"] == ["exception"] and t["ERROR"][0].startswith("load:"), r.stdout


def test_sweep_parse_and_classify():
    log = "noise\nSTATUS,ok\nPERF,1048576,10,9,100\nMETA,4,4,7,12,30\nTIMING_CSV,m,0.500,0.120\nSTATUS,oom\n"
    t = sweep.parse_tagged(log)
    assert t["STATUS"] == ["ok"] and t["META"][3] == "12" and t["TIMING_CSV"] == ["m", "0.500", "0.120"]
    assert sweep.classify(0, t) == "ok"
    assert sweep.classify(None, t) == "timeout"
    assert sweep.classify(139, t) == "error"
    assert sweep.classify(0, {}) == "no_status"
    assert sweep.classify(0, {"STATUS": ["weird"]}) == "unknown"
    for s in sweep.KNOWN:
        assert sweep.classify(0, {"STATUS": [s]}) == s


@needs_drivers
def test_sweep_host_only_outcomes(tmp_path):
    ns = write_mtx(str(tmp_path / "rect.mtx"), random_sparse(20, 25, 3, 2))
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n")
    lst = tmp_path / "list.txt"
    lst.write_text("absent_matrix\n")
    paths = [ns, str(bad)] + sweep.expand([str(lst)], str(tmp_path))
    out = tmp_path / "out"
    counts = sweep.sweep(paths, str(out), timeout=60, echo=lambda s: None)
    assert counts == {"nonsquare": 1, "exception": 1, "missing": 1}
    rows = (out / "summary.csv").read_text().splitlines()
    assert rows[0] == "idx,name,status,rc,wall_s,peak_mb,c_nnz" and len(rows) == 4
    assert rows[1].split(",")[1:3] == ["rect", "nonsquare"] and rows[3].split(",")[2] == "missing"
    assert (out / "timings.csv").read_text().splitlines() == ["name,total_ms,gflops"]


def build_pair_driver(out_dir):
    exe = os.path.join(out_dir, "pair_driver")
    libdir = os.path.dirname(LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{ROOT}/include",
           f"-I{ROOT}/opsparse_amd/csrc/drivers", os.path.join(ROOT, "tests", "cpp", "pair_driver.cpp"), "-o", exe,
           f"-L{libdir}", "-lopsparse_amd", f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{libdir}",
           f"-Wl,-rpath,{ROCM}/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def read_csrs(path, count):
    out = []
    with open(path, "rb") as f:
        for _ in range(count):
            M, N, nnz = (int(x) for x in np.frombuffer(f.read(12), dtype=np.int32))
            rpt = np.frombuffer(f.read(4 * (M + 1)), dtype=np.int32)
            col = np.frombuffer(f.read(4 * nnz), dtype=np.int32)
            val = np.frombuffer(f.read(8 * nnz), dtype=np.float64)
            out.append(sp.csr_matrix((val, col, rpt), shape=(M, N)))
    return out


def same(X, Y):
    X, Y = sp.csr_matrix(X), sp.csr_matrix(Y)
    X.sort_indices()
    Y.sort_indices()
    return (X.shape == Y.shape and np.array_equal(X.indptr, Y.indptr) and np.array_equal(X.indices, Y.indices)
            and np.array_equal(X.data, Y.data))


@pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="library not built")
def test_load_pair_and_submatrix_against_scipy(tmp_path):
    exe = build_pair_driver(str(tmp_path))
    A = random_sparse(60, 50, 6, 3)
    B = random_sparse(70, 45, 5, 4)
    C = random_sparse(40, 30, 4, 5)
    pa, pb, pc = (write_mtx(str(tmp_path / f"{n}.mtx"), X) for n, X in (("a", A), ("b", B), ("c", C)))
    As, Bs, Cs = (scipy.io.mmread(p).tocsr() for p in (pa, pb, pc))
    out = str(tmp_path / "o.bin")
    # A.N (50) < B.M (70): B keeps its first 50 rows
    subprocess.run([exe, "pair", pa, pb, out], check=True)
    ra, rb = read_csrs(out, 2)
    assert same(ra, As) and same(rb, Bs[:50, :])
    # A.N (50) > C.M (40): A keeps its first 40 columns
    subprocess.run([exe, "pair", pa, pc, out], check=True)
    ra, rc = read_csrs(out, 2)
    assert same(ra, As[:, :40]) and same(rc, Cs)
    # same name: B is a copy of A
    subprocess.run([exe, "pair", pa, pa, out], check=True)
    ra, rb = read_csrs(out, 2)
    assert same(ra, As) and same(rb, As)
    # sub-matrix with offsets, an empty one and the whole matrix
    for M_, N_, m0, n0 in ((25, 17, 10, 20), (0, 0, 60, 50), (60, 50, 0, 0), (1, 50, 59, 0)):
        subprocess.run([exe, "sub", pa, str(M_), str(N_), str(m0), str(n0), out], check=True)
        (rs,) = read_csrs(out, 1)
        assert same(rs, As[m0:m0 + M_, n0:n0 + N_]), (M_, N_, m0, n0)
    r = subprocess.run([exe, "sub", pa, "30", "10", "40", "0", out], capture_output=True, text=True)
    assert r.returncode == 3 and "out_of_range" in r.stdout


@pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="library not built")
@pytest.mark.parametrize("symmetry,field", [("general", "real"), ("general", "pattern"), ("symmetric", "real"),
                                            ("symmetric", "pattern"), ("skew-symmetric", "real"),
                                            ("general", "integer")])
def test_mtx_loader_formats_against_scipy(tmp_path, symmetry, field):
    """CSR(path) (the reference's MatrixMarket loader, csr.cu:187-352) on each
    banner the reference accepts: symmetric and skew-symmetric entries are
    mirrored (negated for skew), pattern entries read as 1.0, integer values
    as doubles; the result equals scipy's reading of the same file."""
    exe = build_pair_driver(str(tmp_path))
    rng = np.random.default_rng(8)
    n = 40
    S = random_sparse(n, n, 5, 9)
    if field == "integer":
        S.data = rng.integers(-9, 10, S.data.size).astype(np.float64)
    if symmetry == "symmetric":
        L = sp.tril(S, format="csr")
        S = L + sp.tril(L, k=-1).T
    elif symmetry == "skew-symmetric":
        L = sp.tril(S, k=-1, format="csr")
        S = L - L.T
    if field == "pattern":
        S = S.copy()
        S.data[:] = 1.0
    path = write_mtx(str(tmp_path / "m.mtx"), S, symmetry=symmetry, field=field)
    head = open(path).readline().split()
    assert head[3:5] == [field, symmetry], head
    ref = scipy.io.mmread(path).tocsr()
    out = str(tmp_path / "o.bin")
    subprocess.run([exe, "sub", path, str(n), str(n), "0", "0", out], check=True)
    (got,) = read_csrs(out, 1)
    assert same(got, ref)


def hub_matrix(n):
    """Row 0 and column 0 full: A*A has n*n products through entry (0,0)'s
    column and row, more than INT32_MAX for n >= 46341."""
    rows = np.r_[np.zeros(n, dtype=np.int64), np.arange(1, n)]
    cols = np.r_[np.arange(n), np.zeros(n - 1, dtype=np.int64)]
    return sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(n, n))


@pytest.mark.gpu
@needs_drivers
def test_reg_ok_overflow_and_sweep_on_gpu(gpu, tmp_path):
    S = random_sparse(2000, 2000, 7, 6)
    ok = write_mtx(str(tmp_path / "sq.mtx"), S)
    Sm = scipy.io.mmread(ok).tocsr()
    Sm.sort_indices()
    Ah = HostCSR(2000, 2000, Sm.indptr.astype(np.int32), Sm.indices.astype(np.int32), Sm.data.astype(np.float64))
    ref = oracle.spgemm(Ah, Ah)
    P = oracle.compute_flop(Ah, Ah)
    r = subprocess.run([REG, ok], capture_output=True, text=True, timeout=120)
    t = tagged(r.stdout)
    assert r.returncode == 0 and t["STATUS"] == ["ok"], r.stdout + r.stderr
    assert t["META"] == ["2000", "2000", str(Sm.nnz), str(ref[1].size), str(P)]
    assert t["TIMING_CSV"][0] == "sq" and float(t["TIMING_CSV"][1]) > 0
    assert len(t["PERF"]) == 4 and int(t["PERF"][3]) > 0
    big = write_mtx(str(tmp_path / "hub.mtx"), hub_matrix(50000), field="pattern")
    r = subprocess.run([REG, big], capture_output=True, text=True, timeout=120)
    t = tagged(r.stdout)
    assert r.returncode == 0 and t["STATUS"] == ["int32_overflow"], r.stdout + r.stderr
    assert t["META"][3] == "-1" and int(t["META"][4]) > 2**31 - 1 and "TIMING_CSV" not in t
    ns = write_mtx(str(tmp_path / "rect.mtx"), random_sparse(30, 40, 3, 7))
    counts = sweep.sweep([ok, big, ns], str(tmp_path / "out"), timeout=120, echo=lambda s: None)
    assert counts == {"ok": 1, "int32_overflow": 1, "nonsquare": 1}
    timing_rows = (tmp_path / "out" / "timings.csv").read_text().splitlines()
    assert len(timing_rows) == 2 and timing_rows[1].startswith("sq,")


@pytest.mark.gpu
@needs_drivers
def test_profile_driver_pair(gpu, tmp_path):
    A = random_sparse(3000, 2500, 6, 8)
    B = random_sparse(2600, 2000, 5, 9)
    pa = write_mtx(str(tmp_path / "a.mtx"), A)
    pb = write_mtx(str(tmp_path / "b.mtx"), B)
    r = subprocess.run([PROFILE, pa, pb], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    As = scipy.io.mmread(pa).tocsr()
    Bs = scipy.io.mmread(pb).tocsr()[:2500, :]
    P = int(np.diff(Bs.indptr)[As.indices].sum())
    assert r.stdout.startswith(pa + " ")
    assert f"total flop {2.0 * P:f}" in r.stdout and "perf(Gflops):" in r.stdout
