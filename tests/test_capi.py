"""C ABI checks that need no GPU: exports, host pre-checks, host flop count."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import oracle
import opsparse_amd
import opsparse_amd.gen as G
from opsparse_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^[A-Za-z_][\w \*]*?\b(ops_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25, names
    h = C.CDLL(opsparse_amd.LIB_PATH)
    for n in names:
        assert hasattr(h, n), f"{n} declared in include/ but not exported"
    assert set(names) == set(_lib.SIGNATURES), "ctypes table out of sync with the header"


def test_cpp_dropin_symbols_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", opsparse_amd.LIB_PATH], capture_output=True, text=True).stdout
    for sym in ("opsparse::opsparse_spgemm(", "opsparse::opsparse_spgemm_reg(", "opsparse::opsparse_spgemm_safe(",
                "opsparse::CSR::H2D()", "opsparse::CSR::D2H()", "opsparse::CSR::construct(",
                "opsparse::compute_flop(", "opsparse::SpgemmTimings::print(", "opsparse::to_cstr("):
        assert sym in out, sym


def test_matrix_market_banner_interface(tmp_path):
    """The header-only banner parser keeps the reference's interface
    (include/opsparse/io/matrix_market.hpp:11-65): fields matrix / storage /
    type / symmetry, tokenize(tokens, str, delimiters) and a template
    read_mm_banner over any stream.  A reference-style caller compiles
    against it unchanged and sees the same fields and errors."""
    src = tmp_path / "banner.cpp"
    src.write_text(r"""
#include <cstdio>
#include <sstream>
#include "opsparse/io/matrix_market.hpp"
using namespace opsparse::io;
int main() {
    std::vector<std::string> t;
    tokenize(t, "a,,b;c", ",;");
    if (t.size() != 3 || t[0] != "a" || t[1] != "b" || t[2] != "c") return 1;
    tokenize(t, "  x \t y\n");
    if (t.size() != 2 || t[1] != "y") return 2;
    std::istringstream in("%%MatrixMarket matrix coordinate real Symmetric\n3 3 1\n");
    MatrixMarketBanner b;
    read_mm_banner(in, b);
    if (b.matrix != "matrix" || b.storage != "coordinate" || b.type != "real" || b.symmetry != "symmetric") return 3;
    const char *bad[] = {"%%MatrixMarket vector coordinate real general", "%%MatrixMarket matrix array real general",
                         "%%MatrixMarket matrix coordinate real hermitian", "%%MatrixMarket matrix coordinate bad general",
                         "%MatrixMarket matrix coordinate real general", "%%MatrixMarket matrix coordinate real"};
    for (const char *s : bad) {
        std::istringstream bi(s);
        try { read_mm_banner(bi, b); return 4; } catch (const std::runtime_error &) {}
    }
    std::puts("ok");
    return 0;
}
""")
    exe = tmp_path / "banner"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout)


def test_version():
    assert b"gfx950" in _lib.lib().ops_version()


def _desc(M, N, nnz):
    d = _lib.OpsCsr()
    d.M, d.N, d.nnz = M, N, nnz
    return d


def test_safe_precheck_nonsquare_without_gpu():
    a, b, c = _desc(4, 5, 0), _desc(6, 4, 0), _lib.OpsCsr()
    rpt = np.zeros(8, dtype=np.int32)
    st = _lib.lib().ops_spgemm_safe_f64(None, C.byref(a), C.byref(b), rpt.ctypes.data, rpt.ctypes.data,
                                        rpt.ctypes.data, C.byref(c), None, None, None, None, None, None, None)
    assert st == _lib.OPS_NONSQUARE


def test_safe_precheck_int32_overflow_without_gpu():
    # A: 3 rows with one nonzero each pointing at B rows whose lengths sum past INT32_MAX.
    a, b, c = _desc(3, 3, 3), _desc(3, 3, 0), _lib.OpsCsr()
    arpt = np.array([0, 1, 2, 3], dtype=np.int32)
    acol = np.array([0, 1, 2], dtype=np.int32)
    brpt = np.array([0, 2**30, 2**31 - 2, 2**31 - 1], dtype=np.int64).astype(np.int32)
    brpt[2] = 2**31 - 2
    flop = C.c_int64()
    st = _lib.lib().ops_spgemm_safe_f64(None, C.byref(a), C.byref(b), arpt.ctypes.data, acol.ctypes.data,
                                        brpt.ctypes.data, C.byref(c), None, None, C.byref(flop), None, None, None,
                                        None)
    # row lengths 2^30, 2^31-2-2^30, 1 -> total 2^31 - 1: the pre-check passes, and with no workspace
    # the call then stops at the missing argument
    assert st == _lib.OPS_INVALID and flop.value == 2**31 - 1
    acol2 = np.array([0, 0, 0], dtype=np.int32)  # three times the 2^30 row
    st = _lib.lib().ops_spgemm_safe_f64(None, C.byref(a), C.byref(b), arpt.ctypes.data, acol2.ctypes.data,
                                        brpt.ctypes.data, C.byref(c), None, None, C.byref(flop), None, None, None,
                                        None)
    assert st == _lib.OPS_INT32_OVERFLOW and flop.value == 3 * 2**30


def test_invalid_arguments_without_gpu():
    a, b, c = _desc(4, 4, 0), _desc(4, 4, 0), _lib.OpsCsr()
    assert _lib.lib().ops_spgemm_f64(None, C.byref(a), C.byref(b), C.byref(c), None, None) == _lib.OPS_INVALID
    assert _lib.lib().ops_plan_create(None, C.byref(a), C.byref(b), None, None, None) == _lib.OPS_INVALID
    assert _lib.lib().ops_workspace_create(None) == _lib.OPS_INVALID
    assert _lib.lib().ops_csr_free(None) == _lib.OPS_INVALID
    ex = C.c_void_p()
    L = _lib.lib()
    # executors: no workspace, no operands, a negative instance count, no handle
    assert L.ops_exec_create(None, 0, C.byref(a), C.byref(b), 0, 1, C.byref(ex)) == _lib.OPS_INVALID
    assert L.ops_exec_create_batch(None, 0, 1, C.byref(a), C.byref(b), 0, 1, None, C.byref(ex)) == _lib.OPS_INVALID
    assert L.ops_exec_create_batch(None, 0, -1, None, None, 0, 1, None, C.byref(ex)) == _lib.OPS_INVALID
    assert L.ops_exec_launch(None, None) == _lib.OPS_INVALID
    assert L.ops_exec_result(None, None, None, None) == _lib.OPS_INVALID
    assert L.ops_exec_destroy(None) == _lib.OPS_INVALID
    assert not ex.value


def test_host_flop_matches_oracle():
    A = G.rmat(11, 8, 1)
    assert G.compute_flop_host(A, A) == oracle.compute_flop(A, A)


def test_tuning_defaults_within_caps_without_gpu():
    d, c = _lib.OpsTuning(), _lib.OpsTuning()
    assert _lib.lib().ops_tuning_default(C.byref(d)) == _lib.OPS_OK
    assert _lib.lib().ops_tuning_caps(C.byref(c)) == _lib.OPS_OK
    for i in range(_lib.NUM_BINS - 1):
        assert 0 < d.sym_bound[i] <= c.sym_bound[i] and 0 < d.num_bound[i] <= c.num_bound[i]
        if i:
            assert d.sym_bound[i] >= d.sym_bound[i - 1] and d.num_bound[i] >= d.num_bound[i - 1]
    assert d.sym_bitmap == 0 and c.sym_bitmap == 1
    assert _lib.lib().ops_workspace_set_tuning(None, C.byref(d)) == _lib.OPS_INVALID


def test_mem_cache_counters_without_gpu():
    st = _lib.OpsMemStats()
    assert _lib.lib().ops_mem_cache_stats(C.byref(st)) == _lib.OPS_OK
    assert st.cached_bytes == 0 and st.live_bytes == 0 and st.limit == -1
    assert _lib.lib().ops_mem_cache_stats(None) == _lib.OPS_INVALID
    assert _lib.lib().ops_free(None) == _lib.OPS_OK  # like hipFree(NULL)


def test_vendor_comparator_is_a_separate_library():
    """The rocSPARSE comparator is its own library with one entry point, and
    the SpGEMM library does not link rocSPARSE (it is never a dispatch path)."""
    import os
    import subprocess
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    vend = os.path.join(here, "opsparse_amd", "libvendor_rocsparse.so")
    main = os.path.join(here, "opsparse_amd", "libopsparse_amd.so")
    assert os.path.exists(vend) and os.path.exists(main)
    syms = subprocess.run(["nm", "-D", "--defined-only", vend], capture_output=True, text=True, check=True).stdout
    assert "vendor_rocsparse_spgemm" in syms
    deps = subprocess.run(["readelf", "-d", main], capture_output=True, text=True, check=True).stdout
    assert "rocsparse" not in deps
