// C++ drop-in API (namespace opsparse) over the C ABI.
//
// Mirrors the reference's public C++ surface: CSR (core_ds/csr.cu),
// SpgemmMeta, opsparse_spgemm / _reg / _safe (core_algo/spgemm/spgemm.cu:16-219),
// SpgemmTimings (profile/spgemm_timings.cpp), SpgemmStatus / SpgemmPerf
// (profile/perf.cpp) and the MatrixMarket loader (csr.cu:187-352,
// io/matrix_market.hpp:18-65).  Same argument meaning and error behaviour:
// opsparse_spgemm and _reg throw after printing on a device error; _safe
// never throws.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "internal.hpp"
#include "opsparse/core_algo/spgemm/spgemm.hpp"
#include "outmem.hpp"
#include "opsparse/io/matrix_market.hpp"
#include "opsparse/numeric/compute_flop.hpp"
#include "opsparse/utils/tsc_timer.hpp"
#include "opsparse_amd.h"

namespace opsparse {

// ---------------- CSR ----------------

void CSR::hrelease()
{
    delete[] rpt;
    delete[] col;
    delete[] val;
    rpt = nullptr;
    col = nullptr;
    val = nullptr;
}

void CSR::drelease()
{
    // ops_free: arrays of a library-computed C go back to the output cache,
    // arrays from H2D (plain hipMalloc) to hipFree.
    if (d_rpt) OPSPARSE_CHECK_HIP(opsamd::out_free(d_rpt));
    if (d_col) OPSPARSE_CHECK_HIP(opsamd::out_free(d_col));
    if (d_val) OPSPARSE_CHECK_HIP(opsamd::out_free(d_val));
    d_rpt = nullptr;
    d_col = nullptr;
    d_val = nullptr;
}

void CSR::release()
{
    hrelease();
    drelease();
}

CSR::~CSR()
{
    hrelease();
    // Destructors must not throw; ignore device errors here.
    if (d_rpt) (void)opsamd::out_free(d_rpt);
    if (d_col) (void)opsamd::out_free(d_col);
    if (d_val) (void)opsamd::out_free(d_val);
}

void CSR::H2D()
{
    drelease();
    OPSPARSE_CHECK_HIP(hipMalloc(&d_rpt, (static_cast<size_t>(M) + 1) * sizeof(mint)));
    OPSPARSE_CHECK_HIP(hipMalloc(&d_col, std::max<size_t>(nnz, 1) * sizeof(mint)));
    OPSPARSE_CHECK_HIP(hipMalloc(&d_val, std::max<size_t>(nnz, 1) * sizeof(mdouble)));
    h2d(d_rpt, rpt, (static_cast<size_t>(M) + 1) * sizeof(mint));
    if (nnz > 0) {
        h2d(d_col, col, static_cast<size_t>(nnz) * sizeof(mint));
        h2d(d_val, val, static_cast<size_t>(nnz) * sizeof(mdouble));
    }
}

void CSR::D2H()
{
    hrelease();
    rpt = new mint[static_cast<size_t>(M) + 1];
    col = new mint[std::max<size_t>(nnz, 1)];
    val = new mdouble[std::max<size_t>(nnz, 1)];
    d2h(rpt, d_rpt, (static_cast<size_t>(M) + 1) * sizeof(mint));
    if (nnz > 0) {
        d2h(col, d_col, static_cast<size_t>(nnz) * sizeof(mint));
        d2h(val, d_val, static_cast<size_t>(nnz) * sizeof(mdouble));
    }
}

CSR::CSR(const CSR &A) : M(A.M), N(A.N), nnz(A.nnz)
{
    rpt = new mint[static_cast<size_t>(M) + 1];
    col = new mint[std::max<size_t>(nnz, 1)];
    val = new mdouble[std::max<size_t>(nnz, 1)];
    std::memcpy(rpt, A.rpt, (static_cast<size_t>(M) + 1) * sizeof(mint));
    std::memcpy(col, A.col, static_cast<size_t>(nnz) * sizeof(mint));
    std::memcpy(val, A.val, static_cast<size_t>(nnz) * sizeof(mdouble));
}

CSR &CSR::operator=(const CSR &A)
{
    if (this == &A) return *this;
    release();
    M = A.M;
    N = A.N;
    nnz = A.nnz;
    rpt = new mint[static_cast<size_t>(M) + 1];
    col = new mint[std::max<size_t>(nnz, 1)];
    val = new mdouble[std::max<size_t>(nnz, 1)];
    std::memcpy(rpt, A.rpt, (static_cast<size_t>(M) + 1) * sizeof(mint));
    std::memcpy(col, A.col, static_cast<size_t>(nnz) * sizeof(mint));
    std::memcpy(val, A.val, static_cast<size_t>(nnz) * sizeof(mdouble));
    return *this;
}

CSR::CSR(const CSR &A, mint M_, mint N_, mint M_start, mint N_start)
{
    if (M_ < 0 || N_ < 0 || M_start < 0 || N_start < 0 || M_start + M_ > A.M || N_start + N_ > A.N)
        throw std::out_of_range("CSR sub-matrix out of range");
    M = M_;
    N = N_;
    rpt = new mint[static_cast<size_t>(M) + 1];
    rpt[0] = 0;
    for (mint i = 0; i < M; i++) {
        mint cnt = 0;
        for (mint j = A.rpt[M_start + i]; j < A.rpt[M_start + i + 1]; j++)
            cnt += (A.col[j] >= N_start && A.col[j] < N_start + N_) ? 1 : 0;
        rpt[i + 1] = rpt[i] + cnt;
    }
    nnz = rpt[M];
    col = new mint[std::max<size_t>(nnz, 1)];
    val = new mdouble[std::max<size_t>(nnz, 1)];
    for (mint i = 0; i < M; i++) {
        mint out = rpt[i];
        for (mint j = A.rpt[M_start + i]; j < A.rpt[M_start + i + 1]; j++) {
            if (A.col[j] >= N_start && A.col[j] < N_start + N_) {
                col[out] = A.col[j] - N_start;
                val[out++] = A.val[j];
            }
        }
    }
}

bool CSR::operator==(const CSR &rhs)
{
    if (nnz != rhs.nnz) {
        std::printf("nnz not equal %d %d\n", nnz, rhs.nnz);
        throw std::runtime_error("nnz not equal");
    }
    if (M != rhs.M || N != rhs.N) throw std::runtime_error("dimension not same");
    int errors = 0;
    const double eps = 1e-9;
    for (mint i = 0; i < M; i++) {
        if (errors > 10) throw std::runtime_error("matrix compare: error num exceed threshold");
        if (rpt[i] != rhs.rpt[i]) {
            std::printf("rpt not equal at %d rows, %d != %d\n", i, rpt[i], rhs.rpt[i]);
            errors++;
        }
        for (mint j = rpt[i]; j < rpt[i + 1]; j++) {
            if (errors > 10) throw std::runtime_error("matrix compare: error num exceed threshold");
            if (col[j] != rhs.col[j]) {
                std::printf("col not equal at %d rows, index %d != %d\n", i, col[j], rhs.col[j]);
                errors++;
            }
            const double d = std::fabs(val[j] - rhs.val[j]);
            if (!(d < eps || d < eps * std::fabs(val[j]))) {
                std::printf("val not equal at %d rows, value %.18le != %.18le\n", i, val[j], rhs.val[j]);
                errors++;
            }
        }
    }
    if (rpt[M] != rhs.rpt[M]) throw std::runtime_error("matrix compare: rpt[M] not equal");
    return errors == 0;
}

CSR::CSR(const std::string &mtx_file) { construct(mtx_file); }

// MatrixMarket coordinate -> CSR: 1-based to 0-based, symmetric and
// skew-symmetric expansion, entries sorted by (row, col), then validated
// (sorted, in-range columns).  Duplicate entries are rejected.
void CSR::construct(const std::string &mtx_file)
{
    std::ifstream in(mtx_file.c_str());
    if (!in) throw std::runtime_error("unable to open file \"" + mtx_file + "\" for reading");
    io::MatrixMarketBanner banner;
    io::read_mm_banner(in, banner);
    std::string line;
    do {
        if (!std::getline(in, line)) throw std::runtime_error("MatrixMarket: missing size line");
    } while (line.empty() || line[0] == '%');
    std::vector<std::string> tok;
    io::tokenize(tok, line);
    if (tok.size() != 3) throw std::runtime_error("invalid MatrixMarket coordinate format");
    long long m = 0, n = 0, nz = 0;
    std::istringstream(tok[0]) >> m;
    std::istringstream(tok[1]) >> n;
    std::istringstream(tok[2]) >> nz;
    if (m < 0 || n < 0 || nz < 0 || m > INT_MAX || n > INT_MAX) throw std::runtime_error("MatrixMarket: bad sizes");
    struct E {
        long long key;
        double v;
    };
    std::vector<E> ents;
    ents.reserve(static_cast<size_t>(nz) * (banner.symmetry == "general" ? 1 : 2));
    const bool pattern = banner.type == "pattern";
    const bool cplx = banner.type == "complex";  // the imaginary part is read and dropped
    long long read = 0;
    while (read < nz) {
        long long i, j;
        double v = 1.0, im = 0.0;
        if (!(in >> i >> j)) break;
        if (!pattern && !(in >> v)) break;
        if (cplx && !(in >> im)) break;
        if (i < 1 || i > m || j < 1 || j > n) throw std::runtime_error("csr col_index out of range");
        i -= 1;
        j -= 1;
        ents.push_back({i * n + j, v});
        if (banner.symmetry != "general" && i != j)
            ents.push_back({j * n + i, banner.symmetry == "skew-symmetric" ? -v : v});
        read++;
    }
    if (read != nz) throw std::runtime_error("read nnz not equal to declared nnz " + std::to_string(read));
    std::sort(ents.begin(), ents.end(), [](const E &a, const E &b) { return a.key < b.key; });
    for (size_t t = 1; t < ents.size(); t++)
        if (ents[t].key == ents[t - 1].key) throw std::runtime_error("csr col_index not in assending order");
    if (ents.size() > static_cast<size_t>(INT_MAX)) throw std::runtime_error("nnz exceeds INT32_MAX");
    release();
    M = static_cast<mint>(m);
    N = static_cast<mint>(n);
    nnz = static_cast<mint>(ents.size());
    rpt = new mint[static_cast<size_t>(M) + 1]();
    col = new mint[std::max<size_t>(nnz, 1)];
    val = new mdouble[std::max<size_t>(nnz, 1)];
    for (mint t = 0; t < nnz; t++) {
        const long long r = n > 0 ? ents[t].key / n : 0;
        rpt[r + 1]++;
        col[t] = static_cast<mint>(n > 0 ? ents[t].key % n : 0);
        val[t] = ents[t].v;
    }
    for (mint i = 0; i < M; i++) rpt[i + 1] += rpt[i];
}

// ---------------- SpgemmMeta ----------------

SpgemmMeta::SpgemmMeta(CSR &C) { allocate_rpt(C); }

SpgemmMeta::~SpgemmMeta() { release(); }

void SpgemmMeta::allocate_rpt(CSR &C)
{
    OPSPARSE_CHECK_HIP(hipMalloc(&C.d_rpt, (static_cast<size_t>(C.M) + 1) * sizeof(mint)));
}

// Mirror layout (host and device): bin_size[kNumBin], max_row_nnz,
// total_nnz, bin_offset[kNumBin].
constexpr int kMirror = 2 * kNumBin + 2;

void SpgemmMeta::allocate(CSR &C)
{
    M = C.M;
    N = C.N;
    (void)workspace();
    if (!combined_mem) {
        combined_mem = static_cast<mint *>(std::calloc(kMirror, sizeof(mint)));
        if (!combined_mem) throw std::bad_alloc();
        OPSPARSE_CHECK_HIP(hipMalloc(&d_combined_mem, kMirror * sizeof(mint)));
        OPSPARSE_CHECK_HIP(hipMemset(d_combined_mem, 0, kMirror * sizeof(mint)));
        bin_size = combined_mem;
        max_row_nnz = bin_size + kNumBin;
        total_nnz = bin_size + kNumBin + 1;
        bin_offset = bin_size + kNumBin + 2;
        d_bin_size = d_combined_mem;
        d_max_row_nnz = d_bin_size + kNumBin;
        d_total_nnz = d_bin_size + kNumBin + 1;
        d_bin_offset = d_bin_size + kNumBin + 2;
    }
    if (!stream) {
        for (int i = 0; i < kNumBin; i++) OPSPARSE_CHECK_HIP(hipStreamCreateWithFlags(&pool_[i], hipStreamNonBlocking));
        stream = pool_;
    }
}

ops_workspace *SpgemmMeta::workspace()
{
    if (!ws_) {
        const int st = ops_workspace_create(&ws_);
        if (st != OPS_OK) throw std::runtime_error("ops_workspace_create failed");
    }
    return ws_;
}

void SpgemmMeta::release()
{
    if (ws_) ops_workspace_destroy(ws_);
    ws_ = nullptr;
    if (stream) {
        for (int i = 0; i < kNumBin; i++) (void)hipStreamDestroy(pool_[i]);
        stream = nullptr;
    }
    if (d_combined_mem) (void)hipFree(d_combined_mem);
    std::free(combined_mem);
    d_combined_mem = combined_mem = nullptr;
    d_bin_size = d_bin_offset = d_max_row_nnz = d_total_nnz = d_bins = nullptr;
    bin_size = bin_offset = max_row_nnz = total_nnz = nullptr;
}

hipStream_t SpgemmMeta::at(mint stream_idx) const
{
    return (stream && stream_idx >= 0 && stream_idx < kNumBin) ? stream[stream_idx] : nullptr;
}

void SpgemmMeta::memset_bin_size(mint stream_idx)
{
    OPSPARSE_CHECK_HIP(hipMemsetAsync(d_bin_size, 0, kNumBin * sizeof(mint), at(stream_idx)));
}

void SpgemmMeta::memset_all(mint stream_idx)
{
    OPSPARSE_CHECK_HIP(hipMemsetAsync(d_bin_size, 0, (kNumBin + 2) * sizeof(mint), at(stream_idx)));
}

void SpgemmMeta::d2h_bin_size(mint stream_idx)
{
    OPSPARSE_CHECK_HIP(
        hipMemcpyAsync(bin_size, d_bin_size, kNumBin * sizeof(mint), hipMemcpyDeviceToHost, at(stream_idx)));
}

void SpgemmMeta::d2h_all(mint stream_idx)
{
    OPSPARSE_CHECK_HIP(
        hipMemcpyAsync(bin_size, d_bin_size, (kNumBin + 2) * sizeof(mint), hipMemcpyDeviceToHost, at(stream_idx)));
}

void SpgemmMeta::h2d_bin_offset(mint stream_idx)
{
    OPSPARSE_CHECK_HIP(
        hipMemcpyAsync(d_bin_offset, bin_offset, kNumBin * sizeof(mint), hipMemcpyHostToDevice, at(stream_idx)));
}

void SpgemmMeta::update_from_last_call(const mint *numeric_rows, const mint *symbolic_rows, mint max_nnz,
                                       long long nnz)
{
    mint off = 0;
    for (int b = 0; b < kNumBin; b++) {
        bin_size[b] = numeric_rows[b];
        bin_offset[b] = off;
        off += numeric_rows[b];
        symbolic_bin_size[b] = symbolic_rows[b];
    }
    *max_row_nnz = max_nnz;
    *total_nnz = nnz > INT32_MAX ? -1 : static_cast<mint>(nnz);
    total_nnz64 = nnz;
    d_bins = const_cast<mint *>(opsamd::workspace_bins(ws_));
    OPSPARSE_CHECK_HIP(hipMemcpy(d_combined_mem, combined_mem, kMirror * sizeof(mint), hipMemcpyHostToDevice));
}

// ---------------- entry points ----------------

namespace {

ops_csr view(const CSR &X)
{
    ops_csr d{};
    d.M = X.M;
    d.N = X.N;
    d.nnz = X.nnz;
    d.rpt = X.d_rpt;
    d.col = X.d_col;
    d.val = X.d_val;
    return d;
}

void fill_timing(SpgemmTimings &t, const ops_timings &o)
{
    t.setup = o.setup;
    t.symbolic_binning = o.symbolic_binning;
    t.symbolic = o.symbolic;
    t.reduce = o.reduce;
    t.numeric_binning = o.numeric_binning;
    t.prefix = o.prefix;
    t.allocate = o.allocate;
    t.numeric = o.numeric;
    t.cleanup = o.cleanup;
    t.total = o.total;
}

void fill_meta(SpgemmMeta &meta, const ops_timings &o)
{
    mint num[kNumBin], sym[kNumBin];
    for (int b = 0; b < kNumBin; b++) {
        num[b] = static_cast<mint>(o.num_rows[b]);
        sym[b] = static_cast<mint>(o.sym_rows[b]);
    }
    meta.update_from_last_call(num, sym, o.max_row_nnz, o.nnz_c);
}

void run(const CSR &A, const CSR &B, CSR &C, SpgemmMeta &meta, SpgemmTimings &timing, bool separate)
{
    const double t0 = fast_clock_time();
    meta.allocate(C);
    ops_workspace *ws = meta.workspace();
    meta.M = A.M;
    meta.N = B.N;
    ops_csr a = view(A), b = view(B), c{};
    ops_timings t{};
    t.measure_separate = separate ? 1 : 0;
    const int st = ops_spgemm_f64(ws, &a, &b, &c, nullptr, &t);
    if (st != OPS_OK) {
        std::printf("opsparse_spgemm failed (status %d): %s\n", st, ops_workspace_last_error(ws));
        throw std::runtime_error(ops_workspace_last_error(ws));
    }
    C.drelease();
    C.M = c.M;
    C.N = c.N;
    C.nnz = c.nnz;
    C.d_rpt = c.rpt;
    C.d_col = c.col;
    C.d_val = static_cast<mdouble *>(c.val);
    fill_timing(timing, t);
    fill_meta(meta, t);
    timing.total = fast_clock_time() - t0;
}

}  // namespace

void opsparse_spgemm(const CSR &A, const CSR &B, CSR &C, SpgemmMeta &meta, SpgemmTimings &timing)
{
    run(A, B, C, meta, timing, true);
}

void opsparse_spgemm_reg(const CSR &A, const CSR &B, CSR &C, SpgemmMeta &meta, SpgemmTimings &timing)
{
    run(A, B, C, meta, timing, false);
}

SpgemmStatus opsparse_spgemm_safe(const CSR &A, const CSR &B, CSR &C, SpgemmMeta &meta, SpgemmTimings &timing,
                                  SpgemmPerf &perf)
{
    perf.reset();
    if (A.N != B.M) {
        perf.status = SpgemmStatus::nonsquare;
        perf.error_msg = "A.N != B.M";
        return perf.status;
    }
    ops_workspace *ws = nullptr;
    try {
        meta.allocate(C);
        ws = meta.workspace();
    } catch (const std::exception &e) {
        perf.status = SpgemmStatus::cuda_error;
        perf.error_msg = e.what();
        return perf.status;
    }
    ops_csr a = view(A), b = view(B), c{};
    ops_timings t{};
    int64_t flop = 0, cnnz = -1;
    const double t0 = fast_clock_time();
    const int st = ops_spgemm_safe_f64(ws, &a, &b, A.rpt, A.col, B.rpt, &c, nullptr, &t, &flop, &cnnz,
                                       &perf.free_before_bytes, &perf.free_after_bytes, &perf.device_total_bytes);
    perf.total_flop = flop;
    perf.c_nnz = cnnz;
    switch (st) {
        case OPS_OK: perf.status = SpgemmStatus::ok; break;
        case OPS_NONSQUARE: perf.status = SpgemmStatus::nonsquare; break;
        case OPS_INT32_OVERFLOW: perf.status = SpgemmStatus::int32_overflow; break;
        case OPS_OOM: perf.status = SpgemmStatus::oom; break;
        case OPS_HIP_ERROR: perf.status = SpgemmStatus::cuda_error; break;
        default: perf.status = SpgemmStatus::exception; break;
    }
    if (st != OPS_OK) {
        if (st == OPS_INT32_OVERFLOW)
            perf.error_msg = flop > INT_MAX ? "total_flop exceeds INT_MAX" : "nnz(C) exceeds INT_MAX";
        else if (st != OPS_NONSQUARE)
            perf.error_msg = ops_workspace_last_error(ws);
        return perf.status;
    }
    C.drelease();
    C.M = c.M;
    C.N = c.N;
    C.nnz = c.nnz;
    C.d_rpt = c.rpt;
    C.d_col = c.col;
    C.d_val = static_cast<mdouble *>(c.val);
    fill_timing(timing, t);
    fill_meta(meta, t);
    timing.total = fast_clock_time() - t0;
    if (perf.free_before_bytes >= perf.free_after_bytes)
        perf.peak_used_bytes = perf.free_before_bytes - perf.free_after_bytes;
    return perf.status;
}

// ---------------- profile ----------------

const char *to_cstr(SpgemmStatus s)
{
    switch (s) {
        case SpgemmStatus::ok: return "ok";
        case SpgemmStatus::nonsquare: return "nonsquare";
        case SpgemmStatus::int32_overflow: return "int32_overflow";
        case SpgemmStatus::oom: return "oom";
        case SpgemmStatus::cuda_error: return "cuda_error";
        case SpgemmStatus::exception: return "exception";
    }
    return "unknown";
}

void SpgemmPerf::reset()
{
    status = SpgemmStatus::ok;
    device_total_bytes = free_before_bytes = free_after_bytes = peak_used_bytes = attempted_bytes = 0;
    total_flop = 0;
    c_nnz = -1;
    error_msg.clear();
}

void SpgemmPerf::snapshot_free(std::size_t &out_free) const
{
    std::size_t total = 0;
    query_hip_memory(out_free, total);
}

void query_hip_memory(std::size_t &free_bytes, std::size_t &total_bytes)
{
    if (hipMemGetInfo(&free_bytes, &total_bytes) != hipSuccess) {
        free_bytes = 0;
        total_bytes = 0;
        (void)hipGetLastError();
    }
}

void SpgemmTimings::operator+=(const SpgemmTimings &b)
{
    setup += b.setup;
    symbolic_binning += b.symbolic_binning;
    symbolic += b.symbolic;
    reduce += b.reduce;
    numeric_binning += b.numeric_binning;
    prefix += b.prefix;
    allocate += b.allocate;
    numeric += b.numeric;
    cleanup += b.cleanup;
    total += b.total;
}

void SpgemmTimings::operator/=(double x)
{
    setup /= x;
    symbolic_binning /= x;
    symbolic /= x;
    reduce /= x;
    numeric_binning /= x;
    prefix /= x;
    allocate /= x;
    numeric /= x;
    cleanup /= x;
    total /= x;
}

void SpgemmTimings::print(double total_flop) const
{
    const double g = total_flop / 1e9;
    const double parts[] = {setup, symbolic_binning, symbolic, reduce, numeric_binning, prefix, allocate, numeric, cleanup};
    const char *names[] = {"setup", "symbolic_binning", "symbolic", "reduce", "numeric_binning",
                           "prefix", "allocate", "numeric", "cleanup"};
    double sum = 0;
    for (double p : parts) sum += p;
    std::printf("total flop %lf\n", total_flop);
    if (measure_separate) {
        std::printf("time(ms):\n");
        for (int i = 0; i < 9; i++)
            std::printf("    %-16s %8.3lfms %6.2lf%%\n", names[i], 1000 * parts[i], total > 0 ? parts[i] / total * 100 : 0.0);
        std::printf("    %-16s %8.3lfms %6.2lf%%\n", "sum_total", 1000 * sum, total > 0 ? sum / total * 100 : 0.0);
        std::printf("    %-16s %8.3lfms %6.2lf%%\n", "total", 1000 * total, 100.0);
        std::printf("perf(Gflops):\n");
        for (int i = 0; i < 9; i++) std::printf("    %-16s %6.2lf\n", names[i], parts[i] > 0 ? g / parts[i] : 0.0);
        std::printf("    %-16s %6.2lf\n", "total", total > 0 ? g / total : 0.0);
    }
}

void SpgemmTimings::reg_print(double total_flop) const { std::printf("%6.2lf\n", total > 0 ? total_flop / 1e9 / total : 0.0); }

void SpgemmTimings::perf_print(double total_flop) const
{
    const double g = total_flop / 1e9;
    std::printf("%6.2lf %6.2lf\n", symbolic > 0 ? g / symbolic : 0.0, numeric > 0 ? g / numeric : 0.0);
}

void SpgemmTimings::binning_print(double /*total_flop*/) const
{
    const double b = symbolic_binning + numeric_binning;
    std::printf("%.4le %.4lf\n", b, total > 0 ? 100 * b / total : 0.0);
}

}  // namespace opsparse
