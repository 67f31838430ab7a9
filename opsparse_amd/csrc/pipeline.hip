// Host driver of the SpGEMM pipeline and the C ABI (include/opsparse_amd.h).
//
// Phase order follows the reference driver (src/opsparse/core_algo/spgemm/spgemm.cu:16-126):
// setup -> symbolic binning -> symbolic -> scan + numeric binning ->
// allocate C -> numeric.  Differences, all MI355X-side:
//  * the workspace (streams, events, buffers) persists across calls instead
//    of being created and destroyed per call (meta.cu:16-65);
//  * two host synchronisations per call (after setup, after the scan) instead
//    of 4-7 device-wide syncs;
//  * per-bin kernels run on a pool of 8 HIP streams forked from and joined
//    back to the caller's stream with events, so the call is stream-ordered;
//  * C's row offsets are computed in 64 bits, so products above 2^31 nonzeros
//    are supported by the 64-bit entry points.
#include <hip/hip_runtime.h>

#include <time.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "batch.hpp"
#include "internal.hpp"
#include "launch.hpp"
#include "opsparse_amd.h"
#include "outmem.hpp"

using namespace opsamd;

namespace {

double now_s()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC_RAW, &ts);
    return static_cast<double>(ts.tv_sec) + static_cast<double>(ts.tv_nsec) * 1e-9;
}

struct HipFail : std::runtime_error {
    int status;
    HipFail(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};

void check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        const int st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? OPS_OOM : OPS_HIP_ERROR;
        (void)hipGetLastError();
        throw HipFail(st, std::string(what) + ": " + hipGetErrorString(e));
    }
}
#define CK(x) check((x), #x)

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    void ensure(size_t n)
    {
        if (n <= cap && p) return;
        if (p) CK(hipFree(p));
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(n + n / 8, 256);
        CK(hipMalloc(&p, want * sizeof(T)));
        cap = want;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

bool g_attrs_done = false;

// Bytes of long-row column bitmaps held per device by every workspace and
// plan of the process together, against one limit (ops_kept_bitmap_limit).
constexpr int kMaxDev = 64;
constexpr long long kMaxKeptBitmapBytes = 64ll << 30;
std::mutex g_kb_mu;
long long g_kb_held[kMaxDev] = {};
long long g_kb_limit = -1;  // < 0: min(kMaxKeptBitmapBytes, a quarter of the device's memory)

// Column bitmaps one holder (a workspace or a plan) keeps between the
// symbolic and numeric phases; its bytes count against the device's limit.
struct KeptBitmaps {
    DevBuf<unsigned> bm;
    DevBuf<int> slot;
    int dev = -1;
    long long accounted = 0;
    // Bytes this holder may hold: the device limit minus what the others
    // hold, and at most a quarter of free memory (counting its own as free).
    long long budget(size_t free_b, size_t total_b)
    {
        std::lock_guard<std::mutex> g(g_kb_mu);
        const long long lim = g_kb_limit >= 0 ? g_kb_limit
                                              : std::min<long long>(kMaxKeptBitmapBytes,
                                                                    static_cast<long long>(total_b) / 4);
        const int d = dev >= 0 ? dev : 0;
        const long long others = g_kb_held[d] - accounted;
        const long long by_free = (static_cast<long long>(free_b) + accounted) / 4;
        return std::max<long long>(0, std::min(lim - others, by_free));
    }
    void account()
    {
        std::lock_guard<std::mutex> g(g_kb_mu);
        const long long now = static_cast<long long>(bm.cap * sizeof(unsigned));
        if (dev >= 0 && dev < kMaxDev) g_kb_held[dev] += now - accounted;
        accounted = now;
    }
    void release()
    {
        bm.release();
        slot.release();
        account();
    }
};

}  // namespace

struct ops_workspace {
    int device = 0;
    hipStream_t pool[kNumBins] = {};
    hipStream_t cap = nullptr;  // origin stream for graph capture
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[kNumBins] = {};
    hipEvent_t ev_k0[kNumBins] = {}, ev_k1[kNumBins] = {};
    DevStats *d_stats = nullptr;
    DevStats *h_stats = nullptr;  // pinned
    DevBuf<int> row_flop, row_nnz, bins, cursor, spill, wcut, crpt32_tmp;
    KeptBitmaps kb;  // column bitmaps kept by the symbolic phase for one call
    DevBuf<long long> block_sums, crpt64;
    BinBounds bounds = default_bounds();
    int short_hash = kHashLin;
    bool sym_bitmap = false;
    int sym_shape[kNumBins] = {}, num_shape[kNumBins] = {};  // kernel shapes of bins 0-4
    // Set only on the stack workspace of the allocation-free entry point:
    // every bin runs on the caller's stream, and neither kept column
    // bitmaps nor a window-cut table are made (both would allocate).
    bool no_alloc = false;
    std::string err;
};

// Everything the numeric phase needs, produced by the symbolic phase.
struct PlanCore {
    int M = 0, N = 0;
    long long nnz = 0;
    long long *crpt64 = nullptr;  // M+1, device
    int *bins = nullptr;          // numeric bin lists (nullptr: identity over one bin)
    LongScratch ls;               // long-row bin only
    int *spill = nullptr;         // bin-5 rows handed from the wave kernel to the block kernel
    int num_size[kNumBins] = {};
    int num_off[kNumBins] = {};
    int hash = kHashLin;  // hash family of the short-row bins
    int num_shape[kNumBins] = {};
    int kept_rows = 0;    // long rows whose symbolic column bitmap is kept (ls.bm)
    DevStats stats{};
};

struct ops_plan {
    ops_workspace *ws = nullptr;
    PlanCore core;
    DevBuf<long long> crpt64;
    DevBuf<int> bins, cursor, spill, wcut;
    KeptBitmaps kb;  // column bitmaps of long rows kept for every numeric phase of the plan
    int annz = 0;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;
    // arguments the captured graph was recorded with
    const void *g_args[8] = {};
    int g_dt = -1;
};

namespace {

CsrIn as_in(const ops_csr *m) { return CsrIn{m->rpt, m->col, m->val, m->M, m->N}; }

void validate(const ops_csr *A, const ops_csr *B)
{
    if (!A || !B) throw HipFail(OPS_INVALID, "null matrix descriptor");
    if (A->M < 0 || A->N < 0 || B->M < 0 || B->N < 0 || A->nnz < 0 || B->nnz < 0)
        throw HipFail(OPS_INVALID, "negative dimension");
    if (A->N != B->M) throw HipFail(OPS_NONSQUARE, "A.N != B.M");
    if (A->M > 0 && (!A->rpt || (A->nnz > 0 && !A->col))) throw HipFail(OPS_INVALID, "A device arrays missing");
    if (B->M > 0 && (!B->rpt || (B->nnz > 0 && !B->col))) throw HipFail(OPS_INVALID, "B device arrays missing");
}

void init_attrs()
{
    if (g_attrs_done) return;
    (void)hipGetLastError();
    symbolic_init_attributes();
    numeric_init_attributes();
    backward_init_attributes();
    // An attribute the device refuses (a kernel asking for more LDS than a
    // block may hold) is reported here, not by a later launch check.
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        throw HipFail(OPS_HIP_ERROR, std::string("kernel attribute setup: ") + hipGetErrorString(e));
    g_attrs_done = true;
}

// Runs `launch(bin, stream)` for each non-empty bin on the pool streams,
// forked from and joined back to `s`.  ms[] receives event-timed kernel
// durations when timing is on.
template <typename F>
void dispatch_bins(ops_workspace *ws, hipStream_t s, const int *size, bool timing, F &&launch)
{
    int used = 0;
    for (int b = 0; b < kNumBins; b++) used += size[b] > 0 ? 1 : 0;
    if (used == 0) return;
    if (ws->no_alloc) {
        // one stream, no events (the workspace has none)
        for (int b = kNumBins - 1; b >= 0; b--) {
            if (size[b] <= 0) continue;
            launch(b, s);
            CK(hipGetLastError());
        }
        return;
    }
    // The fork is recorded before the first launch, so the pool streams do
    // not wait for a bin that runs on s itself.
    if (used > 1) CK(hipEventRecord(ws->ev_fork, s));
    // Largest bins first so the long kernels start earliest.  The first one
    // runs on the caller's stream (no fork or join for it), except when the
    // long-row bin has rows: its kernel then runs long enough that a fork
    // costs nothing, and every bin takes a pool stream, the k-th used bin
    // pool[k].  The pool's streams are created back to back and the runtime
    // gives streams hardware queues round-robin, so the largest bins then sit
    // on different queues; the caller's stream may share a queue with the
    // stream of another bin, which then starts only after the long-row kernel.
    const bool all_pool = used > 1 && size[kNumBins - 1] > 0;
    const int order[kNumBins] = {7, 6, 5, 4, 3, 2, 1, 0};
    bool first = true;
    int k = 0;
    for (int oi = 0; oi < kNumBins; oi++) {
        const int b = order[oi];
        if (size[b] <= 0) continue;
        const bool on_s = first && !all_pool;
        hipStream_t ps = on_s ? s : ws->pool[all_pool ? k++ : b];
        if (!on_s) CK(hipStreamWaitEvent(ps, ws->ev_fork, 0));
        if (timing) CK(hipEventRecord(ws->ev_k0[b], ps));
        launch(b, ps);
        CK(hipGetLastError());
        if (timing) CK(hipEventRecord(ws->ev_k1[b], ps));
        if (!on_s) {
            CK(hipEventRecord(ws->ev_join[b], ps));
            CK(hipStreamWaitEvent(s, ws->ev_join[b], 0));
        }
        first = false;
    }
}

void read_kernel_ms(ops_workspace *ws, const int *size, double *ms)
{
    for (int b = 0; b < kNumBins; b++) {
        ms[b] = 0.0;
        if (size[b] <= 0) continue;
        float f = 0.f;
        CK(hipEventSynchronize(ws->ev_k1[b]));
        CK(hipEventElapsedTime(&f, ws->ev_k0[b], ws->ev_k1[b]));
        ms[b] = f;
    }
}

// Column windows of the long-row numeric kernel over B's columns.
int rank_windows(int n) { return static_cast<int>((static_cast<long long>(n) + kRankWindow - 1) / kRankWindow); }

// Whether the window-cut table comes with the unit cut table (sub-window
// ends found between unit cuts).  OPSAMD_UNIT_CUTS=0 turns it off (galloping
// from the sub-window start), for tests.
bool unit_cuts_wanted()
{
    const char *env = std::getenv("OPSAMD_UNIT_CUTS");
    return !(env && env[0] == '0');
}

// Whether the long-row numeric kernel gets a window-cut table of B: more
// than one window and at most kMaxCutBytes.  OPSAMD_CUT_TABLE=0 turns the
// table off (galloping path), for tests.
bool cut_table_wanted(int bm, int bn)
{
    const int nwin = rank_windows(bn);
    const long long bytes = static_cast<long long>(sizeof(int)) * bm * (nwin + 1);
    const char *env = std::getenv("OPSAMD_CUT_TABLE");
    const bool off = env && env[0] == '0';
    return nwin > 1 && bytes <= kMaxCutBytes && !off;
}

// Setup, symbolic and scan phases.  crpt64_dst: where C's 64-bit row
// offsets go; crpt32_dst (optional): int32 copy, written only if nnz fits.
// kb: where long rows may leave their column bitmaps for the numeric phase
// (pc.ls.bm): the workspace's buffers for one call, a plan's own for a
// plan, whose numeric phase runs later and repeatedly.
void symbolic_phase(ops_workspace *ws, const CsrIn &A, const CsrIn &B, PlanCore &pc, long long *crpt64_dst,
                    int *crpt32_dst, int *bins_dst, hipStream_t s, ops_timings *t, KeptBitmaps &kb)
{
    const bool sep = t && t->measure_separate;
    const bool ev = t && t->kernel_events;
    const int M = A.M;
    pc.M = M;
    pc.N = B.N;
    pc.crpt64 = crpt64_dst;
    pc.hash = ws->short_hash;
    for (int b = 0; b < kNumBins; b++) pc.num_shape[b] = ws->num_shape[b];
    // When B's columns fit one bitmap window of the long-row symbolic
    // kernel, rows above the bin-5 bound go to it (bin 7) instead of the
    // 16384-slot hash of bin 6: setting bits costs no probing or CAS, and
    // the one window is walked once.  With more columns every window would
    // walk the row again, and bin 6 keeps hashing.
    BinBounds bb = ws->bounds;
    if (B.N <= kSymLongWindow) bb.sym[kNumBins - 2] = bb.sym[kNumBins - 3];
    double t0 = now_s();

    ws->row_flop.ensure(static_cast<size_t>(M) + 1);
    ws->row_nnz.ensure(static_cast<size_t>(M) + 1);
    ws->block_sums.ensure(static_cast<size_t>(scan_blocks(M)) + 2);
    launch_zero_stats(ws->d_stats, s);
    launch_setup(A, B, ws->row_flop.p, ws->row_nnz.p, ws->d_stats, bb, s);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(ws->h_stats, ws->d_stats, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    DevStats st = *ws->h_stats;
    if (t) {
        t->setup = now_s() - t0;
        t->total_flop = static_cast<int64_t>(st.total_flop);
        t->max_row_flop = st.max_flop;
        for (int b = 0; b < kNumBins; b++) {
            t->sym_rows[b] = st.sym_size[b];
            t->sym_flop[b] = static_cast<int64_t>(st.sym_flop[b]);
            t->sym_arows_nnz[b] = static_cast<int64_t>(st.sym_annz[b]);
        }
    }

    // Symbolic binning: identity when every row is in one bin.
    t0 = now_s();
    int sym_off[kNumBins];
    const int *sym_rows = nullptr;
    {
        int run = 0, nonempty = 0, only = -1;
        for (int b = 0; b < kNumBins; b++) {
            sym_off[b] = run;
            run += st.sym_size[b];
            if (st.sym_size[b]) {
                nonempty++;
                only = b;
            }
        }
        if (!(nonempty == 1 && st.sym_size[only] == M) && nonempty > 0) {
            launch_bin_scatter(ws->row_flop.p, M, 0, ws->d_stats, bins_dst, bb, s);
            CK(hipGetLastError());
            sym_rows = bins_dst;
        }
    }
    if (sep) {
        CK(hipStreamSynchronize(s));
        t->symbolic_binning = now_s() - t0;
    }

    // Bitmaps of rows bound for the long-row numeric kernel are kept when
    // B's columns fit one symbolic window and every numeric window's B
    // ranges are known without a mark walk (one window, or the cut table).
    SymSave save;
    const int long_rows = st.sym_size[kNumBins - 1];
    if (long_rows > 0 && B.N <= kSymLongWindow && !ws->no_alloc &&
        (rank_windows(B.N) == 1 || cut_table_wanted(B.M, B.N))) {
        const long long words = (static_cast<long long>(B.N) + 31) / 32;
        size_t free_b = 0, total_b = 0;
        CK(hipMemGetInfo(&free_b, &total_b));
        if (kb.dev < 0) kb.dev = ws->device;
        long long cap = std::min<long long>(long_rows, kb.budget(free_b, total_b) / (4 * words));
        // OPSAMD_KEPT_BITMAPS=n keeps at most n (0: none), for tests.
        const char *env = std::getenv("OPSAMD_KEPT_BITMAPS");
        if (env && env[0]) cap = std::min(cap, std::atoll(env));
        if (cap > 0) {
            // Keeping bitmaps is an optimisation: when their memory cannot
            // be had, the numeric phase marks the rows again.
            try {
                kb.bm.ensure(static_cast<size_t>(cap * words));
                kb.slot.ensure(static_cast<size_t>(M));
            } catch (const HipFail &e) {
                if (e.status != OPS_OOM) throw;
                kb.release();
                cap = 0;
            }
            kb.account();
        }
        if (cap > 0) {
            // slot = -1 for every row; the kept-row counter (in d_stats) is zero
            launch_fill_int(kb.slot.p, -1, M, s);
            CK(hipGetLastError());
            save = SymSave{kb.bm.p, kb.slot.p, &ws->d_stats->kept_rows, bb.num[kNumBins - 2], static_cast<int>(cap)};
            save.by_index = cap >= long_rows;
            pc.ls.bm = kb.bm.p;
            pc.ls.bm_slot = kb.slot.p;
            pc.ls.bm_words = static_cast<int>(words);
        }
    }

    t0 = now_s();
    dispatch_bins(ws, s, st.sym_size, ev, [&](int b, hipStream_t ps) {
        const RowList rl{sym_rows ? sym_rows + sym_off[b] : nullptr, st.sym_size[b]};
        launch_symbolic_bin(b, A, B, rl, ws->row_nnz.p, pc.hash, ws->sym_bitmap, save, ps, ws->sym_shape[b]);
    });
    if (sep) {
        CK(hipStreamSynchronize(s));
        t->symbolic = now_s() - t0;
    }
    if (ev) read_kernel_ms(ws, st.sym_size, t->sym_ms);

    t0 = now_s();
    launch_scan(ws->row_nnz.p, ws->row_flop.p, A.rpt, M, ws->block_sums.p, crpt64_dst, crpt32_dst, ws->d_stats, bb,
                s);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(ws->h_stats, ws->d_stats, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    st = *ws->h_stats;
    pc.stats = st;
    pc.nnz = static_cast<long long>(st.total_nnz);
    pc.kept_rows = save.bm ? std::min(st.kept_rows, save.cap) : 0;
    if (t) {
        t->prefix = now_s() - t0;
        t->nnz_c = pc.nnz;
        t->kept_rows = pc.kept_rows;
        t->max_row_nnz = st.max_nnz;
        for (int b = 0; b < kNumBins; b++) {
            t->num_rows[b] = st.num_size[b];
            t->num_flop[b] = static_cast<int64_t>(st.num_flop[b]);
            t->num_nnz[b] = static_cast<int64_t>(st.num_nnz[b]);
            t->num_arows_nnz[b] = static_cast<int64_t>(st.num_annz[b]);
        }
    }

    // Numeric binning (the rows' bins are final once nnz is known).
    t0 = now_s();
    int run = 0, nonempty = 0, only = -1;
    for (int b = 0; b < kNumBins; b++) {
        pc.num_size[b] = st.num_size[b];
        pc.num_off[b] = run;
        run += st.num_size[b];
        if (st.num_size[b]) {
            nonempty++;
            only = b;
        }
    }
    pc.bins = nullptr;
    if (!(nonempty == 1 && st.num_size[only] == M) && nonempty > 0) {
        launch_bin_scatter(ws->row_nnz.p, M, 1, ws->d_stats, bins_dst, bb, s);
        CK(hipGetLastError());
        pc.bins = bins_dst;
    }
    pc.ls.nheavy = pc.bins ? st.num_heavy : -1;
    if (sep) {
        CK(hipStreamSynchronize(s));
        t->numeric_binning = now_s() - t0;
    }
}

template <typename V>
void numeric_phase(ops_workspace *ws, const CsrIn &A, const CsrIn &B, const PlanCore &pc, int *ccol, V *cval,
                   hipStream_t s, ops_timings *t)
{
    const bool ev = t && t->kernel_events;
    dispatch_bins(ws, s, pc.num_size, ev, [&](int b, hipStream_t ps) {
        const RowList rl{pc.bins ? pc.bins + pc.num_off[b] : nullptr, pc.num_size[b]};
        launch_numeric_bin<V>(b, A, B, rl, pc.crpt64, ccol, cval, pc.ls, pc.spill, pc.hash, ps, pc.num_shape[b]);
    });
}

template <typename F>
int guarded(ops_workspace *ws, F &&f)
{
    try {
        return f();
    } catch (const HipFail &e) {
        if (ws) ws->err = e.what();
        return e.status;
    } catch (const std::bad_alloc &e) {
        if (ws) ws->err = e.what();
        return OPS_OOM;
    } catch (const std::exception &e) {
        if (ws) ws->err = e.what();
        return OPS_EXCEPTION;
    } catch (...) {
        if (ws) ws->err = "unknown exception";
        return OPS_EXCEPTION;
    }
}

// Column-index span of A's descriptor: the highest A-nonzero index + 1,
// which is what per-A-nonzero scratch must cover (descriptors may be row
// slices whose rpt does not start at 0).
// B entry indices below this bound let the numeric long-row walk address
// B by 32-bit byte offsets (index * 8 < 2^32).
constexpr long long kOff32Entries = 1ll << 29;

long long a_index_span(const ops_csr *A, hipStream_t s)
{
    // (also used for B: its entry indices end at the value returned)
    if (A->M == 0) return 0;
    int last = 0;
    CK(hipMemcpyAsync(&last, A->rpt + A->M, sizeof(int), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    return last;
}

// Entries [rpt[0], rpt[M]) of a descriptor, read from the device.
struct EntryRange {
    long long begin = 0, end = 0;
};
EntryRange entry_range(const ops_csr *X, hipStream_t s)
{
    if (X->M == 0) return {};
    int v[2] = {0, 0};
    CK(hipMemcpyAsync(&v[0], X->rpt, sizeof(int), hipMemcpyDeviceToHost, s));
    CK(hipMemcpyAsync(&v[1], X->rpt + X->M, sizeof(int), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    if (v[1] < v[0]) throw HipFail(OPS_INVALID, "descriptor row offsets decrease");
    return EntryRange{v[0], v[1]};
}

// Scratch of the numeric phase: the spill list of bin kSpillBin and, for the
// long-row bin, the per-A-nonzero cursors and the window-cut table of B
// (built here on stream s when B has more than one column window and the
// table stays within kMaxCutBytes; it depends on B's pattern only).
void numeric_scratch(PlanCore &pc, const ops_csr *A, const ops_csr *B, DevBuf<int> &cursor, DevBuf<int> &spill,
                     DevBuf<int> &wcut, hipStream_t s, bool allow_cut = true)
{
    if (pc.num_size[kNumBins - 1] > 0) {
        pc.ls.span = static_cast<int>(a_index_span(A, s));
        pc.ls.off32 = a_index_span(B, s) < kOff32Entries;
        cursor.ensure(4 * static_cast<size_t>(pc.ls.span) + 4);
        pc.ls.cursor = cursor.p;
        const int nwin = rank_windows(B->N);
        if (allow_cut && cut_table_wanted(B->M, B->N)) {
            // the unit cut table follows the window-cut table when it fits
            // kMaxCutBytes too
            const long long nunit = (static_cast<long long>(B->N) + (1 << kUnitCutShift) - 1) >> kUnitCutShift;
            const size_t coarse = static_cast<size_t>(B->M) * (nwin + 1);
            const size_t fine = static_cast<size_t>(B->M) * static_cast<size_t>(nunit + 1);
            const bool units = static_cast<long long>(sizeof(int) * fine) <= kMaxCutBytes && unit_cuts_wanted();
            wcut.ensure(coarse + (units ? fine : 0));
            launch_window_cuts(as_in(B), nwin, wcut.p, s);
            CK(hipGetLastError());
            pc.ls.wcut = wcut.p;
            pc.ls.nwin = nwin;
            if (units) {
                launch_window_cuts(as_in(B), static_cast<int>(nunit), wcut.p + coarse, s, kUnitCutShift);
                CK(hipGetLastError());
                pc.ls.ucut = wcut.p + coarse;
                pc.ls.nunit = static_cast<int>(nunit);
            }
        }
        // Kept bitmaps need a window's B ranges without a mark walk.
        if (!pc.ls.wcut && nwin > 1) {
            pc.ls.bm = nullptr;
            pc.ls.bm_slot = nullptr;
            pc.ls.bm_words = 0;
        }
    }
    if (pc.num_size[kSpillBin] > 0) {
        spill.ensure(static_cast<size_t>(pc.num_size[kSpillBin]) + 2);
        pc.spill = spill.p;
    }
}

template <typename V>
int spgemm_impl(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_csr *C32, ops_csr64 *C64, hipStream_t s,
                ops_timings *t)
{
    const double t_start = now_s();
    validate(A, B);
    if (!C32 && !C64) throw HipFail(OPS_INVALID, "null output descriptor");
    init_attrs();
    const CsrIn a = as_in(A), b = as_in(B);
    const int M = A->M;
    PlanCore pc;
    ws->bins.ensure(static_cast<size_t>(M) + 1);
    // On any failure below, C's partial arrays are released.
    struct Cleanup {
        ops_csr *c32;
        ops_csr64 *c64;
        bool armed = true;
        ~Cleanup()
        {
            if (!armed) return;
            if (c32) ops_csr_free(c32);
            if (c64) ops_csr64_free(c64);
        }
    } cleanup{C32, C64};

    double t0 = now_s();
    long long *crpt64_dst;
    int *crpt32_dst = nullptr;
    if (C64) {
        *C64 = ops_csr64{};
        C64->M = M;
        C64->N = B->N;
        CK(out_malloc(reinterpret_cast<void **>(&C64->rpt), sizeof(int64_t) * (static_cast<size_t>(M) + 1)));
        crpt64_dst = reinterpret_cast<long long *>(C64->rpt);
    } else {
        *C32 = ops_csr{};
        C32->M = M;
        C32->N = B->N;
        CK(out_malloc(reinterpret_cast<void **>(&C32->rpt), sizeof(int32_t) * (static_cast<size_t>(M) + 1)));
        ws->crpt64.ensure(static_cast<size_t>(M) + 1);
        crpt64_dst = ws->crpt64.p;
        crpt32_dst = C32->rpt;
    }
    const double t_alloc_rpt = now_s() - t0;

    symbolic_phase(ws, a, b, pc, crpt64_dst, crpt32_dst, ws->bins.p, s, t, ws->kb);
    if (t) t->setup += t_alloc_rpt;

    if (C32 && pc.nnz > static_cast<long long>(INT32_MAX))
        throw HipFail(OPS_INT32_OVERFLOW, "nnz(C) exceeds INT32_MAX; use the 64-bit entry point");

    t0 = now_s();
    const size_t nz = static_cast<size_t>(std::max<long long>(pc.nnz, 1));
    int *ccol = nullptr;
    V *cval = nullptr;
    CK(out_malloc(reinterpret_cast<void **>(&ccol), nz * sizeof(int)));
    CK(out_malloc(reinterpret_cast<void **>(&cval), nz * sizeof(V)));
    if (C64) {
        C64->nnz = pc.nnz;
        C64->col = ccol;
        C64->val = cval;
    } else {
        C32->nnz = static_cast<int32_t>(pc.nnz);
        C32->col = ccol;
        C32->val = cval;
    }
    numeric_scratch(pc, A, B, ws->cursor, ws->spill, ws->wcut, s);
    if (t) t->allocate = now_s() - t0;

    t0 = now_s();
    numeric_phase<V>(ws, a, b, pc, ccol, cval, s, t);
    CK(hipStreamSynchronize(s));
    if (t) {
        t->numeric = now_s() - t0;
        if (t->kernel_events) read_kernel_ms(ws, pc.num_size, t->num_ms);
        t->cleanup = 0.0;
        t->total = now_s() - t_start;
    }
    cleanup.armed = false;
    return OPS_OK;
}

template <typename V>
void backward_impl(ops_workspace *ws, const ops_csr *A, const ops_csr *B, const int64_t *crpt64, const int32_t *ccol,
                   const V *dc, V *da, V *db, hipStream_t s)
{
    validate(A, B);
    init_attrs();
    const int M = A->M;
    if (M > 0 && !crpt64) throw HipFail(OPS_INVALID, "C row offsets missing");
    const CsrIn a = as_in(A), b = as_in(B);
    ws->row_nnz.ensure(static_cast<size_t>(M) + 1);
    ws->bins.ensure(static_cast<size_t>(M) + 1);
    ws->block_sums.ensure(static_cast<size_t>(scan_blocks(M)) + 2);
    const long long *crpt = reinterpret_cast<const long long *>(crpt64);
    // da and db are indexed from the first entry of A's and B's descriptors
    // (row slices need not start at entry 0).  dA entries of rows without
    // products and all of dB start at zero.
    const EntryRange ea = entry_range(A, s), eb = entry_range(B, s);
    launch_zero_values(da, ea.end - ea.begin, sizeof(V) == 4, s);
    launch_zero_values(db, eb.end - eb.begin, sizeof(V) == 4, s);
    launch_zero_stats(ws->d_stats, s);
    launch_rownnz_from_rpt(crpt, M, ws->row_nnz.p, s);
    CK(hipGetLastError());
    if (M == 0) return;
    // Numeric bin sizes of C's rows (the scan's own output is unused here).
    ws->crpt64.ensure(static_cast<size_t>(M) + 1);
    launch_scan(ws->row_nnz.p, ws->row_nnz.p, A->rpt, M, ws->block_sums.p, ws->crpt64.p, nullptr, ws->d_stats,
                ws->bounds, s);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(ws->h_stats, ws->d_stats, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    const DevStats st = *ws->h_stats;
    int off[kNumBins];
    int run = 0, nonempty = 0, only = -1;
    for (int bb = 0; bb < kNumBins; bb++) {
        off[bb] = run;
        run += st.num_size[bb];
        if (st.num_size[bb]) {
            nonempty++;
            only = bb;
        }
    }
    const int *rows = nullptr;
    if (!(nonempty == 1 && st.num_size[only] == M) && nonempty > 0) {
        launch_bin_scatter(ws->row_nnz.p, M, 1, ws->d_stats, ws->bins.p, ws->bounds, s);
        CK(hipGetLastError());
        rows = ws->bins.p;
    }
    // Rows of C above the largest LDS table walk dense column windows with
    // per-A-nonzero cursors (two ints per A-nonzero index, double-buffered).
    int *cursor = nullptr;
    int span = 0;
    const int *ucut = nullptr;
    int nunit = 0;
    if (st.num_size[kNumBins - 1] > 0) {
        span = static_cast<int>(ea.end);
        ws->cursor.ensure(2 * static_cast<size_t>(span) + 2);
        cursor = ws->cursor.p;
        // B's unit cut table bounds the search for each sub-window's end
        const long long nu = (static_cast<long long>(B->N) + (1 << kUnitCutShift) - 1) >> kUnitCutShift;
        const size_t fine = static_cast<size_t>(B->M) * static_cast<size_t>(nu + 1);
        if (static_cast<long long>(sizeof(int) * fine) <= kMaxCutBytes && unit_cuts_wanted()) {
            ws->wcut.ensure(fine);
            launch_window_cuts(b, static_cast<int>(nu), ws->wcut.p, s, kUnitCutShift);
            CK(hipGetLastError());
            ucut = ws->wcut.p;
            nunit = static_cast<int>(nu);
        }
    }
    const int a0 = static_cast<int>(ea.begin), b0 = static_cast<int>(eb.begin);
    dispatch_bins(ws, s, st.num_size, false, [&](int bin, hipStream_t ps) {
        launch_backward_bin<V>(bin, a, b, rows ? rows + off[bin] : nullptr, st.num_size[bin], crpt, ccol, dc, da, db,
                               a0, b0, cursor, span, ucut, nunit, ps);
    });
}

}  // namespace

// ---- allocation-free entry point (ops_spgemm64_noalloc) -------------------
//
// Its scratch (row products, row nnz, bin lists, the spill list, the scan's
// block sums, the counter block and the long-row cursors) lives in the
// file-scope array below, which the code object brings with it: every GPU
// the library runs on holds one copy until the process ends.  Limits: A
// with at most kNaMaxRows rows and at most kNaMaxSpan entries (index of its
// last entry + 1).
constexpr int kNaMaxRows = 1 << 18;
constexpr int kNaMaxSpan = 1 << 20;
constexpr size_t kNaAlign = 256;
constexpr size_t na_round(size_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr size_t kNaRowBytes = na_round(sizeof(int) * (static_cast<size_t>(kNaMaxRows) + 1));
constexpr size_t kNaSumEntries = 256;  // >= scan_blocks(kNaMaxRows) + 2
constexpr size_t kNaOffFlop = 0;
constexpr size_t kNaOffNnz = kNaOffFlop + kNaRowBytes;
constexpr size_t kNaOffBins = kNaOffNnz + kNaRowBytes;
constexpr size_t kNaOffSpill = kNaOffBins + kNaRowBytes;
constexpr size_t kNaOffSums = kNaOffSpill + kNaRowBytes;
constexpr size_t kNaOffStats = kNaOffSums + na_round(sizeof(long long) * kNaSumEntries);
constexpr size_t kNaOffCursor = kNaOffStats + na_round(sizeof(DevStats));
constexpr size_t kNaBytes = kNaOffCursor + na_round(sizeof(int) * (4 * static_cast<size_t>(kNaMaxSpan) + 4));

static __device__ __align__(256) unsigned long long scratch[kNaBytes / sizeof(unsigned long long)];

namespace {

// Device address of `scratch` on each device, looked up once per device
// with that device current (the array's name in host code is only a
// stand-in for the symbol, not an address).
std::mutex g_na_mu;
unsigned char *g_na_ptr[kMaxDev] = {};

unsigned char *na_scratch(int dev)
{
    if (dev < 0 || dev >= kMaxDev) throw HipFail(OPS_INVALID, "device index out of range");
    std::lock_guard<std::mutex> g(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        CK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_na_ptr[dev] = static_cast<unsigned char *>(p);
    }
    return g_na_ptr[dev];
}

template <typename T>
void na_view(DevBuf<T> &b, unsigned char *base, size_t off, size_t bytes)
{
    b.p = reinterpret_cast<T *>(base + off);
    b.cap = bytes / sizeof(T);
}

template <typename V>
int spgemm_noalloc_impl(const ops_csr *A, const ops_csr *B, ops_csr64 *C, int64_t capacity, hipStream_t s)
{
    validate(A, B);
    if (!C || (capacity > 0 && (!C->rpt || !C->col || !C->val)) || (A->M > 0 && !C->rpt) || capacity < 0)
        throw HipFail(OPS_INVALID, "output arrays missing");
    if (A->M > kNaMaxRows) throw HipFail(OPS_INVALID, "A has more rows than the allocation-free limit (262144)");
    const EntryRange ea = entry_range(A, s);
    if (ea.end > kNaMaxSpan)
        throw HipFail(OPS_INVALID, "A's entries exceed the allocation-free limit (1048576 entry indices)");
    init_attrs();
    int dev = 0;
    CK(hipGetDevice(&dev));
    unsigned char *base = na_scratch(dev);
    // A workspace on the stack whose buffers are views of the scratch (each
    // sized to the limits above, so none ever grows) and which has no
    // streams, events or pinned memory.
    ops_workspace ws;
    ws.device = dev;
    ws.no_alloc = true;
    DevStats h_stats{};
    ws.d_stats = reinterpret_cast<DevStats *>(base + kNaOffStats);
    ws.h_stats = &h_stats;
    na_view(ws.row_flop, base, kNaOffFlop, kNaRowBytes);
    na_view(ws.row_nnz, base, kNaOffNnz, kNaRowBytes);
    na_view(ws.bins, base, kNaOffBins, kNaRowBytes);
    na_view(ws.spill, base, kNaOffSpill, kNaRowBytes);
    na_view(ws.block_sums, base, kNaOffSums, sizeof(long long) * kNaSumEntries);
    na_view(ws.cursor, base, kNaOffCursor, kNaBytes - kNaOffCursor);
    const CsrIn a = as_in(A), b = as_in(B);
    const int M = A->M;
    C->M = M;
    C->N = B->N;
    C->nnz = 0;
    PlanCore pc;
    symbolic_phase(&ws, a, b, pc, reinterpret_cast<long long *>(C->rpt), nullptr, ws.bins.p, s, nullptr, ws.kb);
    C->nnz = pc.nnz;
    if (pc.nnz > capacity) throw HipFail(OPS_INVALID, "nnz(C) exceeds the output capacity (C->nnz holds it)");
    numeric_scratch(pc, A, B, ws.cursor, ws.spill, ws.wcut, s, false);
    numeric_phase<V>(&ws, a, b, pc, C->col, static_cast<V *>(C->val), s, nullptr);
    CK(hipStreamSynchronize(s));
    return OPS_OK;
}

}  // namespace

// ---- whole calls with device-side sizes (ops_exec_*) ----------------------
//
// Every step of a call is issued on one stream without a host
// synchronisation: bin sizes, bin offsets and nnz(C) stay on the device.
// Each per-bin kernel is launched for every bin at a width set by the row
// count and reads its bin's size and offset from DevStats (RowList with
// dsize), looping over its tiles; C goes to buffers of fixed capacity, and a
// check between the scan and the numeric phase empties every numeric bin
// when nnz(C) exceeds it (the launch then writes no entry and the overflow
// is reported).  The sequence is captured once into a hipGraph and replayed:
// any patterns of A and B with the same descriptors (pointers, sizes, entry
// counts) give a correct product.
//
// A measuring pass at creation records which bins hold rows.  Those run
// concurrently on the pool streams; the others are launched after the join,
// on the main stream, where they find their bin empty and exit at once (or
// do the work, if a later pattern fills them).  Launched concurrently, an
// empty kernel whose blocks need much LDS would wait for a whole CU's LDS
// and hold back the kernels queued behind it.
//
// A batched executor packs the A_i and B_i in front of the sequence (pack
// kernels over fixed instance tables).  The A_i are packed block-diagonally
// and the B_i stacked (rows one after the other, columns not shifted), so
// the product's rows of instance i are A_i * B_i in B_i's own columns.
struct ops_exec {
    ops_workspace *ws = nullptr;
    ops_dtype dt = OPS_F64;
    bool use_graph = true;
    ops_csr A{}, B{};  // the product's operands (the packed arrays for a batch)
    // batch: instance tables and packed operand arrays
    bool batch = false;
    int count = 0;
    PackedSide pa, pb;
    DevBuf<int> a_rpt, a_col, b_rpt, b_col;
    DevBuf<double> a_val, b_val;  // FP32 batches use the first half of each
    // scratch of the sequence (its own: a graph keeps these addresses)
    DevBuf<int> row_flop, row_nnz, bins, spill, cursor, wcut;
    DevBuf<long long> block_sums, info;
    DevStats *stats = nullptr;
    int span = 0, nwin = 0;
    bool off32 = false;  // B's entry indices stay below kOff32Entries
    bool cut = false;
    // bins holding rows in the measuring pass (symbolic, numeric)
    int busy_sym[kNumBins] = {}, busy_num[kNumBins] = {};
    // C: row offsets (M + 1), columns and values with room for `capacity`
    long long capacity = 0;
    DevBuf<long long> crpt;
    DevBuf<int> ccol;
    DevBuf<double> cval;
    hipGraph_t src = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Bins that held rows at creation run concurrently on the pool streams; the
// rest follow on s after the join.
template <typename F>
void dispatch_device_sized(ops_workspace *ws, hipStream_t s, const int *busy, F &&launch)
{
    dispatch_bins(ws, s, busy, false, launch);
    for (int b = kNumBins - 1; b >= 0; b--) {
        if (busy[b] > 0) continue;
        launch(b, s);
        CK(hipGetLastError());
    }
}

// The whole call of `ex` on stream s (no host synchronisation).  measure:
// stop after the capacity check (no numeric phase), with every numeric bin
// left as binned.
void exec_sequence(ops_exec *ex, hipStream_t s, bool measure = false)
{
    ops_workspace *ws = ex->ws;
    const bool f64 = ex->dt == OPS_F64;
    if (ex->batch) {
        pack_side_launch(ex->pa, f64, ex->a_rpt.p, ex->a_col.p, ex->a_val.p, s);
        pack_side_launch(ex->pb, f64, ex->b_rpt.p, ex->b_col.p, ex->b_val.p, s);
        CK(hipGetLastError());
    }
    const CsrIn A = as_in(&ex->A), B = as_in(&ex->B);
    const int M = A.M;
    BinBounds bb = ws->bounds;
    if (B.N <= kSymLongWindow) bb.sym[kNumBins - 2] = bb.sym[kNumBins - 3];
    DevStats *st = ex->stats;
    launch_zero_stats(st, s);
    if (M > 0) {
        launch_setup(A, B, ex->row_flop.p, ex->row_nnz.p, st, bb, s);
        launch_bin_scatter(ex->row_flop.p, M, 0, st, ex->bins.p, bb, s);
        CK(hipGetLastError());
        dispatch_device_sized(ws, s, ex->busy_sym, [&](int b, hipStream_t ps) {
            launch_symbolic_bin(b, A, B, RowList{ex->bins.p, M, st->sym_size, b, st->next_sym + b}, ex->row_nnz.p,
                                ws->short_hash, ws->sym_bitmap, SymSave{}, ps, ws->sym_shape[b]);
        });
    }
    launch_scan(ex->row_nnz.p, ex->row_flop.p, A.rpt, M, ex->block_sums.p, ex->crpt.p, nullptr, st, bb, s);
    if (M > 0) launch_bin_scatter(ex->row_nnz.p, M, 1, st, ex->bins.p, bb, s);
    launch_capacity_check(st, measure ? LLONG_MAX : ex->capacity, ex->info.p, s);
    CK(hipGetLastError());
    if (M == 0 || measure) return;
    LongScratch ls;
    ls.cursor = ex->cursor.p;
    ls.span = ex->span;
    ls.off32 = ex->off32;
    if (ex->cut) {
        launch_window_cuts(B, ex->nwin, ex->wcut.p, s);
        CK(hipGetLastError());
        ls.wcut = ex->wcut.p;
        ls.nwin = ex->nwin;
    }
    dispatch_device_sized(ws, s, ex->busy_num, [&](int b, hipStream_t ps) {
        const RowList rl{ex->bins.p, M, st->num_size, b, st->next_num + b, st->next_num2 + b};
        if (f64)
            launch_numeric_bin<double>(b, A, B, rl, ex->crpt.p, ex->ccol.p, ex->cval.p, ls, ex->spill.p,
                                       ws->short_hash, ps, ws->num_shape[b]);
        else
            launch_numeric_bin<float>(b, A, B, rl, ex->crpt.p, ex->ccol.p, reinterpret_cast<float *>(ex->cval.p), ls,
                                      ex->spill.p, ws->short_hash, ps, ws->num_shape[b]);
    });
}

// Scratch of the sequence, sized from the operand descriptors (and, for the
// long-row bin, from A's last entry index, read once here).
void exec_scratch(ops_exec *ex, hipStream_t s)
{
    const int M = ex->A.M;
    ex->row_flop.ensure(static_cast<size_t>(M) + 1);
    ex->row_nnz.ensure(static_cast<size_t>(M) + 1);
    ex->bins.ensure(static_cast<size_t>(M) + 1);
    ex->spill.ensure(static_cast<size_t>(M) + 2);
    ex->block_sums.ensure(static_cast<size_t>(scan_blocks(M)) + 2);
    ex->info.ensure(2);
    ex->crpt.ensure(static_cast<size_t>(M) + 1);
    if (!ex->stats) CK(hipMalloc(&ex->stats, sizeof(DevStats)));
    // For a batch the packed A's last entry index is its entry count; a
    // descriptor is read on the device (its rpt[M]).
    ex->span = ex->batch ? ex->A.nnz : static_cast<int>(a_index_span(&ex->A, s));
    // A batch's packed B starts at entry 0; a given B keeps the entry range
    // it has at creation (patterns may change, entry counts may not).
    const long long b_end = ex->batch ? ex->B.nnz : std::max<long long>(ex->B.nnz, a_index_span(&ex->B, s));
    ex->off32 = b_end < kOff32Entries;
    ex->cursor.ensure(4 * static_cast<size_t>(ex->span) + 4);
    ex->nwin = rank_windows(ex->B.N);
    ex->cut = cut_table_wanted(ex->B.M, ex->B.N);
    if (ex->cut) ex->wcut.ensure(static_cast<size_t>(ex->B.M) * (ex->nwin + 1));
}

void exec_set_capacity(ops_exec *ex, long long cap)
{
    ex->capacity = cap;
    const size_t n = static_cast<size_t>(std::max<long long>(cap, 1));
    ex->ccol.ensure(n);
    // FP32 values use the first half of the FP64-sized buffer's bytes
    ex->cval.ensure(ex->dt == OPS_F64 ? n : (n + 1) / 2);
}

// Reads nnz(C) and the overflow flag of the last completed sequence.
void exec_info(ops_exec *ex, long long &nnz, bool &over, hipStream_t s)
{
    long long h[2] = {0, 0};
    CK(hipMemcpyAsync(h, ex->info.p, sizeof(h), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    nnz = h[0];
    over = h[1] != 0;
}

// The measuring pass (which bins hold rows, nnz(C)), then C's arrays:
// `capacity` entries, or exactly nnz(C) when capacity <= 0.  Then the graph
// is captured (on the workspace's capture stream) when asked for.
void exec_finish(ops_exec *ex, long long capacity, hipStream_t s)
{
    exec_scratch(ex, s);
    exec_sequence(ex, s, true);
    DevStats h{};
    CK(hipMemcpyAsync(&h, ex->stats, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (int b = 0; b < kNumBins; b++) {
        ex->busy_sym[b] = h.sym_size[b];
        ex->busy_num[b] = h.num_size[b];
    }
    if (capacity <= 0) capacity = static_cast<long long>(h.total_nnz);
    exec_set_capacity(ex, capacity);
    const long long zero[2] = {0, 0};  // nothing launched yet: nnz 0, no overflow
    CK(hipMemcpyAsync(ex->info.p, zero, sizeof(zero), hipMemcpyHostToDevice, s));
    CK(hipStreamSynchronize(s));
    if (!ex->use_graph) return;
    hipStream_t cs = ex->ws->cap;
    CK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
        exec_sequence(ex, cs);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cs, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    CK(hipStreamEndCapture(cs, &ex->src));
    CK(hipGraphInstantiate(&ex->exec, ex->src, nullptr, nullptr, 0));
}

void exec_release(ops_exec *ex)
{
    if (ex->exec) (void)hipGraphExecDestroy(ex->exec);
    if (ex->src) (void)hipGraphDestroy(ex->src);
    if (ex->stats) (void)hipFree(ex->stats);
    ex->pa.release();
    ex->pb.release();
    // DevBuf frees nothing by itself (the allocation-free entry point points
    // its buffers at static memory), so every owned buffer is released here.
    for (DevBuf<int> *b : {&ex->a_rpt, &ex->a_col, &ex->b_rpt, &ex->b_col, &ex->row_flop, &ex->row_nnz, &ex->bins,
                           &ex->spill, &ex->cursor, &ex->wcut, &ex->ccol})
        b->release();
    ex->a_val.release();
    ex->b_val.release();
    ex->cval.release();
    ex->block_sums.release();
    ex->info.release();
    ex->crpt.release();
    delete ex;
}

template <typename F>
int exec_create(ops_workspace *ws, ops_dtype dt, int use_graph, ops_exec **out, F &&fill)
{
    if (!ws || !out) return OPS_INVALID;
    if (dt != OPS_F64 && dt != OPS_F32) return OPS_INVALID;
    *out = nullptr;
    ops_exec *ex = new (std::nothrow) ops_exec;
    if (!ex) return OPS_OOM;
    ex->ws = ws;
    ex->dt = dt;
    ex->use_graph = use_graph != 0;
    const int st = guarded(ws, [&]() {
        CK(hipSetDevice(ws->device));
        init_attrs();
        fill(ex);
        return OPS_OK;
    });
    if (st != OPS_OK) {
        exec_release(ex);
        return st;
    }
    *out = ex;
    return OPS_OK;
}

}  // namespace

namespace opsamd {
const int *workspace_bins(const ops_workspace *ws) { return ws ? ws->bins.p : nullptr; }
}  // namespace opsamd

extern "C" {

const char *ops_version(void) { return "opsparse_amd 0.1 (gfx950)"; }

int ops_workspace_create(ops_workspace **out)
{
    if (!out) return OPS_INVALID;
    *out = nullptr;
    ops_workspace *ws = new (std::nothrow) ops_workspace();
    if (!ws) return OPS_OOM;
    const int st = guarded(ws, [&]() {
        CK(hipGetDevice(&ws->device));
        for (int b = 0; b < kNumBins; b++) {
            CK(hipStreamCreateWithFlags(&ws->pool[b], hipStreamNonBlocking));
            CK(hipEventCreateWithFlags(&ws->ev_join[b], hipEventDisableTiming));
            CK(hipEventCreate(&ws->ev_k0[b]));
            CK(hipEventCreate(&ws->ev_k1[b]));
        }
        CK(hipEventCreateWithFlags(&ws->ev_fork, hipEventDisableTiming));
        CK(hipStreamCreateWithFlags(&ws->cap, hipStreamNonBlocking));
        CK(hipMalloc(&ws->d_stats, sizeof(DevStats)));
        CK(hipHostMalloc(&ws->h_stats, sizeof(DevStats), hipHostMallocDefault));
        return OPS_OK;
    });
    if (st != OPS_OK) {
        ops_workspace_destroy(ws);
        return st;
    }
    *out = ws;
    return OPS_OK;
}

int ops_workspace_destroy(ops_workspace *ws)
{
    if (!ws) return OPS_INVALID;
    for (int b = 0; b < kNumBins; b++) {
        if (ws->pool[b]) (void)hipStreamDestroy(ws->pool[b]);
        if (ws->ev_join[b]) (void)hipEventDestroy(ws->ev_join[b]);
        if (ws->ev_k0[b]) (void)hipEventDestroy(ws->ev_k0[b]);
        if (ws->ev_k1[b]) (void)hipEventDestroy(ws->ev_k1[b]);
    }
    if (ws->ev_fork) (void)hipEventDestroy(ws->ev_fork);
    if (ws->cap) (void)hipStreamDestroy(ws->cap);
    if (ws->d_stats) (void)hipFree(ws->d_stats);
    if (ws->h_stats) (void)hipHostFree(ws->h_stats);
    ws->row_flop.release();
    ws->row_nnz.release();
    ws->bins.release();
    ws->cursor.release();
    ws->spill.release();
    ws->wcut.release();
    ws->kb.release();
    ws->crpt32_tmp.release();
    ws->block_sums.release();
    ws->crpt64.release();
    delete ws;
    return OPS_OK;
}

const char *ops_workspace_last_error(const ops_workspace *ws) { return ws ? ws->err.c_str() : "null workspace"; }

int ops_workspace_trim(ops_workspace *ws)
{
    if (!ws) return OPS_INVALID;
    return guarded(ws, [&]() {
        CK(hipDeviceSynchronize());
        ws->row_flop.release();
        ws->row_nnz.release();
        ws->bins.release();
        ws->cursor.release();
        ws->spill.release();
        ws->wcut.release();
        ws->kb.release();
        ws->crpt32_tmp.release();
        ws->block_sums.release();
        ws->crpt64.release();
        return OPS_OK;
    });
}

int ops_kept_bitmap_limit(int64_t bytes)
{
    std::lock_guard<std::mutex> g(g_kb_mu);
    g_kb_limit = bytes < 0 ? -1 : bytes;
    return OPS_OK;
}

int ops_plan_kept_rows(const ops_plan *plan, int64_t *rows)
{
    if (!plan || !rows) return OPS_INVALID;
    *rows = plan->core.kept_rows;
    return OPS_OK;
}

int ops_tuning_default(ops_tuning *t)
{
    if (!t) return OPS_INVALID;
    for (int i = 0; i < kNumBins - 1; i++) {
        t->sym_bound[i] = kSymBound[i];
        t->num_bound[i] = kNumBound[i];
    }
    t->short_hash = kHashLin;
    t->sym_bitmap = 0;
    for (int i = 0; i < kNumBins - 1; i++) t->sym_shape[i] = t->num_shape[i] = 0;
    return OPS_OK;
}

int ops_tuning_caps(ops_tuning *c)
{
    if (!c) return OPS_INVALID;
    for (int i = 0; i < kNumBins - 1; i++) {
        c->sym_bound[i] = kSymCap[i];
        c->num_bound[i] = kNumCap[i];
    }
    c->short_hash = kHashFib;
    c->sym_bitmap = 1;
    for (int i = 0; i < kNumBins - 1; i++) c->sym_shape[i] = c->num_shape[i] = i < 5 ? kShapes - 1 : 0;
    return OPS_OK;
}

int ops_workspace_set_tuning(ops_workspace *ws, const ops_tuning *t)
{
    if (!ws || !t) return OPS_INVALID;
    if (t->short_hash != kHashLin && t->short_hash != kHashFib) return OPS_INVALID;
    if (t->sym_bitmap != 0 && t->sym_bitmap != 1) return OPS_INVALID;
    for (int i = 0; i < kNumBins - 1; i++) {
        if (t->sym_bound[i] < 0 || t->sym_bound[i] > kSymCap[i]) return OPS_INVALID;
        if (t->num_bound[i] < 0 || t->num_bound[i] > kNumCap[i]) return OPS_INVALID;
        if (i > 0 && (t->sym_bound[i] < t->sym_bound[i - 1] || t->num_bound[i] < t->num_bound[i - 1]))
            return OPS_INVALID;
        const int top = i < 5 ? kShapes - 1 : 0;
        if (t->sym_shape[i] < 0 || t->sym_shape[i] > top || t->num_shape[i] < 0 || t->num_shape[i] > top)
            return OPS_INVALID;
    }
    for (int i = 0; i < kNumBins - 1; i++) {
        ws->bounds.sym[i] = t->sym_bound[i];
        ws->bounds.num[i] = t->num_bound[i];
    }
    ws->short_hash = t->short_hash;
    ws->sym_bitmap = t->sym_bitmap != 0;
    for (int i = 0; i < kNumBins - 1; i++) {
        ws->sym_shape[i] = t->sym_shape[i];
        ws->num_shape[i] = t->num_shape[i];
    }
    return OPS_OK;
}

int ops_workspace_get_tuning(const ops_workspace *ws, ops_tuning *t)
{
    if (!ws || !t) return OPS_INVALID;
    for (int i = 0; i < kNumBins - 1; i++) {
        t->sym_bound[i] = ws->bounds.sym[i];
        t->num_bound[i] = ws->bounds.num[i];
    }
    t->short_hash = ws->short_hash;
    t->sym_bitmap = ws->sym_bitmap ? 1 : 0;
    for (int i = 0; i < kNumBins - 1; i++) {
        t->sym_shape[i] = ws->sym_shape[i];
        t->num_shape[i] = ws->num_shape[i];
    }
    return OPS_OK;
}

int ops_csr_free(ops_csr *C)
{
    if (!C) return OPS_INVALID;
    if (C->rpt) (void)out_free(C->rpt);
    if (C->col) (void)out_free(C->col);
    if (C->val) (void)out_free(C->val);
    *C = ops_csr{};
    return OPS_OK;
}

int ops_csr64_free(ops_csr64 *C)
{
    if (!C) return OPS_INVALID;
    if (C->rpt) (void)out_free(C->rpt);
    if (C->col) (void)out_free(C->col);
    if (C->val) (void)out_free(C->val);
    *C = ops_csr64{};
    return OPS_OK;
}

int ops_spgemm_f64(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_csr *C, void *stream, ops_timings *t)
{
    if (!ws) return OPS_INVALID;
    return guarded(ws, [&]() {
        return spgemm_impl<double>(ws, A, B, C, nullptr, static_cast<hipStream_t>(stream), t);
    });
}

int ops_spgemm_f32(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_csr *C, void *stream, ops_timings *t)
{
    if (!ws) return OPS_INVALID;
    return guarded(ws, [&]() {
        return spgemm_impl<float>(ws, A, B, C, nullptr, static_cast<hipStream_t>(stream), t);
    });
}

int ops_spgemm64_f64(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_csr64 *C, void *stream, ops_timings *t)
{
    if (!ws) return OPS_INVALID;
    return guarded(ws, [&]() {
        return spgemm_impl<double>(ws, A, B, nullptr, C, static_cast<hipStream_t>(stream), t);
    });
}

int ops_spgemm64_f32(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_csr64 *C, void *stream, ops_timings *t)
{
    if (!ws) return OPS_INVALID;
    return guarded(ws, [&]() {
        return spgemm_impl<float>(ws, A, B, nullptr, C, static_cast<hipStream_t>(stream), t);
    });
}

int ops_spgemm64_noalloc(ops_dtype dt, const ops_csr *A, const ops_csr *B, ops_csr64 *C, int64_t capacity,
                         void *stream)
{
    if (dt != OPS_F64 && dt != OPS_F32) return OPS_INVALID;
    // Errors are reported through the status only (no workspace to hold a message).
    return guarded(nullptr, [&]() {
        const hipStream_t s = static_cast<hipStream_t>(stream);
        return dt == OPS_F64 ? spgemm_noalloc_impl<double>(A, B, C, capacity, s)
                             : spgemm_noalloc_impl<float>(A, B, C, capacity, s);
    });
}

int ops_spgemm_safe_f64(ops_workspace *ws, const ops_csr *A, const ops_csr *B, const int32_t *a_rpt_host,
                        const int32_t *a_col_host, const int32_t *b_rpt_host, ops_csr *C, void *stream,
                        ops_timings *t, int64_t *total_flop_out, int64_t *nnz_c_out, size_t *free_before,
                        size_t *free_after, size_t *total_bytes)
{
    if (total_flop_out) *total_flop_out = 0;
    if (nnz_c_out) *nnz_c_out = -1;
    if (!A || !B || !C) return OPS_INVALID;
    // Pre-flight on the host, before any GPU work (spgemm.cu:154-177).
    if (A->N != B->M) return OPS_NONSQUARE;
    if (!a_rpt_host || !a_col_host || !b_rpt_host) return OPS_INVALID;
    const int64_t flop = ops_compute_flop_host(A->M, a_rpt_host, a_col_host, b_rpt_host, nullptr);
    if (total_flop_out) *total_flop_out = flop;
    if (flop > static_cast<int64_t>(INT32_MAX)) return OPS_INT32_OVERFLOW;
    if (!ws) return OPS_INVALID;
    size_t fb = 0, tb = 0;
    if (hipMemGetInfo(&fb, &tb) != hipSuccess) fb = tb = 0;
    (void)hipGetLastError();
    if (free_before) *free_before = fb;
    if (total_bytes) *total_bytes = tb;
    const int st = guarded(ws, [&]() {
        return spgemm_impl<double>(ws, A, B, C, nullptr, static_cast<hipStream_t>(stream), t);
    });
    size_t fa = 0;
    if (hipMemGetInfo(&fa, &tb) != hipSuccess) fa = 0;
    (void)hipGetLastError();
    if (free_after) *free_after = fa;
    if (st == OPS_OK && nnz_c_out) *nnz_c_out = C->nnz;
    return st;
}

int ops_compute_flop_device(ops_workspace *ws, const ops_csr *A, const ops_csr *B, int32_t *row_flop,
                            int64_t *total_out, void *stream)
{
    if (!ws || !row_flop) return OPS_INVALID;
    return guarded(ws, [&]() {
        validate(A, B);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        ws->row_nnz.ensure(static_cast<size_t>(A->M) + 1);
        launch_zero_stats(ws->d_stats, s);
        launch_setup(as_in(A), as_in(B), row_flop, ws->row_nnz.p, ws->d_stats, ws->bounds, s);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(ws->h_stats, ws->d_stats, sizeof(DevStats), hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        if (total_out) *total_out = static_cast<int64_t>(ws->h_stats->total_flop);
        return OPS_OK;
    });
}

// ---- plans ----

int ops_plan_create(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_plan **plan, int64_t *nnz_c,
                    void *stream)
{
    if (!ws || !plan) return OPS_INVALID;
    *plan = nullptr;
    ops_plan *p = new (std::nothrow) ops_plan();
    if (!p) return OPS_OOM;
    p->ws = ws;
    const int st = guarded(ws, [&]() {
        validate(A, B);
        init_attrs();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const int M = A->M;
        p->crpt64.ensure(static_cast<size_t>(M) + 1);
        p->bins.ensure(static_cast<size_t>(M) + 1);
        symbolic_phase(ws, as_in(A), as_in(B), p->core, p->crpt64.p, nullptr, p->bins.p, s, nullptr, p->kb);
        numeric_scratch(p->core, A, B, p->cursor, p->spill, p->wcut, s);
        CK(hipStreamSynchronize(s));
        if (nnz_c) *nnz_c = p->core.nnz;
        return OPS_OK;
    });
    if (st != OPS_OK) {
        ops_plan_destroy(p);
        return st;
    }
    *plan = p;
    return OPS_OK;
}

int ops_plan_destroy(ops_plan *plan)
{
    if (!plan) return OPS_INVALID;
    if (plan->graph) (void)hipGraphExecDestroy(plan->graph);
    if (plan->graph_src) (void)hipGraphDestroy(plan->graph_src);
    plan->crpt64.release();
    plan->bins.release();
    plan->cursor.release();
    plan->spill.release();
    plan->wcut.release();
    plan->kb.release();
    delete plan;
    return OPS_OK;
}

int ops_plan_rowptr64(const ops_plan *plan, int64_t *rpt, void *stream)
{
    if (!plan || !rpt) return OPS_INVALID;
    return guarded(plan->ws, [&]() {
        CK(hipMemcpyAsync(rpt, plan->core.crpt64, sizeof(int64_t) * (static_cast<size_t>(plan->core.M) + 1),
                          hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
        return OPS_OK;
    });
}

int ops_plan_rowptr32(const ops_plan *plan, int32_t *rpt, void *stream)
{
    if (!plan || !rpt) return OPS_INVALID;
    if (plan->core.nnz > static_cast<long long>(INT32_MAX)) return OPS_INT32_OVERFLOW;
    return guarded(plan->ws, [&]() {
        launch_narrow_rpt(plan->core.crpt64, plan->core.M, rpt, static_cast<hipStream_t>(stream));
        CK(hipGetLastError());
        return OPS_OK;
    });
}

int ops_plan_numeric(ops_plan *plan, ops_dtype dt, const ops_csr *A, const ops_csr *B, int32_t *ccol, void *cval,
                     void *stream)
{
    if (!plan || !A || !B || (plan->core.nnz > 0 && (!ccol || !cval))) return OPS_INVALID;
    if (dt != OPS_F64 && dt != OPS_F32) return OPS_INVALID;
    if (A->M != plan->core.M || B->N != plan->core.N) return OPS_INVALID;
    return guarded(plan->ws, [&]() {
        const hipStream_t s = static_cast<hipStream_t>(stream);
        if (dt == OPS_F64)
            numeric_phase<double>(plan->ws, as_in(A), as_in(B), plan->core, ccol, static_cast<double *>(cval), s,
                                  nullptr);
        else
            numeric_phase<float>(plan->ws, as_in(A), as_in(B), plan->core, ccol, static_cast<float *>(cval), s,
                                 nullptr);
        return OPS_OK;
    });
}

int ops_spgemm_backward(ops_workspace *ws, ops_dtype dt, const ops_csr *A, const ops_csr *B, const int64_t *crpt64,
                        const int32_t *ccol, const void *dcval, void *da, void *db, void *stream)
{
    if (!ws || !A || !B) return OPS_INVALID;
    if (dt != OPS_F64 && dt != OPS_F32) return OPS_INVALID;
    if ((A->nnz > 0 && !da) || (B->nnz > 0 && !db)) return OPS_INVALID;
    return guarded(ws, [&]() {
        const hipStream_t s = static_cast<hipStream_t>(stream);
        if (dt == OPS_F64)
            backward_impl<double>(ws, A, B, crpt64, ccol, static_cast<const double *>(dcval), static_cast<double *>(da),
                                  static_cast<double *>(db), s);
        else
            backward_impl<float>(ws, A, B, crpt64, ccol, static_cast<const float *>(dcval), static_cast<float *>(da),
                                 static_cast<float *>(db), s);
        return OPS_OK;
    });
}

int ops_plan_numeric_graph(ops_plan *plan, ops_dtype dt, const ops_csr *A, const ops_csr *B, int32_t *ccol,
                           void *cval, void *stream)
{
    if (!plan || !A || !B || (plan->core.nnz > 0 && (!ccol || !cval))) return OPS_INVALID;
    if (dt != OPS_F64 && dt != OPS_F32) return OPS_INVALID;
    if (A->M != plan->core.M || B->N != plan->core.N) return OPS_INVALID;
    return guarded(plan->ws, [&]() {
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const void *args[8] = {A->rpt, A->col, A->val, B->rpt, B->col, B->val, ccol, cval};
        bool same = plan->graph && plan->g_dt == static_cast<int>(dt);
        for (int i = 0; i < 8 && same; i++) same = plan->g_args[i] == args[i];
        if (!same) {
            if (plan->graph) CK(hipGraphExecDestroy(plan->graph));
            if (plan->graph_src) CK(hipGraphDestroy(plan->graph_src));
            plan->graph = nullptr;
            plan->graph_src = nullptr;
            // Capture on a private stream so the caller's stream may be the null stream.
            hipStream_t cs = plan->ws->cap;
            CK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            try {
                if (dt == OPS_F64)
                    numeric_phase<double>(plan->ws, as_in(A), as_in(B), plan->core, ccol, static_cast<double *>(cval),
                                          cs, nullptr);
                else
                    numeric_phase<float>(plan->ws, as_in(A), as_in(B), plan->core, ccol, static_cast<float *>(cval),
                                         cs, nullptr);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(cs, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            CK(hipStreamEndCapture(cs, &plan->graph_src));
            CK(hipGraphInstantiate(&plan->graph, plan->graph_src, nullptr, nullptr, 0));
            for (int i = 0; i < 8; i++) plan->g_args[i] = args[i];
            plan->g_dt = static_cast<int>(dt);
        }
        CK(hipGraphLaunch(plan->graph, s));
        return OPS_OK;
    });
}

int ops_exec_create(ops_workspace *ws, ops_dtype dt, const ops_csr *A, const ops_csr *B, int64_t capacity,
                    int use_graph, ops_exec **out)
{
    if (!A || !B) return OPS_INVALID;
    return exec_create(ws, dt, use_graph, out, [&](ops_exec *ex) {
        validate(A, B);
        ex->A = *A;
        ex->B = *B;
        exec_finish(ex, capacity, ws->cap);
    });
}

int ops_exec_create_batch(ops_workspace *ws, ops_dtype dt, int32_t count, const ops_csr *A, const ops_csr *B,
                          int64_t capacity, int use_graph, int64_t *row_base, ops_exec **out)
{
    if (count < 0 || (count > 0 && (!A || !B))) return OPS_INVALID;
    return exec_create(ws, dt, use_graph, out, [&](ops_exec *ex) {
        const hipStream_t s = ws->cap;
        for (int i = 0; i < count; i++)
            if (A[i].N != B[i].M) throw HipFail(OPS_NONSQUARE, "instance with A_i.N != B_i.M");
        // A_i block-diagonal (A_i's columns shifted to B_i's rows), B_i stacked
        int st = pack_side_prepare(count, A, ex->pa, row_base, nullptr, s);
        if (st == OPS_OK) st = pack_side_prepare(count, B, ex->pb, nullptr, nullptr, s, true);
        if (st != OPS_OK) throw HipFail(st, "batch instances");
        ex->batch = true;
        ex->count = count;
        const size_t wv = dt == OPS_F64 ? 1 : 2;  // values per double of storage
        ex->a_rpt.ensure(static_cast<size_t>(ex->pa.rows) + 1);
        ex->a_col.ensure(std::max<size_t>(ex->pa.nnz, 1));
        ex->a_val.ensure(std::max<size_t>((ex->pa.nnz + wv - 1) / wv, 1));
        ex->b_rpt.ensure(static_cast<size_t>(ex->pb.rows) + 1);
        ex->b_col.ensure(std::max<size_t>(ex->pb.nnz, 1));
        ex->b_val.ensure(std::max<size_t>((ex->pb.nnz + wv - 1) / wv, 1));
        ex->A = ops_csr{ex->pa.rows, ex->pa.cols, ex->pa.nnz, ex->a_rpt.p, ex->a_col.p, ex->a_val.p};
        ex->B = ops_csr{ex->pb.rows, ex->pb.cols, ex->pb.nnz, ex->b_rpt.p, ex->b_col.p, ex->b_val.p};
        validate(&ex->A, &ex->B);
        exec_finish(ex, capacity, s);
    });
}

int ops_exec_launch(ops_exec *ex, void *stream)
{
    if (!ex) return OPS_INVALID;
    return guarded(ex->ws, [&]() {
        const hipStream_t s = static_cast<hipStream_t>(stream);
        if (ex->exec)
            CK(hipGraphLaunch(ex->exec, s));
        else
            exec_sequence(ex, s);
        return OPS_OK;
    });
}

int ops_exec_result(ops_exec *ex, ops_csr64 *C, int64_t *nnz_c, void *stream)
{
    if (!ex) return OPS_INVALID;
    return guarded(ex->ws, [&]() {
        long long nnz = 0;
        bool over = false;
        exec_info(ex, nnz, over, static_cast<hipStream_t>(stream));
        if (nnz_c) *nnz_c = nnz;
        if (C) {
            *C = ops_csr64{};
            C->M = ex->A.M;
            C->N = ex->B.N;
            C->nnz = nnz;
            C->rpt = reinterpret_cast<int64_t *>(ex->crpt.p);
            C->col = ex->ccol.p;
            C->val = ex->cval.p;
        }
        return over ? OPS_INVALID : OPS_OK;
    });
}

int ops_exec_destroy(ops_exec *ex)
{
    if (!ex) return OPS_INVALID;
    (void)hipDeviceSynchronize();
    exec_release(ex);
    return OPS_OK;
}

}  // extern "C"
