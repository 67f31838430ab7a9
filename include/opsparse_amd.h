/*
 * opsparse_amd: C ABI of the MI355X (gfx950) SpGEMM library, C = A * B in CSR.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never
 * throws, and returns a status code.  Device arrays are HIP device pointers;
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *
 * The reference (hongyx11/opsparse) exposes C++ entry points only; each
 * function below names the reference interface it replaces.  The C++
 * drop-in (namespace opsparse, include/opsparse/...) is built on top of it.
 */
#ifndef OPSPARSE_AMD_H_
#define OPSPARSE_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  0-5 keep the values of the reference's SpgemmStatus
 * (include/opsparse/profile/perf.hpp:16-23); 6 is new (bad arguments). */
enum {
    OPS_OK = 0,
    OPS_NONSQUARE = 1,      /* A.N != B.M */
    OPS_INT32_OVERFLOW = 2, /* total products or nnz(C) above INT32_MAX on an int32 output */
    OPS_OOM = 3,            /* hipMalloc failed */
    OPS_HIP_ERROR = 4,      /* any other HIP runtime error */
    OPS_EXCEPTION = 5,      /* internal exception caught at the boundary */
    OPS_INVALID = 6         /* invalid argument (null pointer, negative size, bad dtype) */
};

typedef enum { OPS_F64 = 0, OPS_F32 = 1 } ops_dtype;

/* Device CSR, int32 offsets: the device mirror of opsparse::CSR
 * (include/opsparse/core_ds/csr.hpp:12-21: M, N, nnz, d_rpt, d_col, d_val).
 * rpt[r]..rpt[r+1] index col/val directly, so rpt need not start at 0
 * (a row slice of a larger CSR is a valid descriptor). */
typedef struct ops_csr {
    int32_t M;
    int32_t N;
    int32_t nnz;
    int32_t *rpt; /* M+1 */
    int32_t *col; /* nnz, strictly ascending per row */
    void *val;    /* nnz, double (OPS_F64) or float (OPS_F32) */
} ops_csr;

/* Output CSR with 64-bit row offsets, for products whose nnz(C) exceeds
 * INT32_MAX (the reference refuses these: spgemm.cu:173-177, 210-213). */
typedef struct ops_csr64 {
    int64_t M;
    int64_t N;
    int64_t nnz;
    int64_t *rpt; /* M+1 */
    int32_t *col;
    void *val;
} ops_csr64;

#define OPS_NUM_BINS 8

/* Per-phase timings in seconds (the fields of opsparse::SpgemmTimings,
 * include/opsparse/profile/spgemm_timings.hpp:8-18), plus per-bin kernel
 * statistics used for the roofline report. */
typedef struct ops_timings {
    int measure_separate; /* in: 1 = synchronise after every phase and time it on the host clock */
    int kernel_events;    /* in: 1 = time every per-bin kernel with HIP events on its own stream */
    double setup;
    double symbolic_binning;
    double symbolic;
    double reduce;
    double numeric_binning;
    double prefix;
    double allocate;
    double numeric;
    double cleanup;
    double total;
    /* out: problem statistics */
    int64_t total_flop; /* intermediate products P */
    int64_t nnz_c;      /* nnz(C) */
    int32_t max_row_flop;
    int32_t max_row_nnz;
    /* out: per-bin rows, products and event-timed kernel milliseconds
     * (ms are 0 unless kernel_events was set) */
    int64_t sym_rows[OPS_NUM_BINS];
    int64_t sym_flop[OPS_NUM_BINS];
    int64_t sym_arows_nnz[OPS_NUM_BINS]; /* nnz of A over the rows of the bin */
    double sym_ms[OPS_NUM_BINS];
    int64_t num_rows[OPS_NUM_BINS];
    int64_t num_flop[OPS_NUM_BINS];
    int64_t num_nnz[OPS_NUM_BINS];
    int64_t num_arows_nnz[OPS_NUM_BINS]; /* nnz of A over the rows of the bin */
    double num_ms[OPS_NUM_BINS];
    /* out: long rows whose symbolic column bitmap was kept for the numeric
     * phase (see ops_kept_bitmap_limit) */
    int64_t kept_rows;
} ops_timings;

/* Persistent per-device workspace: stream pool, events, growable device
 * buffers.  Replaces opsparse::SpgemmMeta (include/opsparse/core_algo/spgemm/meta.hpp:15-60),
 * which the reference allocates and frees inside every call (meta.cu:16-65). */
typedef struct ops_workspace ops_workspace;

int ops_workspace_create(ops_workspace **ws);
int ops_workspace_destroy(ops_workspace *ws);
/* Last error message recorded on this workspace (never NULL). */
const char *ops_workspace_last_error(const ops_workspace *ws);
/* Frees the workspace's growable device buffers (bin lists, scan sums,
 * cursors, window-cut table, kept column bitmaps); streams and events stay,
 * and the next call allocates what it needs again.  The counterpart of the
 * per-call SpgemmMeta::release (meta.cu:50-65) for callers that want the
 * memory back between products. */
int ops_workspace_trim(ops_workspace *ws);

/* Long-row column bitmaps.  For rows above the numeric bin-6 bound, the
 * symbolic phase can keep each row's column bitmap (ceil(B.N / 32) words,
 * 128 KB at 2^20 columns) so the numeric phase skips a second walk over the
 * row's products.  A workspace keeps them from one call to the next and a
 * plan until it is destroyed.  Every workspace and plan of the process on
 * one device shares one limit: `bytes` (< 0, the default: min(64 GiB, a
 * quarter of the device's memory); 0: keep none).  Each holder also takes
 * at most a quarter of the free memory, and when the allocation fails the
 * rows are marked again instead (never an error). */
int ops_kept_bitmap_limit(int64_t bytes);

/*
 * C = A * B.  Replaces opsparse::opsparse_spgemm / opsparse_spgemm_reg
 * (include/opsparse/core_algo/spgemm/spgemm.hpp:12-17; src/opsparse/core_algo/spgemm/spgemm.cu:16-126).
 * The library allocates C->rpt, C->col and C->val (from its output cache,
 * below; release them with ops_csr_free, the counterpart of CSR::release,
 * csr.cu:40-56) and sets
 * C->M, C->N, C->nnz.  Returns when C is complete on `stream`.  A and B must
 * have sorted, duplicate-free rows; A->N must equal B->M.
 */
int ops_spgemm_f64(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_csr *C, void *stream, ops_timings *t);
int ops_spgemm_f32(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_csr *C, void *stream, ops_timings *t);

/* C = A * B with 64-bit row offsets for C (no counterpart in the reference,
 * which returns int32_overflow instead: spgemm.cu:163-177). */
int ops_spgemm64_f64(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_csr64 *C, void *stream, ops_timings *t);
int ops_spgemm64_f32(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_csr64 *C, void *stream, ops_timings *t);

/* C = A * B into caller-owned device arrays, allocating no device memory:
 * no hipMalloc, hipFree or allocating library routine.  No counterpart in
 * the reference; for callers that cannot allocate (serving loops with
 * preallocated buffers, harnesses that forbid allocation; it synchronises
 * the host, so it cannot itself be captured in a graph).  Variant of
 * ops_spgemm64_f64 / _f32 (dt selects the value
 * type) that takes no workspace: its scratch is a fixed array of about
 * 21 MB inside the library's code object on each GPU, held until the
 * process ends.  Limits: A has at most 262,144 rows and its last entry
 * index is below 1,048,576 (A->rpt[M] <= 2^20); larger inputs return
 * OPS_INVALID before any kernel runs.  The caller passes C->rpt (M+1 int64),
 * C->col and C->val with room for `capacity` entries; on return C->M, C->N
 * and C->nnz are set.  When nnz(C) > capacity the call returns OPS_INVALID
 * with C->nnz = nnz(C), C->rpt filled and C->col / C->val untouched
 * (retry with larger arrays).  All bins run on `stream`; the call synchronises the host
 * twice and returns when C is complete.  Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (from several host threads
 * or streams) race and must be serialised by the caller.  Long rows take
 * the cursor path (no kept column bitmaps, no window-cut table). */
int ops_spgemm64_noalloc(ops_dtype dt, const ops_csr *A, const ops_csr *B, ops_csr64 *C, int64_t capacity,
                         void *stream);

/* Releases library-allocated C arrays and zeroes the descriptor (CSR::drelease, csr.cu:40-48). */
int ops_csr_free(ops_csr *C);
int ops_csr64_free(ops_csr64 *C);

/*
 * Output memory (no counterpart in the reference, which calls cudaMalloc for
 * C on every product: spgemm.cu:96-114).  C's arrays and packed batches come
 * from a per-device cache of blocks released earlier through ops_free,
 * ops_csr_free or ops_csr64_free; only a miss calls hipMalloc, whose fresh
 * VRAM the driver clears at a cost comparable to a large product.
 * Releasing synchronises the device, as hipFree does.  Library-allocated
 * arrays must be released through these calls, not hipFree; ops_free passes
 * any other pointer to hipFree.  The cache keeps at most `bytes` per device
 * (bytes < 0, the default: half of the device's memory; 0: no caching),
 * freeing the least recently released blocks to make room for a new one.
 */
typedef struct {
    int64_t cached_bytes, cached_blocks;  /* held by the cache */
    int64_t live_bytes, live_blocks;      /* handed out, not yet released */
    int64_t hits, misses;                 /* allocations served from the cache / by hipMalloc */
    int64_t limit;                        /* -1: default cap */
} ops_mem_stats;

int ops_free(void *p);
int ops_mem_cache_set_limit(int64_t bytes);
int ops_mem_cache_release(void);
int ops_mem_cache_stats(ops_mem_stats *out);

/*
 * Safe variant: never throws, classifies failures.  Replaces
 * opsparse::opsparse_spgemm_safe (spgemm.cu:148-219).  Reads the HOST arrays
 * a_rpt_host, a_col_host and b_rpt_host for the int32 pre-check exactly as
 * the reference does (compute_flop.cu:59-65); returns OPS_NONSQUARE or
 * OPS_INT32_OVERFLOW before any GPU work.  total_flop_out / nnz_c_out may be
 * NULL; free_before / free_after receive hipMemGetInfo snapshots.
 */
int ops_spgemm_safe_f64(ops_workspace *ws, const ops_csr *A, const ops_csr *B, const int32_t *a_rpt_host,
                        const int32_t *a_col_host, const int32_t *b_rpt_host, ops_csr *C, void *stream,
                        ops_timings *t, int64_t *total_flop_out, int64_t *nnz_c_out, size_t *free_before,
                        size_t *free_after, size_t *total_bytes);

/* Host product count with OpenMP.  Replaces opsparse::compute_flop
 * (include/opsparse/numeric/compute_flop.hpp:11-13; compute_flop.cu:33-65);
 * row_flop_out (M entries) may be NULL.  Returns the 64-bit total. */
int64_t ops_compute_flop_host(int32_t M, const int32_t *a_rpt, const int32_t *a_col, const int32_t *b_rpt,
                              int64_t *row_flop_out);

/* Device product count: row_flop[M] (int32, saturated) and the 64-bit total.
 * Replaces k_compute_flop (compute_flop.cu:9-31). */
int ops_compute_flop_device(ops_workspace *ws, const ops_csr *A, const ops_csr *B, int32_t *row_flop,
                            int64_t *total_out, void *stream);

/* ---- Structure reuse: plan once, run the numeric phase many times ------- */

/* A plan holds C's row offsets and the numeric row bins for fixed patterns
 * of A and B (no counterpart in the reference).  ops_plan_create runs the
 * setup, symbolic and scan phases and reports nnz(C).  A plan holds device
 * memory until it is destroyed: C's row offsets, the bin lists, the
 * long-row scratch and the kept column bitmaps of its long rows (within
 * ops_kept_bitmap_limit). */
typedef struct ops_plan ops_plan;

int ops_plan_create(ops_workspace *ws, const ops_csr *A, const ops_csr *B, ops_plan **plan, int64_t *nnz_c,
                    void *stream);
int ops_plan_destroy(ops_plan *plan);
/* Long rows whose column bitmap the plan keeps. */
int ops_plan_kept_rows(const ops_plan *plan, int64_t *rows);
/* Copies C's row offsets (M+1) into caller device memory, int32 or int64. */
int ops_plan_rowptr32(const ops_plan *plan, int32_t *rpt, void *stream);
int ops_plan_rowptr64(const ops_plan *plan, int64_t *rpt, void *stream);
/* Numeric phase into caller-owned device buffers ccol[nnz_c], cval[nnz_c]:
 * no allocation and no host synchronisation, so it can be captured in a
 * hipGraph.  A and B must have the patterns the plan was made with. */
int ops_plan_numeric(ops_plan *plan, ops_dtype dt, const ops_csr *A, const ops_csr *B, int32_t *ccol, void *cval,
                     void *stream);
/* Captures ops_plan_numeric into a hipGraph on first use and replays it. */
int ops_plan_numeric_graph(ops_plan *plan, ops_dtype dt, const ops_csr *A, const ops_csr *B, int32_t *ccol,
                           void *cval, void *stream);

/* ---- Whole calls replayed from a hipGraph (no counterpart in the reference)
 *
 * An executor runs the whole product - setup, binning, symbolic, scan and
 * numeric - with bin sizes, bin offsets and nnz(C) kept on the device: no
 * host synchronisation, every per-bin kernel launched for every bin (each
 * reads its bin's size on the device and loops over its rows).  The
 * sequence is captured once into a hipGraph (use_graph = 1; 0 issues it
 * again at every launch) and each ops_exec_launch replays it.  The
 * operands' descriptors (pointers, M, N, nnz) are fixed at creation; their
 * row offsets, columns and values may change between launches (the entry
 * counts staying the same, and B's entries within the index range
 * [rpt[0], rpt[M]) they span at creation).  C lives in the executor: 64-bit row offsets
 * and room for `capacity` entries (capacity <= 0: exactly nnz(C) of the
 * operands at creation, measured by one untimed run).  The workspace lends
 * its streams and tuning (taken at creation) and must outlive the executor.
 * Long rows take the cursor path (no kept column bitmaps). */
typedef struct ops_exec ops_exec;
int ops_exec_create(ops_workspace *ws, ops_dtype dt, const ops_csr *A, const ops_csr *B, int64_t capacity,
                    int use_graph, ops_exec **ex);
/* Batched: C_i = A_i * B_i for `count` independent instances (host arrays
 * of device descriptors, as for ops_batch_pack; A_i.N == B_i.M).  Each
 * launch packs the A_i block-diagonally and stacks the B_i (their columns
 * unshifted) with kernels over instance tables built here, then multiplies:
 * C_i is rows [row_base[i], row_base[i+1]) of the executor's C, in B_i's
 * own column indices.  row_base (count + 1 host entries, may be NULL)
 * receives the A_i row bases. */
int ops_exec_create_batch(ops_workspace *ws, ops_dtype dt, int32_t count, const ops_csr *A, const ops_csr *B,
                          int64_t capacity, int use_graph, int64_t *row_base, ops_exec **ex);
/* Enqueues one whole call on `stream` (returns at once). */
int ops_exec_launch(ops_exec *ex, void *stream);
/* Waits for `stream`, then describes the last launch's C (arrays owned by
 * the executor, valid until the next launch or ops_exec_destroy) and its
 * nnz.  Returns OPS_INVALID when nnz(C) exceeded the capacity: C->nnz and
 * the row offsets are then set, no entry was written. */
int ops_exec_result(ops_exec *ex, ops_csr64 *C, int64_t *nnz_c, void *stream);
int ops_exec_destroy(ops_exec *ex);

/* ---- Tuning: bin boundaries and the short-row hash family ---------------- */

/* The reference hard-codes its bin bounds (kernels/binning.cuh:21, 51, 105,
 * 143) and hash (utils/config.hpp:32); here they are run-time settings of a
 * workspace, chosen by the autotuner (opsparse_amd/autotune.py).  A bound may
 * not exceed the capacity of its bin's kernel (ops_tuning_caps) and bounds
 * must not decrease from bin to bin; bin 7 is unbounded. */
typedef struct ops_tuning {
    int32_t sym_bound[OPS_NUM_BINS - 1]; /* inclusive upper bound of row products per symbolic bin */
    int32_t num_bound[OPS_NUM_BINS - 1]; /* inclusive upper bound of row nnz per numeric bin */
    int32_t short_hash;                  /* bins 0-4: 0 = linear (col * 107 mod T), 1 = Fibonacci */
    int32_t sym_bitmap;                  /* symbolic bins 0-4: 1 = a row whose products span fewer than
                                            32*T columns sets bits in its T-word table and counts them
                                            instead of hashing; 0 = always hash */
    /* Kernel shape of bins 0-4 (0 = default; bins 5 and 6 take 0 only):
     * lanes per row G, table slots T and block size BS.  No shape lowers a
     * bin's capacity, so any shape goes with any valid bounds.
     *   symbolic  bin: shape 0        shape 1        shape 2
     *             0:   16/32/256      16/64/256      32/64/256
     *             1:   32/64/256      32/128/256     64/128/256
     *             2:   64/128/256     64/256/256     32/128/256
     *             3:   64/512/256     64/1024/256    64/512/64
     *             4:   64/1024/256    64/2048/256    64/1024/64
     *   numeric   bin: shape 0        shape 1        shape 2       (G/T/BS)
     *             0:   16/32/256      16/64/256      16/32/64
     *             1:   32/64/256      32/128/256     16/64/256
     *             2:   64/128/256     64/256/256     32/128/256
     *             3:   32/256/64      64/256/64      32/512/64
     *             4:   64/512/64      64/1024/64     64/512/256 */
    int32_t sym_shape[OPS_NUM_BINS - 1];
    int32_t num_shape[OPS_NUM_BINS - 1];
} ops_tuning;

int ops_tuning_default(ops_tuning *t);
int ops_tuning_caps(ops_tuning *caps);
/* Validates and installs; later calls and plans on this workspace use it. */
int ops_workspace_set_tuning(ops_workspace *ws, const ops_tuning *t);
int ops_workspace_get_tuning(const ops_workspace *ws, ops_tuning *t);

/* ---- Batches of independent products (no counterpart in the reference) --- */

/* Packs `count` device CSR matrices block-diagonally into one library-
 * allocated CSR (free with ops_csr_free): instance i occupies rows
 * [row_base[i], row_base[i+1]) and columns [col_base[i], col_base[i+1]).
 * Packing the A_i and the B_i of independent products C_i = A_i * B_i gives
 * packed A and B whose product is the block-diagonal of the C_i, so one plan
 * (one setup, one set of per-bin kernels, bins shared across instances)
 * computes every C_i; C_i is rows [rowA_base[i], rowA_base[i+1]) of the
 * packed C with columns shifted by colB_base[i].  `instances` is a host
 * array; each descriptor's nnz must equal rpt[M] - rpt[0].  row_base and
 * col_base (count + 1 entries each, host) may be NULL. */
int ops_batch_pack(int32_t count, const ops_csr *instances, ops_dtype dt, ops_csr *packed, int64_t *row_base,
                   int64_t *col_base, void *stream);

/* ---- Masked backward at fixed sparsity (no counterpart in the reference) -- */

/* Given dC on pattern(C) (ccol sorted per row, row offsets crpt64),
 * dA = (dC * B^T) sampled on pattern(A), dB = (A^T * dC) sampled on
 * pattern(B).  da and db are overwritten and are indexed from the first
 * entry of each descriptor: da[e - A->rpt[0]] is the gradient of A entry e
 * (A->rpt[M] - A->rpt[0] values), db[e - B->rpt[0]] that of B entry e, so a
 * row-slice descriptor's gradients fill a buffer of the slice's nnz. */
int ops_spgemm_backward(ops_workspace *ws, ops_dtype dt, const ops_csr *A, const ops_csr *B, const int64_t *crpt64,
                        const int32_t *ccol, const void *dcval, void *da, void *db, void *stream);

/* ---- Synthetic inputs (host, deterministic across machines) ------------- */

/* Host CSR produced by the generators; free with ops_host_csr_free. */
typedef struct ops_host_csr {
    int32_t M, N;
    int64_t nnz;
    int32_t *rpt;
    int32_t *col;
    double *val;
} ops_host_csr;

/* M x N, exactly d distinct uniform columns per row, values U[-1,1]. */
int ops_gen_uniform_rows(int32_t M, int32_t N, int32_t d, uint64_t seed, ops_host_csr *out);
/* 27-point stencil on nx*ny*nz (x fastest), diagonal `diag`, off-diagonal
 * `off` + U[-noise, noise]. */
int ops_gen_stencil27(int32_t nx, int32_t ny, int32_t nz, double diag, double off, double noise, uint64_t seed,
                      ops_host_csr *out);
/* R-MAT scale s (2^s x 2^s), edge_factor * 2^s edges drawn with (a,b,c,d),
 * duplicates merged, no vertex permutation; values U[-1,1] (normal == 0) or
 * N(0,1) (normal == 1) drawn after merging. */
int ops_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed, int normal,
                 ops_host_csr *out);
/* Batch of `count` instances packed block-diagonally: instance i is n_i x n_i
 * with n_i = floor(exp(U[ln n_lo, ln n_hi])), d_i uniform in [d_lo, d_hi]
 * distinct uniform columns per row, values U[-1,1]; A and B are drawn
 * independently.  row_base (count+1, host memory supplied by the caller)
 * receives the instance offsets; d_out (count) the per-instance d. */
int ops_gen_batch(int32_t count, int32_t n_lo, int32_t n_hi, int32_t d_lo, int32_t d_hi, uint64_t seed,
                  ops_host_csr *A, ops_host_csr *B, int64_t *row_base, int32_t *d_out);
int ops_host_csr_free(ops_host_csr *m);

/* Library version string. */
const char *ops_version(void);

#ifdef __cplusplus
}
#endif

#endif /* OPSPARSE_AMD_H_ */
