/*
 * CPU oracle for C = A*B in CSR (test infrastructure only).
 *
 * This file is the checker for the HIP SpGEMM library.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark load it; the
 * library itself never links or calls it.
 *
 * What it restates from the reference (hongyx11/opsparse):
 *   - CSR layout: zero-based, int32 row pointers and column indices, columns
 *     strictly ascending per row, FP64 values
 *     (include/opsparse/core_ds/csr.hpp:10-37; include/opsparse/system/cuda_common.hpp:13-14).
 *   - Product count per row ("flop"): row_flop[i] = sum over A(i,k) != 0 of
 *     nnz(B row k), total returned as a 64-bit integer
 *     (src/opsparse/numeric/compute_flop.cu:33-52).
 *   - Output semantics of the pipeline: C's pattern is structural (entries
 *     that cancel to 0.0 are kept; the symbolic count runs over patterns,
 *     include/opsparse/core_algo/spgemm/kernels/symbolic.cuh:36-67), every
 *     output row is sorted by column (numeric.cuh:97-106, 194-204), and
 *     row_ptr is the exclusive sum of row nnz (spgemm.cu:49).
 *   - Masked backward (no reference counterpart): for every product
 *     (i,k,j) with A(i,k) and B(k,j) present, dA(i,k) += dC(i,j) * B(k,j)
 *     and dB(k,j) += A(i,k) * dC(i,j).
 *
 * The multiply itself is the textbook row-wise algorithm of Gustavson
 * (ACM TOMS 4(3), 1978): a dense accumulator plus a marker array per
 * thread, then a sort of the touched columns.  It is not in the reference.
 *
 * Every function also accumulates sum_k |A(i,k) * B(k,j)| per output entry
 * when asked, which bounds the rounding error of any summation order and is
 * what the parity tests normalise value errors by.
 *
 * Parity pins (see tests/test_oracle.py): scipy.sparse products on
 * positive-valued inputs (structure and values), a dense triple loop on tiny
 * inputs, identity / diagonal / permutation special cases, and the closed
 * forms of the 4096x4096 8-per-row workload (262,144 products) and the
 * 27-point stencil on a 100^3 grid (704,969,000 products, 120,553,784
 * output nonzeros).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Restates compute_flop.cu:33-52: per-row product count, 64-bit total. */
i64 oracle_compute_flop(int M, const int *arpt, const int *acol, const int *brpt, i64 *row_flop)
{
    i64 total = 0;
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int i = 0; i < M; i++) {
        i64 s = 0;
        for (int j = arpt[i]; j < arpt[i + 1]; j++) s += (i64)(brpt[acol[j] + 1] - brpt[acol[j]]);
        if (row_flop) row_flop[i] = s;
        total += s;
    }
    return total;
}

/*
 * Symbolic phase: crpt[0..M] (int64) = exclusive sum of |union of B rows|.
 * Returns nnz(C).  Marker array per thread, size N.
 */
i64 oracle_symbolic(int M, int N, const int *arpt, const int *acol, const int *brpt, const int *bcol, i64 *crpt)
{
    crpt[0] = 0;
#pragma omp parallel
    {
        int *mark = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
        for (int c = 0; c < N; c++) mark[c] = -1;
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < M; i++) {
            i64 cnt = 0;
            for (int j = arpt[i]; j < arpt[i + 1]; j++) {
                int k = acol[j];
                for (int l = brpt[k]; l < brpt[k + 1]; l++) {
                    int c = bcol[l];
                    if (mark[c] != i) {
                        mark[c] = i;
                        cnt++;
                    }
                }
            }
            crpt[i + 1] = cnt;
        }
        free(mark);
    }
    for (int i = 0; i < M; i++) crpt[i + 1] += crpt[i];
    return crpt[M];
}

/*
 * Numeric phase, FP64 accumulation.  aval/bval are double (is_f32 == 0) or
 * float (is_f32 == 1); the products and sums are always formed in double.
 * Writes sorted columns into ccol and values into cval (double); cabs, if
 * non-NULL, receives sum |a*b| per entry.  crpt must come from
 * oracle_symbolic on the same patterns.
 */
void oracle_numeric(int M, int N, const int *arpt, const int *acol, const void *aval, const int *brpt, const int *bcol,
                    const void *bval, int is_f32, const i64 *crpt, int *ccol, double *cval, double *cabs)
{
#pragma omp parallel
    {
        double *acc = (double *)calloc((size_t)(N > 0 ? N : 1), sizeof(double));
        double *aab = cabs ? (double *)calloc((size_t)(N > 0 ? N : 1), sizeof(double)) : NULL;
        int *mark = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
        for (int c = 0; c < N; c++) mark[c] = -1;
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < M; i++) {
            i64 base = crpt[i];
            i64 n = 0;
            for (int j = arpt[i]; j < arpt[i + 1]; j++) {
                int k = acol[j];
                double a = is_f32 ? (double)((const float *)aval)[j] : ((const double *)aval)[j];
                for (int l = brpt[k]; l < brpt[k + 1]; l++) {
                    int c = bcol[l];
                    double b = is_f32 ? (double)((const float *)bval)[l] : ((const double *)bval)[l];
                    if (mark[c] != i) {
                        mark[c] = i;
                        ccol[base + n++] = c;
                        acc[c] = 0.0;
                        if (aab) aab[c] = 0.0;
                    }
                    acc[c] += a * b;
                    if (aab) aab[c] += a * b < 0 ? -(a * b) : a * b;
                }
            }
            qsort(ccol + base, (size_t)n, sizeof(int), cmp_int);
            for (i64 t = 0; t < n; t++) {
                int c = ccol[base + t];
                cval[base + t] = acc[c];
                if (cabs) cabs[base + t] = aab[c];
            }
        }
        free(acc);
        free(aab);
        free(mark);
    }
}

/*
 * Symbolic + numeric for a contiguous block of rows [r0, r1) only: used to
 * check large products on a sample of rows.  crpt_local has r1-r0+1 entries
 * (starting at 0).  Returns nnz of the block, or -1 when cap is too small
 * (call with ccol == NULL to size first).
 */
i64 oracle_rows(int r0, int r1, int N, const int *arpt, const int *acol, const void *aval, const int *brpt,
                const int *bcol, const void *bval, int is_f32, i64 *crpt_local, int *ccol, double *cval, double *cabs,
                i64 cap)
{
    int M = r1 - r0;
    i64 nnz = oracle_symbolic(M, N, arpt + r0, acol, brpt, bcol, crpt_local);
    if (ccol == NULL) return nnz;
    if (nnz > cap) return -1;
    oracle_numeric(M, N, arpt + r0, acol, aval, brpt, bcol, bval, is_f32, crpt_local, ccol, cval, cabs);
    return nnz;
}

/*
 * Masked backward in FP64: given the forward patterns (A, B, C) and dC on
 * pattern(C), dA(i,k) = sum_j dC(i,j) * B(k,j) over j in pattern(B row k),
 * dB(k,j) = sum_i A(i,k) * dC(i,j).  dC(i,j) is found by binary search in
 * the sorted C row.  dB is accumulated with atomic adds across threads, so
 * its summation order varies (FP64, far below any tolerance used).
 */
void oracle_backward(int M, const int *arpt, const int *acol, const void *aval, const int *brpt, const int *bcol,
                     const void *bval, int is_f32, const i64 *crpt, const int *ccol, const void *dcval,
                     double *da, double *db, int nnzB)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < M; i++) {
        const int *row = ccol + crpt[i];
        i64 n = crpt[i + 1] - crpt[i];
        for (int j = arpt[i]; j < arpt[i + 1]; j++) {
            int k = acol[j];
            double s = 0.0;
            for (int l = brpt[k]; l < brpt[k + 1]; l++) {
                int c = bcol[l];
                i64 lo = 0, hi = n;
                while (lo < hi) {
                    i64 mid = (lo + hi) >> 1;
                    if (row[mid] < c)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                double g = is_f32 ? (double)((const float *)dcval)[crpt[i] + lo] : ((const double *)dcval)[crpt[i] + lo];
                double b = is_f32 ? (double)((const float *)bval)[l] : ((const double *)bval)[l];
                s += g * b;
            }
            da[j] = s;
        }
    }
    for (int l = 0; l < nnzB; l++) db[l] = 0.0;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < M; i++) {
        const int *row = ccol + crpt[i];
        i64 n = crpt[i + 1] - crpt[i];
        for (int j = arpt[i]; j < arpt[i + 1]; j++) {
            int k = acol[j];
            double a = is_f32 ? (double)((const float *)aval)[j] : ((const double *)aval)[j];
            for (int l = brpt[k]; l < brpt[k + 1]; l++) {
                int c = bcol[l];
                i64 lo = 0, hi = n;
                while (lo < hi) {
                    i64 mid = (lo + hi) >> 1;
                    if (row[mid] < c)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                double g = is_f32 ? (double)((const float *)dcval)[crpt[i] + lo] : ((const double *)dcval)[crpt[i] + lo];
#pragma omp atomic
                db[l] += a * g;
            }
        }
    }
}
