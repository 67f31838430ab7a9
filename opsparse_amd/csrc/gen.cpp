// Deterministic synthetic CSR generators (host code).
//
// Every random draw is a pure function of (seed, tag, index, level) through a
// splitmix64-style mixer, so the matrices are bit-identical on every machine
// and for every OpenMP thread count.  The reference has no generators (its
// inputs are SuiteSparse .mtx files read by CSR::construct, csr.cu:187-352);
// these build the synthetic workloads the benchmark and tests run on.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "opsparse_amd.h"

namespace {

using u64 = uint64_t;

inline u64 mix64(u64 z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline u64 draw(u64 seed, u64 tag, u64 a, u64 b) { return mix64(mix64(mix64(seed ^ (tag * 0xD1B54A32D192ED03ull)) + a) + b); }

// Uniform in [0, 1) with 53 random bits.
inline double unit(u64 h) { return static_cast<double>(h >> 11) * 0x1.0p-53; }

// Uniform integer in [0, n) by multiply-high (n < 2^63).
inline u64 below(u64 h, u64 n) { return static_cast<u64>((static_cast<unsigned __int128>(h) * n) >> 64); }

enum : u64 { kTagCol = 1, kTagVal = 2, kTagRmat = 3, kTagBatchN = 4, kTagBatchD = 5, kTagNormal = 6 };

bool alloc_csr(ops_host_csr *m, int32_t M, int32_t N, int64_t nnz)
{
    m->M = M;
    m->N = N;
    m->nnz = nnz;
    m->rpt = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * (static_cast<size_t>(M) + 1)));
    m->col = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * static_cast<size_t>(nnz > 0 ? nnz : 1)));
    m->val = static_cast<double *>(std::malloc(sizeof(double) * static_cast<size_t>(nnz > 0 ? nnz : 1)));
    if (!m->rpt || !m->col || !m->val) {
        ops_host_csr_free(m);
        return false;
    }
    return true;
}

// d distinct columns in [0, N) for one row by Floyd's algorithm, written
// sorted into out[0..d).  `mark` is an N-byte scratch that is all zero on
// entry and on exit.
void floyd_row(u64 seed, u64 tag, u64 row, int32_t N, int32_t d, int32_t *out, std::vector<unsigned char> &mark)
{
    int32_t n = 0;
    for (int64_t j = static_cast<int64_t>(N) - d; j < N; j++) {
        const int32_t t = static_cast<int32_t>(below(draw(seed, tag, row, static_cast<u64>(j)), static_cast<u64>(j) + 1));
        const int32_t pick = mark[t] ? static_cast<int32_t>(j) : t;
        mark[pick] = 1;
        out[n++] = pick;
    }
    std::sort(out, out + d);
    for (int32_t i = 0; i < d; i++) mark[out[i]] = 0;
}

void fill_uniform_values(u64 seed, u64 tag, double *val, int64_t nnz)
{
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nnz; e++) val[e] = 2.0 * unit(draw(seed, tag, static_cast<u64>(e), 0)) - 1.0;
}

void fill_normal_values(u64 seed, double *val, int64_t nnz)
{
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nnz; e++) {
        // Box-Muller on two independent uniforms; u1 in (0, 1].
        const double u1 = 1.0 - unit(draw(seed, kTagNormal, static_cast<u64>(e), 0));
        const double u2 = unit(draw(seed, kTagNormal, static_cast<u64>(e), 1));
        val[e] = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
}

}  // namespace

extern "C" {

int ops_host_csr_free(ops_host_csr *m)
{
    if (!m) return OPS_INVALID;
    std::free(m->rpt);
    std::free(m->col);
    std::free(m->val);
    m->rpt = nullptr;
    m->col = nullptr;
    m->val = nullptr;
    m->M = m->N = 0;
    m->nnz = 0;
    return OPS_OK;
}

int ops_gen_uniform_rows(int32_t M, int32_t N, int32_t d, uint64_t seed, ops_host_csr *out)
{
    if (!out || M < 0 || N < 0 || d < 0 || d > N) return OPS_INVALID;
    const int64_t nnz = static_cast<int64_t>(M) * d;
    if (nnz > INT32_MAX) return OPS_INT32_OVERFLOW;
    if (!alloc_csr(out, M, N, nnz)) return OPS_OOM;
    for (int32_t r = 0; r <= M; r++) out->rpt[r] = r * d;
#pragma omp parallel
    {
        std::vector<unsigned char> mark(static_cast<size_t>(N > 0 ? N : 1), 0);
#pragma omp for schedule(static)
        for (int32_t r = 0; r < M; r++) floyd_row(seed, kTagCol, static_cast<u64>(r), N, d, out->col + static_cast<int64_t>(r) * d, mark);
    }
    fill_uniform_values(seed, kTagVal, out->val, nnz);
    return OPS_OK;
}

int ops_gen_stencil27(int32_t nx, int32_t ny, int32_t nz, double diag, double off, double noise, uint64_t seed,
                      ops_host_csr *out)
{
    if (!out || nx <= 0 || ny <= 0 || nz <= 0) return OPS_INVALID;
    const int64_t M64 = static_cast<int64_t>(nx) * ny * nz;
    if (M64 > INT32_MAX) return OPS_INT32_OVERFLOW;
    const int32_t M = static_cast<int32_t>(M64);
    auto span = [](int32_t i, int32_t n) { return (i > 0 ? 1 : 0) + 1 + (i < n - 1 ? 1 : 0); };
    // Row lengths are products of per-axis neighbour counts.
    std::vector<int64_t> rp(static_cast<size_t>(M) + 1, 0);
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < M; r++) {
        const int32_t x = r % nx, y = (r / nx) % ny, z = r / (nx * ny);
        rp[r + 1] = static_cast<int64_t>(span(x, nx)) * span(y, ny) * span(z, nz);
    }
    for (int32_t r = 0; r < M; r++) rp[r + 1] += rp[r];
    if (rp[M] > INT32_MAX) return OPS_INT32_OVERFLOW;
    if (!alloc_csr(out, M, M, rp[M])) return OPS_OOM;
    for (int32_t r = 0; r <= M; r++) out->rpt[r] = static_cast<int32_t>(rp[r]);
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < M; r++) {
        const int32_t x = r % nx, y = (r / nx) % ny, z = r / (nx * ny);
        int64_t e = rp[r];
        // dz outermost, dx innermost: columns come out ascending.
        for (int dz = -1; dz <= 1; dz++) {
            if (z + dz < 0 || z + dz >= nz) continue;
            for (int dy = -1; dy <= 1; dy++) {
                if (y + dy < 0 || y + dy >= ny) continue;
                for (int dx = -1; dx <= 1; dx++) {
                    if (x + dx < 0 || x + dx >= nx) continue;
                    const int32_t c = r + dx + nx * (dy + ny * dz);
                    out->col[e] = c;
                    if (c == r) {
                        out->val[e] = diag;
                    } else {
                        const double u = noise != 0.0 ? (2.0 * unit(draw(seed, kTagVal, static_cast<u64>(e), 0)) - 1.0) : 0.0;
                        out->val[e] = off + noise * u;
                    }
                    e++;
                }
            }
        }
    }
    return OPS_OK;
}

int ops_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed, int normal,
                 ops_host_csr *out)
{
    if (!out || scale < 1 || scale > 30 || edge_factor < 1) return OPS_INVALID;
    const int64_t n = int64_t(1) << scale;
    const int64_t E = static_cast<int64_t>(edge_factor) * n;
    if (E > INT32_MAX) return OPS_INT32_OVERFLOW;
    const double ab = a + b, abc = a + b + c;
    std::vector<int32_t> er(static_cast<size_t>(E)), ec(static_cast<size_t>(E));
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; e++) {
        int64_t r = 0, q = 0;
        for (int l = 0; l < scale; l++) {
            const double u = unit(draw(seed, kTagRmat, static_cast<u64>(e), static_cast<u64>(l)));
            const int rb = u >= ab ? 1 : 0;
            const int cb = (u >= a && u < ab) || u >= abc ? 1 : 0;
            r = (r << 1) | rb;
            q = (q << 1) | cb;
        }
        er[e] = static_cast<int32_t>(r);
        ec[e] = static_cast<int32_t>(q);
    }
    // Counting sort by row, then sort + unique within rows.
    std::vector<std::atomic<int64_t>> cnt(static_cast<size_t>(n));
    for (auto &x : cnt) x.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; e++) cnt[er[e]].fetch_add(1, std::memory_order_relaxed);
    std::vector<int64_t> start(static_cast<size_t>(n) + 1, 0);
    for (int64_t r = 0; r < n; r++) start[r + 1] = start[r] + cnt[r].load(std::memory_order_relaxed);
    for (int64_t r = 0; r < n; r++) cnt[r].store(start[r], std::memory_order_relaxed);
    std::vector<int32_t> cols(static_cast<size_t>(E));
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; e++) cols[cnt[er[e]].fetch_add(1, std::memory_order_relaxed)] = ec[e];
    std::vector<int64_t> len(static_cast<size_t>(n) + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; r++) {
        int32_t *p = cols.data() + start[r];
        const int64_t m = start[r + 1] - start[r];
        std::sort(p, p + m);
        len[r + 1] = std::unique(p, p + m) - p;
    }
    for (int64_t r = 0; r < n; r++) len[r + 1] += len[r];
    if (!alloc_csr(out, static_cast<int32_t>(n), static_cast<int32_t>(n), len[n])) return OPS_OOM;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r <= n; r++) out->rpt[r] = static_cast<int32_t>(len[r]);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; r++)
        std::memcpy(out->col + len[r], cols.data() + start[r], sizeof(int32_t) * static_cast<size_t>(len[r + 1] - len[r]));
    if (normal)
        fill_normal_values(seed, out->val, len[n]);
    else
        fill_uniform_values(seed, kTagVal, out->val, len[n]);
    return OPS_OK;
}

int ops_gen_batch(int32_t count, int32_t n_lo, int32_t n_hi, int32_t d_lo, int32_t d_hi, uint64_t seed,
                  ops_host_csr *A, ops_host_csr *B, int64_t *row_base, int32_t *d_out)
{
    if (!A || !B || !row_base || count < 0 || n_lo < 1 || n_hi < n_lo || d_lo < 0 || d_hi < d_lo || d_hi > n_lo)
        return OPS_INVALID;
    std::vector<int32_t> ni(static_cast<size_t>(count)), di(static_cast<size_t>(count));
    const double llo = std::log(static_cast<double>(n_lo)), lhi = std::log(static_cast<double>(n_hi));
    row_base[0] = 0;
    int64_t nnz = 0;
    for (int32_t i = 0; i < count; i++) {
        double v = std::floor(std::exp(llo + (lhi - llo) * unit(draw(seed, kTagBatchN, static_cast<u64>(i), 0))));
        v = std::min(std::max(v, static_cast<double>(n_lo)), static_cast<double>(n_hi));
        ni[i] = static_cast<int32_t>(v);
        di[i] = d_lo + static_cast<int32_t>(below(draw(seed, kTagBatchD, static_cast<u64>(i), 0), static_cast<u64>(d_hi - d_lo + 1)));
        if (d_out) d_out[i] = di[i];
        row_base[i + 1] = row_base[i] + ni[i];
        nnz += static_cast<int64_t>(ni[i]) * di[i];
    }
    const int64_t M = row_base[count];
    if (M > INT32_MAX || nnz > INT32_MAX) return OPS_INT32_OVERFLOW;
    if (!alloc_csr(A, static_cast<int32_t>(M), static_cast<int32_t>(M), nnz)) return OPS_OOM;
    if (!alloc_csr(B, static_cast<int32_t>(M), static_cast<int32_t>(M), nnz)) {
        ops_host_csr_free(A);
        return OPS_OOM;
    }
    // Row pointers: instance i contributes n_i rows of d_i entries.
    {
        int64_t e = 0;
        int64_t r = 0;
        for (int32_t i = 0; i < count; i++)
            for (int32_t k = 0; k < ni[i]; k++, r++) {
                A->rpt[r] = B->rpt[r] = static_cast<int32_t>(e);
                e += di[i];
            }
        A->rpt[M] = B->rpt[M] = static_cast<int32_t>(e);
    }
    ops_host_csr *mats[2] = {A, B};
    for (int s = 0; s < 2; s++) {
        ops_host_csr *X = mats[s];
        const u64 tag = 16 + static_cast<u64>(s);
#pragma omp parallel
        {
            std::vector<unsigned char> mark(static_cast<size_t>(n_hi), 0);
#pragma omp for schedule(dynamic, 4)
            for (int32_t i = 0; i < count; i++) {
                const int32_t base = static_cast<int32_t>(row_base[i]);
                for (int32_t k = 0; k < ni[i]; k++) {
                    int32_t *out = X->col + X->rpt[base + k];
                    floyd_row(seed, tag, static_cast<u64>(base + k), ni[i], di[i], out, mark);
                    for (int32_t t = 0; t < di[i]; t++) out[t] += base;
                }
            }
        }
        fill_uniform_values(seed, 32 + static_cast<u64>(s), X->val, nnz);
    }
    return OPS_OK;
}

}  // extern "C"
