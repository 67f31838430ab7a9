// Masked backward of C = A*B at fixed sparsity (no reference counterpart):
//   dA(i,k) = sum_j dC(i,j) * B(k,j)   for (i,k) in pattern(A)
//   dB(k,j) = sum_i A(i,k) * dC(i,j)   for (k,j) in pattern(B)
// Every forward product (i,k,j) contributes once to each.  Row i of C (its
// sorted columns and dC values) is loaded into the same LDS open-addressing
// table the forward numeric kernels use, keyed by column, and the products
// of row i are walked with the same flattened per-unit iterator; dA is
// reduced per A-nonzero in LDS, dB is scattered with global atomic adds.
// Rows of C above 4096 entries look dC up by binary search in the sorted
// C row instead of an LDS table.
#include <algorithm>

#include "device.hpp"
#include "launch.hpp"

namespace opsamd {

namespace {

__device__ __forceinline__ void global_add(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void global_add(float *p, float v) { unsafeAtomicAdd(p, v); }

// Products of one A row with per-A-nonzero reduction: f(b_index, a_value)
// returns the dA contribution of that product; the sum per A-nonzero e is
// written to da[e].  sda: U values of LDS owned by this unit.
template <int U, typename V, typename F>
__device__ __forceinline__ void reduce_products(int a_beg, int a_end, int unit, int nunits,
                                                const int *__restrict__ acol, const V *__restrict__ aval,
                                                const int *__restrict__ brpt, V *sda, V *__restrict__ da, F &&f)
{
    const int lane = lane_id();
    const int gl = lane & (U - 1);
    const int gbase = lane & ~(U - 1);
    for (int c = a_beg + unit * U; c < a_end; c += nunits * U) {
        const int e = c + gl;
        int bs = 0, len = 0;
        V av = V(0);
        if (e < a_end) {
            const int k = acol[e];
            bs = brpt[k];
            len = brpt[k + 1] - bs;
            av = aval[e];
        }
        sda[gl] = V(0);
        wave_lds_sync();
        const int incl = unit_incl_scan<U>(len);
        const int excl = incl - len;
        const int total = bperm(incl, gbase + U - 1);
        for (int base = 0; base < total; base += U) {
            const int p = base + gl;
            int q = 0, qv = 0;
#pragma unroll
            for (int s = U >> 1; s >= 1; s >>= 1) {
                const int cand = q + s;
                const int v = bperm(excl, gbase + cand);
                if (v <= p) {
                    q = cand;
                    qv = v;
                }
            }
            const int qb = bperm(bs, gbase + q);
            const V qa = bperm(av, gbase + q);
            if (p < total) atomicAdd(sda + q, f(qb + (p - qv), qa));
        }
        wave_lds_sync();
        if (e < a_end) da[e] = sda[gl];
        wave_lds_sync();
    }
}

// G lanes per row (G <= 64: BS/G rows per block; G == BS: one row per block).
// Dynamic LDS: keys[RPB*T] | vals[RPB*T] | sda[BS].
template <int G, int T, int BS, typename V>
__global__ __launch_bounds__(BS) void k_bwd_hash(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                 const V *__restrict__ aval, const int *__restrict__ brpt,
                                                 const int *__restrict__ bcol, const V *__restrict__ bval,
                                                 const int *__restrict__ rows, int nrows,
                                                 const long long *__restrict__ crpt, const int *__restrict__ ccol,
                                                 const V *__restrict__ dc, V *__restrict__ da, V *__restrict__ db)
{
    constexpr int U = G < kWave ? G : kWave;
    constexpr int RPB = BS / G;
    constexpr int NU = G / U;
    extern __shared__ unsigned long long smem_u64[];
    V *vals = reinterpret_cast<V *>(smem_u64);
    V *sda = vals + RPB * T;
    int *keys = reinterpret_cast<int *>(sda + BS);

    const int tid = threadIdx.x;
    const int slot = tid / G;
    const int unit = (tid % G) / U;
    for (int i = tid; i < RPB * T; i += BS) keys[i] = -1;
    __syncthreads();
    const int ridx = blockIdx.x * RPB + slot;
    int r = 0;
    int *kt = keys + slot * T;
    V *vt = vals + slot * T;
    if (ridx < nrows) {
        r = rows ? rows[ridx] : ridx;
        const long long c0 = crpt[r], c1 = crpt[r + 1];
        for (long long j = c0 + (tid % G); j < c1; j += G) {
            bool fresh;
            const int h = hash_insert<T, (G <= kWave ? kHashLin : kHashFib)>(kt, ccol[j], fresh);
            vt[h] = dc[j];
        }
    }
    __syncthreads();
    if (ridx >= nrows) return;
    reduce_products<U, V>(arpt[r], arpt[r + 1], unit, NU, acol, aval, brpt, sda + (tid / U) * U, da,
                          [&](int k, V a) -> V {
                              const int h = hash_find<T, (G <= kWave ? kHashLin : kHashFib)>(kt, bcol[k]);
                              const V g = h >= 0 ? vt[h] : V(0);
                              global_add(db + k, a * g);
                              return g * bval[k];
                          });
}

constexpr int kLongBS = 1024;

__global__ void k_zero_vals_f64(double *p, long long n)
{
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_zero_vals_f32(float *p, long long n)
{
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        p[i] = 0.f;
}

// Rows of C above 4096 entries: binary search of the sorted C row.
template <typename V>
__global__ __launch_bounds__(kLongBS) void k_bwd_long(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                      const V *__restrict__ aval, const int *__restrict__ brpt,
                                                      const int *__restrict__ bcol, const V *__restrict__ bval,
                                                      const int *__restrict__ rows, int nrows,
                                                      const long long *__restrict__ crpt,
                                                      const int *__restrict__ ccol, const V *__restrict__ dc,
                                                      V *__restrict__ da, V *__restrict__ db)
{
    __shared__ V sda[kLongBS];
    constexpr int NU = kLongBS / kWave;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    for (int ridx = blockIdx.x; ridx < nrows; ridx += gridDim.x) {
        const int r = rows ? rows[ridx] : ridx;
        const long long c0 = crpt[r], c1 = crpt[r + 1];
        reduce_products<kWave, V>(arpt[r], arpt[r + 1], wave, NU, acol, aval, brpt, sda + wave * kWave, da,
                                  [&](int k, V a) -> V {
                                      const int col = bcol[k];
                                      long long lo = c0, hi = c1;
                                      while (lo < hi) {
                                          const long long m = (lo + hi) >> 1;
                                          if (ccol[m] < col)
                                              lo = m + 1;
                                          else
                                              hi = m;
                                      }
                                      const V g = (lo < c1 && ccol[lo] == col) ? dc[lo] : V(0);
                                      global_add(db + k, a * g);
                                      return g * bval[k];
                                  });
    }
}

template <int G, int T, int BS, typename V>
constexpr size_t bwd_lds()
{
    return (sizeof(V) * ((BS / G) * T + BS)) + sizeof(int) * (BS / G) * T;
}

template <int G, int T, int BS, typename V>
void launch_hash(const CsrIn &A, const CsrIn &B, const int *rows, int nrows, const long long *crpt, const int *ccol,
                 const V *dc, V *da, V *db, hipStream_t s)
{
    constexpr int RPB = BS / G;
    const int grid = (nrows + RPB - 1) / RPB;
    k_bwd_hash<G, T, BS, V><<<grid, BS, bwd_lds<G, T, BS, V>(), s>>>(
        A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col, static_cast<const V *>(B.val), rows, nrows, crpt,
        ccol, dc, da, db);
}

template <int G, int T, int BS, typename V>
void set_attr()
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_bwd_hash<G, T, BS, V>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bwd_lds<G, T, BS, V>()));
}

template <typename V>
void set_attrs()
{
    set_attr<16, 32, 256, V>();
    set_attr<32, 64, 256, V>();
    set_attr<64, 128, 256, V>();
    set_attr<64, 256, 256, V>();
    set_attr<64, 512, 256, V>();
    set_attr<256, 2048, 256, V>();
    set_attr<1024, 8192, 1024, V>();
}

}  // namespace

void backward_init_attributes()
{
    set_attrs<double>();
    set_attrs<float>();
}

template <typename V>
void launch_backward_bin(int bin, const CsrIn &A, const CsrIn &B, const int *rows, int nrows, const long long *crpt,
                         const int *ccol, const V *dc, V *da, V *db, hipStream_t s)
{
    if (nrows <= 0) return;
    switch (bin) {
        case 0: launch_hash<16, 32, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, s); break;
        case 1: launch_hash<32, 64, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, s); break;
        case 2: launch_hash<64, 128, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, s); break;
        case 3: launch_hash<64, 256, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, s); break;
        case 4: launch_hash<64, 512, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, s); break;
        case 5: launch_hash<256, 2048, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, s); break;
        case 6: launch_hash<1024, 8192, 1024, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, s); break;
        default:
            k_bwd_long<V><<<nrows, kLongBS, 0, s>>>(A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col,
                                                    static_cast<const V *>(B.val), rows, nrows, crpt, ccol, dc, da,
                                                    db);
            break;
    }
}

template void launch_backward_bin<double>(int, const CsrIn &, const CsrIn &, const int *, int, const long long *,
                                          const int *, const double *, double *, double *, hipStream_t);
template void launch_backward_bin<float>(int, const CsrIn &, const CsrIn &, const int *, int, const long long *,
                                         const int *, const float *, float *, float *, hipStream_t);

void launch_zero_values(void *p, long long n, bool f32, hipStream_t s)
{
    if (n <= 0) return;
    const int grid = static_cast<int>(std::min<long long>((n + 255) / 256, 65536));
    if (f32)
        k_zero_vals_f32<<<grid, 256, 0, s>>>(static_cast<float *>(p), n);
    else
        k_zero_vals_f64<<<grid, 256, 0, s>>>(static_cast<double *>(p), n);
}

}  // namespace opsamd
