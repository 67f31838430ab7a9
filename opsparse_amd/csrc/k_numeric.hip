// Numeric phase: values and sorted columns of every row of C.
//
// Reference counterparts: k_numeric_shared_hash_pwarp,
// k_numeric_shared_hash_tb_full_occu<SH_ROW,BS>,
// k_numeric_max_shared_hash_tb_half_occu, k_numeric_global_hash_tb_full_occu
// (include/opsparse/core_algo/spgemm/kernels/numeric.cuh:9-393) and their
// dispatch h_numeric_full_occu (src/opsparse/core_algo/spgemm/spgemm_phases.cu:148-206).
//
// The accumulation is the same open-addressing hash (CAS on the key, atomic
// add of a*b on the value), with tables of >= 2n slots for a row of n
// entries.  The ordering step is different: the reference ranks every entry
// against every other entry, O(n^2) (numeric.cuh:97-106, 194-204, 382-392);
// here rows of up to 256 entries are compacted and bitonic-sorted in
// registers across their unit (cross-lane ds_bpermute), rows up to 4096 are
// bitonic-sorted in LDS, and rows above 4096 entries never need a sort: they
// accumulate into dense LDS column windows whose occupancy bitmap is read
// out in column order.
#include "device.hpp"
#include "launch.hpp"

namespace opsamd {

namespace {

// Consumers of the product walk (device.hpp): accumulate a*b into the
// row's hash table, or into a dense column window.
template <int T, int H, typename V>
struct NumInsert {
    const int *__restrict__ bcol;
    const V *__restrict__ bval;
    int *kt;
    V *vt;
    struct Item {
        int c;
        V b;
    };
    __device__ __forceinline__ Item load(int k) const { return Item{bcol[k], bval[k]}; }
    __device__ __forceinline__ void apply(const Item &it, V a, int)
    {
        bool fresh;
        const int h = hash_insert<T, H>(kt, it.c, fresh);
        atomicAdd(vt + h, a * it.b);
    }
};

template <typename V>
struct DenseInsert {
    const int *__restrict__ bcol;
    const V *__restrict__ bval;
    V *acc;
    unsigned *bm;
    int w0;
    struct Item {
        int c;
        V b;
    };
    __device__ __forceinline__ Item load(int k) const { return Item{bcol[k], bval[k]}; }
    __device__ __forceinline__ void apply(const Item &it, V a, int)
    {
        const int c = it.c - w0;
        atomicAdd(acc + c, a * it.b);
        const unsigned bit = 1u << (c & 31);
        if (!(bm[c >> 5] & bit)) atomicOr(bm + (c >> 5), bit);
    }
};

// Bins 0-4: G <= 64 lanes per row, BS/G rows per block, register sort.
// T table slots, at most NMAX entries per row (T >= 2 * NMAX).
template <int G, int T, int NMAX, int BS, int H, typename V>
__global__ __launch_bounds__(BS) void k_num_small(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                  const V *__restrict__ aval, const int *__restrict__ brpt,
                                                  const int *__restrict__ bcol, const V *__restrict__ bval,
                                                  const int *__restrict__ rows, int nrows,
                                                  const long long *__restrict__ crpt, int *__restrict__ ccol,
                                                  V *__restrict__ cval)
{
    constexpr int U = G;
    constexpr int RPB = BS / G;
    constexpr int S = T / U;     // table slots per lane
    constexpr int K = NMAX / U;  // sorted entries per lane
    static_assert(G <= kWave && BS % G == 0 && K >= 1 && T >= 2 * NMAX, "bad config");
    __shared__ __align__(16) int keys[RPB * T];
    __shared__ __align__(16) V vals[RPB * T];
    __shared__ int ckeys[RPB * NMAX];

    const int tid = threadIdx.x;
    const int slot = tid / G;
    const int gl = tid % G;
    for (int i = tid; i < RPB * T / 4; i += BS) reinterpret_cast<int4 *>(keys)[i] = make_int4(-1, -1, -1, -1);
    for (int i = tid; i < RPB * T * static_cast<int>(sizeof(V)) / 16; i += BS)
        reinterpret_cast<int4 *>(vals)[i] = make_int4(0, 0, 0, 0);
    __syncthreads();

    const int ridx = xcd_tile() * RPB + slot;
    int r = 0;
    int *kt = keys + slot * T;
    V *vt = vals + slot * T;
    if (ridx < nrows) {
        r = rows ? rows[ridx] : ridx;
        const FullRow range{brpt};
        NumInsert<T, H, V> op{bcol, bval, kt, vt};
        for_each_product<U, true, V>(arpt[r], arpt[r + 1], 0, 1, acol, aval, range, op);
    }
    __syncthreads();
    if (ridx >= nrows) return;

    // Compact the row's keys (slot order) into its key list; the table stays
    // intact and supplies the values after the keys are sorted.
    int *ck = ckeys + slot * NMAX;
    int n = 0;
    const unsigned long long below = gl ? ((1ull << gl) - 1ull) : 0ull;
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int key = kt[s * U + gl];
        const bool valid = key != -1;
        const unsigned long long m = unit_ballot<U>(valid);
        const int pos = n + __popcll(m & below);
        if (valid && pos < NMAX) ck[pos] = key;
        n += __popcll(m);
    }
    n = min(n, NMAX);
    wave_lds_sync();

    int sk[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = k * U + gl;
        sk[k] = j < n ? ck[j] : INT_MAX;
    }
    unit_bitonic_sort_keys<U, K>(sk);
    const long long off = crpt[r];
    const int n_out = min(n, static_cast<int>(crpt[r + 1] - off));  // never write past the row
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = k * U + gl;
        if (j < n_out) {
            const int h = hash_find<T, H>(kt, sk[k]);
            ccol[off + j] = sk[k];
            cval[off + j] = h >= 0 ? vt[h] : V(0);
        }
    }
}

// Bins 5-6: one row per block of BS threads; table of T slots in dynamic
// LDS, then compaction of the keys and an LDS bitonic sort of the keys; the
// values are read back from the table by key.
template <int T, int BS, typename V>
__global__ __launch_bounds__(BS) void k_num_block(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                  const V *__restrict__ aval, const int *__restrict__ brpt,
                                                  const int *__restrict__ bcol, const V *__restrict__ bval,
                                                  const int *__restrict__ rows, int nrows,
                                                  const long long *__restrict__ crpt, int *__restrict__ ccol,
                                                  V *__restrict__ cval)
{
    extern __shared__ unsigned long long smem_u64[];
    V *vt = reinterpret_cast<V *>(smem_u64);   // T values
    int *kt = reinterpret_cast<int *>(vt + T);  // T keys
    int *sk = kt + T;                           // T/2 sorted keys
    __shared__ int s_pos;
    constexpr int NU = BS / kWave;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;

    for (int ridx = blockIdx.x; ridx < nrows; ridx += gridDim.x) {
        const int r = rows ? rows[ridx] : ridx;
        for (int i = tid; i < T; i += BS) {
            kt[i] = -1;
            vt[i] = V(0);
        }
        if (tid == 0) s_pos = 0;
        __syncthreads();
        const FullRow range{brpt};
        NumInsert<T, kHashFib, V> op{bcol, bval, kt, vt};
        for_each_product<kWave, true, V>(arpt[r], arpt[r + 1], wave, NU, acol, aval, range, op);
        __syncthreads();
        const long long off = crpt[r];
        const int n = min(static_cast<int>(crpt[r + 1] - off), T / 2);  // rows of this bin have n <= T/2
        for (int i = tid; i < T; i += BS) {
            const int key = kt[i];
            const bool valid = key != -1;
            const unsigned long long m = __ballot(valid);
            int base = 0;
            if (lane_id() == 0 && m) base = atomicAdd(&s_pos, __popcll(m));
            base = bperm(base, 0);
            const int pos = base + __popcll(m & ((1ull << lane_id()) - 1ull));
            if (valid && pos < T / 2) sk[pos] = key;
        }
        int p2 = 1;
        while (p2 < n) p2 <<= 1;
        __syncthreads();
        for (int i = n + tid; i < p2; i += BS) sk[i] = INT_MAX;
        __syncthreads();
        for (int size = 2; size <= p2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = tid; i < (p2 >> 1); i += BS) {
                    // pair index i -> (lo, lo + stride); stride is a power of two
                    const int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));
                    const int hi = lo + stride;
                    const bool asc = (lo & size) == 0;
                    const int a = sk[lo], b = sk[hi];
                    if ((a > b) == asc) {
                        sk[lo] = b;
                        sk[hi] = a;
                    }
                }
                __syncthreads();
            }
        }
        for (int j = tid; j < n; j += BS) {
            const int key = sk[j];
            const int h = hash_find<T, kHashFib>(kt, key);
            ccol[off + j] = key;
            cval[off + j] = h >= 0 ? vt[h] : V(0);
        }
        __syncthreads();
    }
}

constexpr int kLongBS = 1024;

// Bin 7: rows above 4096 entries.  Dense LDS accumulator + occupancy bitmap
// per window of W columns; windows are written out in column order, so the
// output is sorted without a sort.  cursor[] holds two ints per A-nonzero
// index (span = highest A-nonzero index + 1): windows alternate which half
// they read and which they write, so waves that share an A chunk (and all
// compute the same cursor) never read a position another wave has already
// advanced.
template <typename V>
__global__ __launch_bounds__(kLongBS) void k_num_long(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                      const V *__restrict__ aval, const int *__restrict__ brpt,
                                                      const int *__restrict__ bcol, const V *__restrict__ bval,
                                                      const int *__restrict__ rows, int nrows, int ncols,
                                                      const long long *__restrict__ crpt, int *__restrict__ ccol,
                                                      V *__restrict__ cval, int *__restrict__ cursor, int span)
{
    using D = DenseWin<V>;
    extern __shared__ unsigned long long smem_u64[];
    V *acc = reinterpret_cast<V *>(smem_u64);
    unsigned *bm = reinterpret_cast<unsigned *>(acc + D::W);
    __shared__ int s_scan[kLongBS / kWave + 1];
    __shared__ int s_min[kLongBS / kWave];
    constexpr int NU = kLongBS / kWave;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;

    for (int i = tid; i < D::W; i += kLongBS) acc[i] = V(0);
    for (int i = tid; i < D::WORDS; i += kLongBS) bm[i] = 0u;
    __syncthreads();

    for (int ridx = blockIdx.x; ridx < nrows; ridx += gridDim.x) {
        const int r = rows ? rows[ridx] : ridx;
        const int a_beg = arpt[r], a_end = arpt[r + 1];
        long long out = crpt[r];
        const long long row_end = crpt[r + 1];
        // First window: the one holding the smallest column of the row.
        int mn = INT_MAX;
        for (int e = a_beg + tid; e < a_end; e += kLongBS) {
            const int k = acol[e];
            if (brpt[k] < brpt[k + 1]) mn = min(mn, bcol[brpt[k]]);
        }
        mn = unit_min<kWave>(mn);
        if (lane_id() == 0) s_min[wave] = mn;
        __syncthreads();
        mn = INT_MAX;
        for (int w = 0; w < NU; w++) mn = min(mn, s_min[w]);
        __syncthreads();
        int w0 = mn == INT_MAX ? ncols : (mn / D::W) * D::W;
        bool first = true;
        int parity = 0;
        while (w0 < ncols) {
            const int wend = min(ncols, w0 + D::W);
            int nxt = INT_MAX;
            int *cur_a = cursor + (parity ? span : 0);
            int *cur_b = cursor + (parity ? 0 : span);
            const CursorRange range{brpt, bcol, cur_a, cur_b, wend, first, wend >= ncols, &nxt};
            DenseInsert<V> op{bcol, bval, acc, bm, w0};
            for_each_product<kWave, true, V>(a_beg, a_end, wave, NU, acol, aval, range, op);
            nxt = unit_min<kWave>(nxt);
            if (lane_id() == 0) s_min[wave] = nxt;
            __syncthreads();
            // Read the window out in column order.
            for (int wb = 0; wb < D::WORDS; wb += kLongBS) {
                const int word = wb + tid;
                const unsigned bits = word < D::WORDS ? bm[word] : 0u;
                int tot;
                const int pos = block_excl_scan<kLongBS>(__popc(bits), s_scan, tot);
                unsigned b = bits;
                int i = 0;
                while (b) {
                    const int bit = __ffs(b) - 1;
                    b &= b - 1;
                    const int c = word * 32 + bit;
                    const long long at = out + pos + i;
                    if (at < row_end) {
                        ccol[at] = w0 + c;
                        cval[at] = acc[c];
                    }
                    acc[c] = V(0);
                    i++;
                }
                if (word < D::WORDS) bm[word] = 0u;
                out += tot;
            }
            nxt = INT_MAX;
            for (int w = 0; w < NU; w++) nxt = min(nxt, s_min[w]);
            __syncthreads();
            first = false;
            parity ^= 1;
            w0 = nxt == INT_MAX ? ncols : (nxt / D::W) * D::W;
        }
    }
}

template <int G, int T, int NMAX, int BS, typename V>
void launch_small(const CsrIn &A, const CsrIn &B, const int *rows, int nrows, const long long *crpt, int *ccol,
                  V *cval, int hash, hipStream_t s)
{
    constexpr int RPB = BS / G;
    const int grid = xcd_grid((nrows + RPB - 1) / RPB);
    if (hash == kHashLin)
        k_num_small<G, T, NMAX, BS, kHashLin, V><<<grid, BS, 0, s>>>(A.rpt, A.col, static_cast<const V *>(A.val),
                                                                     B.rpt, B.col, static_cast<const V *>(B.val), rows,
                                                                     nrows, crpt, ccol, cval);
    else
        k_num_small<G, T, NMAX, BS, kHashFib, V><<<grid, BS, 0, s>>>(A.rpt, A.col, static_cast<const V *>(A.val),
                                                                     B.rpt, B.col, static_cast<const V *>(B.val), rows,
                                                                     nrows, crpt, ccol, cval);
}

template <int T, int BS, typename V>
constexpr size_t block_lds()
{
    return sizeof(V) * T + sizeof(int) * T + sizeof(int) * (T / 2);
}

template <int T, int BS, typename V>
void launch_block(const CsrIn &A, const CsrIn &B, const int *rows, int nrows, const long long *crpt, int *ccol,
                  V *cval, hipStream_t s)
{
    k_num_block<T, BS, V><<<nrows, BS, block_lds<T, BS, V>(), s>>>(A.rpt, A.col, static_cast<const V *>(A.val), B.rpt,
                                                                   B.col, static_cast<const V *>(B.val), rows, nrows,
                                                                   crpt, ccol, cval);
}

template <typename V>
void set_attrs()
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_num_block<8192, 1024, V>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(block_lds<8192, 1024, V>()));
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_num_block<2048, 256, V>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(block_lds<2048, 256, V>()));
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_num_long<V>), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(DenseWin<V>::lds));
}

}  // namespace

void numeric_init_attributes()
{
    set_attrs<double>();
    set_attrs<float>();
}

template <typename V>
void launch_numeric_bin(int bin, const CsrIn &A, const CsrIn &B, const int *rows, int nrows, const long long *crpt,
                        int *ccol, V *cval, int *cursor, int span, int hash, hipStream_t s)
{
    if (nrows <= 0) return;
    switch (bin) {
        case 0: launch_small<16, 32, 16, 256, V>(A, B, rows, nrows, crpt, ccol, cval, hash, s); break;
        case 1: launch_small<32, 64, 32, 256, V>(A, B, rows, nrows, crpt, ccol, cval, hash, s); break;
        case 2: launch_small<64, 128, 64, 256, V>(A, B, rows, nrows, crpt, ccol, cval, hash, s); break;
        case 3: launch_small<64, 256, 128, 256, V>(A, B, rows, nrows, crpt, ccol, cval, hash, s); break;
        case 4: launch_small<64, 512, 256, 256, V>(A, B, rows, nrows, crpt, ccol, cval, hash, s); break;
        case 5: launch_block<2048, 256, V>(A, B, rows, nrows, crpt, ccol, cval, s); break;
        case 6: launch_block<8192, 1024, V>(A, B, rows, nrows, crpt, ccol, cval, s); break;
        default:
            k_num_long<V><<<nrows, kLongBS, DenseWin<V>::lds, s>>>(
                A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col, static_cast<const V *>(B.val), rows, nrows,
                B.N, crpt, ccol, cval, cursor, span);
            break;
    }
}

template void launch_numeric_bin<double>(int, const CsrIn &, const CsrIn &, const int *, int, const long long *, int *,
                                         double *, int *, int, int, hipStream_t);
template void launch_numeric_bin<float>(int, const CsrIn &, const CsrIn &, const int *, int, const long long *, int *,
                                        float *, int *, int, int, hipStream_t);

}  // namespace opsamd
