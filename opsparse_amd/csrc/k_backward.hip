// Masked backward of C = A*B at fixed sparsity (no reference counterpart):
//   dA(i,k) = sum_j dC(i,j) * B(k,j)   for (i,k) in pattern(A)
//   dB(k,j) = sum_i A(i,k) * dC(i,j)   for (k,j) in pattern(B)
// Every forward product (i,k,j) contributes once to each.  Rows are binned
// by the length of row i of C, and the products of row i are walked with
// the forward kernels' product walk (for_each_product); dC(i,j) is looked
// up per product:
//  * rows up to 4096 entries: row i of C (its columns and dC values) is
//    loaded into the same LDS open-addressing table the forward numeric
//    kernels use, keyed by column;
//  * longer rows: bitmap-rank windows built from row i of C (k_bwd_rank):
//    dC is staged in LDS in C-row order and looked up by the column's rank;
//    B rows are consumed window by window through per-A-nonzero cursors (as
//    in the forward's long-row kernel).  Windows without entries in row i of
//    C hold no products and are skipped by reading the next column of the C
//    row.
// dA is reduced per A-nonzero in LDS, then added to da once per chunk; dB is
// scattered with global atomic adds.  da and db hold one value per entry of
// A's and B's descriptors counted from their first entry: entry e of A (an
// index into A.col) is da[e - a0], a0 = A.rpt[0], and likewise for B, so a
// row slice's gradients fill a buffer of the slice's own nnz.
#include <algorithm>

#include "device.hpp"
#include "launch.hpp"

namespace opsamd {

namespace {

__device__ __forceinline__ void global_add(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void global_add(float *p, float v) { unsafeAtomicAdd(p, v); }

// dA reduction of one unit: products add their dA term into the LDS slot of
// their A-nonzero (q, 0..U-1); at the end of a chunk each lane moves its
// slot into da[e].  The slots are zero between chunks.  When other units
// walked slices of the same chunk the move is an atomic add; otherwise this
// unit is the only one that ever handles e, so a plain read-modify-write is
// enough (da starts at zero).
template <int U, typename V>
struct DaReduce {
    Acc *sda;
    V *__restrict__ da;
    int a0;  // first entry of A's descriptor (da[e - a0] holds entry e)
    __device__ __forceinline__ void add(int q, Acc v) { atomicAdd(sda + q, v); }
    __device__ __forceinline__ void flush(int e, bool valid, bool shared)
    {
        wave_lds_sync();
        const int gl = lane_id() & (U - 1);
        const Acc v = sda[gl];
        sda[gl] = Acc(0);
        wave_lds_sync();
        if (valid && v != Acc(0)) {
            if (shared)
                global_add(da + (e - a0), static_cast<V>(v));
            else
                da[e - a0] += static_cast<V>(v);
        }
    }
};

// Consumer for rows whose C row sits in an LDS hash table (kt keys, vt dC).
template <int T, int H, int U, typename V>
struct BwdHash {
    const int *__restrict__ bcol;
    const V *__restrict__ bval;
    const int *kt;
    const V *vt;
    V *__restrict__ db;
    int b0;  // first entry of B's descriptor
    DaReduce<U, V> red;
    struct Item {
        int k;
        int c;
        V b;
    };
    __device__ __forceinline__ Item load(int k) const { return Item{k, bcol[k], bval[k]}; }
    __device__ __forceinline__ void apply(const Item &it, V a, int q)
    {
        const int h = hash_find<T, H>(kt, it.c);
        const V g = h >= 0 ? vt[h] : V(0);
        global_add(db + (it.k - b0), a * g);
        red.add(q, static_cast<Acc>(g) * static_cast<Acc>(it.b));
    }
    __device__ __forceinline__ void chunk_end(int e, bool valid, bool shared) { red.flush(e, valid, shared); }
};

// G lanes per row (G <= 64: BS/G rows per block; G > 64: one row per block,
// G/64 waves share the table).  Dynamic LDS: vals[RPB*T] | sda[BS] | keys[RPB*T].
template <int G, int T, int BS, typename V>
__global__ __launch_bounds__(BS) void k_bwd_hash(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                 const V *__restrict__ aval, const int *__restrict__ brpt,
                                                 const int *__restrict__ bcol, const V *__restrict__ bval,
                                                 const int *__restrict__ rows, int nrows,
                                                 const long long *__restrict__ crpt, const int *__restrict__ ccol,
                                                 const V *__restrict__ dc, V *__restrict__ da, V *__restrict__ db,
                                                 int a0, int b0)
{
    constexpr int U = G < kWave ? G : kWave;
    constexpr int RPB = BS / G;
    constexpr int NU = G / U;
    constexpr int H = G <= kWave ? kHashLin : kHashFib;
    extern __shared__ unsigned long long smem_u64[];
    V *vals = reinterpret_cast<V *>(smem_u64);
    Acc *sda = reinterpret_cast<Acc *>(vals + RPB * T);  // RPB * T is even: 8-byte aligned
    int *keys = reinterpret_cast<int *>(sda + BS);

    const int tid = threadIdx.x;
    const int slot = tid / G;
    const int unit = (tid % G) / U;
    for (int i = tid; i < RPB * T; i += BS) keys[i] = -1;
    sda[tid] = Acc(0);
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
            const int h = hash_insert<T, H>(kt, ccol[j], fresh);
            vt[h] = dc[j];
        }
    }
    __syncthreads();
    if (ridx >= nrows) return;
    BwdHash<T, H, U, V> op{bcol, bval, kt, vt, db, b0, DaReduce<U, V>{sda + (tid / U) * U, da, a0}};
    const FullRow range{brpt};
    for_each_product<U, true, V, (G > 256)>(arpt[r], arpt[r + 1], unit, NU, acol, aval, range, op);
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

// Rows of C above 4096 entries: bitmap-rank windows, as in the forward's
// long-row kernel, with the bitmap built from row i of C itself.  Per
// window of WB = 2^18 columns: the columns of the C row inside it set their
// bits, a block scan gives every word its rank (the rank of column c is
// then its position among the row's entries of the window), and for each
// sub-window of up to VCAP entries the dC values are staged in LDS in C-row
// order (a contiguous copy).  The products of the sub-window, found through
// per-A-nonzero cursors, look their dC up by rank.  cursor[] holds two ints
// per A-nonzero index (span = highest A-nonzero index + 1); sub-windows
// alternate which half they read and which they write, so waves that share
// an A chunk never read a position another wave has already advanced.
template <typename V>
struct BwdRankWin {
    static constexpr int WB = 1 << 18;
    static constexpr int WORDS = WB / 32;
    static constexpr int VCAP = sizeof(V) == 8 ? 8192 : 16384;
    static constexpr size_t lds = sizeof(uint2) * WORDS + sizeof(V) * VCAP;
};

// Consumer for one sub-window: dC of the product's column at rank - r0.
// The dA term goes to the LDS slot of the batch's A-nonzero (its index e
// minus the batch start).
template <typename V>
struct BwdRank {
    const int *__restrict__ bcol;
    const V *__restrict__ bval;
    const uint2 *br;
    const V *gv;
    int w0;
    int r0;
    int base;
    V *__restrict__ db;
    int b0;  // first entry of B's descriptor
    Acc *sda;
    struct Item {
        int k;
        int c;
        V b;
    };
    __device__ __forceinline__ Item load(int k) const { return Item{k, bcol[k], bval[k]}; }
    __device__ __forceinline__ uint2 prep(const Item &it) const { return br[(it.c - w0) >> 5]; }
    __device__ __forceinline__ void commit(const Item &it, const uint2 &e, V a, int q)
    {
        const int c = it.c - w0;
        const V g = gv[static_cast<int>(e.y) + __popc(e.x & ((1u << (c & 31)) - 1u)) - r0];
        global_add(db + (it.k - b0), a * g);
        atomicAdd(sda + (q - base), static_cast<Acc>(g) * static_cast<Acc>(it.b));
    }
};

// Rows above 4096 C entries.  Per window of the C row (WB columns from the
// window of its next column): the row's C columns set an LDS bitmap, a scan
// of the word popcounts ranks them (DPP inside each wave, wave totals
// through LDS), and dC is staged by rank in sub-windows of VCAP entries.
// Each sub-window walks its products with the whole block (block_stage +
// block_walk over batches of kLongBS A-nonzeros; B ranges start at
// per-A-nonzero cursors and end at the sub-window's last column, found
// between the cuts of B's unit cut table, or by galloping without one);
// after a batch, each thread adds its A-nonzero's dA slot to da.
template <typename V>
__global__ __launch_bounds__(kLongBS) void k_bwd_rank(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                      const V *__restrict__ aval, const int *__restrict__ brpt,
                                                      const int *__restrict__ bcol, const V *__restrict__ bval,
                                                      const int *__restrict__ rows, int nrows, int ncols,
                                                      const long long *__restrict__ crpt,
                                                      const int *__restrict__ ccol, const V *__restrict__ dc,
                                                      V *__restrict__ da, V *__restrict__ db, int *__restrict__ cursor,
                                                      int span, int a0, int b0, const int *__restrict__ ucut,
                                                      int ustride)
{
    using R = BwdRankWin<V>;
    constexpr int NU = kLongBS / kWave;
    constexpr int WPT = R::WORDS / kLongBS;
    constexpr int WPW = R::WORDS / NU;
    constexpr int RPW = WPW / kWave;
    extern __shared__ unsigned long long smem_u64[];
    uint2 *br = reinterpret_cast<uint2 *>(smem_u64);
    V *gv = reinterpret_cast<V *>(br + R::WORDS);
    __shared__ Acc sda[kLongBS];
    __shared__ int s_cnt[kLongBS / kWave];
    __shared__ BlockStage<kLongBS, V> s_bst;
    static_assert(R::lds + sizeof(Acc) * kLongBS + sizeof(BlockStage<kLongBS, V>) + 256 <= kMaxLds, "LDS");
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int lane = lane_id();
    sda[tid] = Acc(0);
    for (int i = 0; i < WPT; i++) br[tid + i * kLongBS] = make_uint2(0u, 0u);
    __syncthreads();

    for (int ridx = blockIdx.x; ridx < nrows; ridx += gridDim.x) {
        const int r = rows ? rows[ridx] : ridx;
        const int a_beg = arpt[r], a_end = arpt[r + 1];
        const long long c1 = crpt[r + 1];
        long long cpos = crpt[r];
        bool first = true;
        int parity = 0;
        while (cpos < c1) {
            // Window of the next column of row i of C: its entries below
            // wend are a prefix of [cpos, c1) of at most WB entries.
            const int w0 = (ccol[cpos] / R::WB) * R::WB;
            const int wend = (ncols - w0 <= R::WB) ? ncols : w0 + R::WB;
            const long long lim = min(c1, cpos + static_cast<long long>(R::WB));
            for (long long j = cpos + tid; j < lim; j += kLongBS) {
                const int c = ccol[j];
                if (c < wend) atomicOr(&br[(c - w0) >> 5].x, 1u << ((c - w0) & 31));
            }
            __syncthreads();
            int ci[RPW], inc[RPW];
            int wtot = 0;
#pragma unroll
            for (int i = 0; i < RPW; i++) {
                ci[i] = __popc(br[wave * WPW + i * kWave + lane].x);
                inc[i] = wave_incl_scan(ci[i]);
                wtot += __builtin_amdgcn_readlane(inc[i], kWave - 1);
            }
            if (lane == 0) s_cnt[wave] = wtot;
            __syncthreads();
            int tot = 0, run = 0;
            for (int w = 0; w < NU; w++) {
                const int t = s_cnt[w];
                run += w < wave ? t : 0;
                tot += t;
            }
#pragma unroll
            for (int i = 0; i < RPW; i++) {
                br[wave * WPW + i * kWave + lane].y = static_cast<unsigned>(run + inc[i] - ci[i]);
                run += __builtin_amdgcn_readlane(inc[i], kWave - 1);
            }
            __syncthreads();
            // Sub-windows of up to VCAP entries [r0, r1) of the window.
            for (int r0 = 0; r0 < tot; r0 += R::VCAP) {
                const int r1 = min(tot, r0 + R::VCAP);
                const int cb = r1 < tot ? ccol[cpos + r1] : wend;
                for (int i = tid; i < r1 - r0; i += kLongBS) gv[i] = dc[cpos + r0 + i];
                __syncthreads();
                drain_vmem();
                const CursorRange range{brpt,  bcol,        cursor + (parity ? span : 0), cursor + (parity ? 0 : span),
                                        cb,    first,       cb >= ncols,                  nullptr,
                                        ucut,  ustride};
                BwdRank<V> op{bcol, bval, br, gv, w0, r0, 0, db, b0, sda};
                for (int c = a_beg; c < a_end; c += kLongBS) {
                    int total;
                    const int n = block_stage<kLongBS, V>(c, a_end, acol, aval, range, s_bst, total);
                    op.base = c;
                    if (total > 0) block_walk<kLongBS, true, V>(s_bst, n, total, op);
                    const int e = c + tid;
                    if (e < a_end) {
                        const Acc v = sda[tid];
                        if (v != Acc(0)) {
                            da[e - a0] += static_cast<V>(v);
                            sda[tid] = Acc(0);
                        }
                    }
                }
                // Every read of gv is done before the next stage.
                __syncthreads();
                first = false;
                parity ^= 1;
            }
            for (int i = 0; i < WPT; i++) br[tid + i * kLongBS] = make_uint2(0u, 0u);
            __syncthreads();
            cpos += tot;
        }
    }
}

template <int G, int T, int BS, typename V>
constexpr size_t bwd_lds()
{
    return sizeof(V) * (BS / G) * T + sizeof(Acc) * BS + sizeof(int) * (BS / G) * T;
}

template <int G, int T, int BS, typename V>
void launch_hash(const CsrIn &A, const CsrIn &B, const int *rows, int nrows, const long long *crpt, const int *ccol,
                 const V *dc, V *da, V *db, int a0, int b0, hipStream_t s)
{
    constexpr int RPB = BS / G;
    const int grid = (nrows + RPB - 1) / RPB;
    k_bwd_hash<G, T, BS, V><<<grid, BS, bwd_lds<G, T, BS, V>(), s>>>(
        A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col, static_cast<const V *>(B.val), rows, nrows, crpt,
        ccol, dc, da, db, a0, b0);
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
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_bwd_rank<V>), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(BwdRankWin<V>::lds));
}

}  // namespace

void backward_init_attributes()
{
    set_attrs<double>();
    set_attrs<float>();
}

template <typename V>
void launch_backward_bin(int bin, const CsrIn &A, const CsrIn &B, const int *rows, int nrows, const long long *crpt,
                         const int *ccol, const V *dc, V *da, V *db, int a0, int b0, int *cursor, int span,
                         const int *ucut, int nunit, hipStream_t s)
{
    if (nrows <= 0) return;
    switch (bin) {
        case 0: launch_hash<16, 32, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, a0, b0, s); break;
        case 1: launch_hash<32, 64, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, a0, b0, s); break;
        case 2: launch_hash<64, 128, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, a0, b0, s); break;
        case 3: launch_hash<64, 256, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, a0, b0, s); break;
        case 4: launch_hash<64, 512, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, a0, b0, s); break;
        case 5: launch_hash<256, 2048, 256, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, a0, b0, s); break;
        case 6: launch_hash<1024, 8192, 1024, V>(A, B, rows, nrows, crpt, ccol, dc, da, db, a0, b0, s); break;
        default:
            k_bwd_rank<V><<<nrows, kLongBS, BwdRankWin<V>::lds, s>>>(
                A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col, static_cast<const V *>(B.val), rows, nrows,
                B.N, crpt, ccol, dc, da, db, cursor, span, a0, b0, ucut, nunit + 1);
            break;
    }
}

template void launch_backward_bin<double>(int, const CsrIn &, const CsrIn &, const int *, int, const long long *,
                                          const int *, const double *, double *, double *, int, int, int *, int,
                                          const int *, int, hipStream_t);
template void launch_backward_bin<float>(int, const CsrIn &, const CsrIn &, const int *, int, const long long *,
                                         const int *, const float *, float *, float *, int, int, int *, int,
                                         const int *, int, hipStream_t);

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
