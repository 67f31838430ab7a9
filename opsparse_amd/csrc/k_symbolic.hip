// Symbolic phase: nnz of every row of C.
//
// Reference counterparts: k_symbolic_shared_hash_pwarp / _tb / _large_ /
// _max_..._with_fail / k_symbolic_global_hash_tb
// (include/opsparse/core_algo/spgemm/kernels/symbolic.cuh:9-325) and their
// dispatch h_symbolic (src/opsparse/core_algo/spgemm/spgemm_phases.cu:52-121).
//
// Bins 0-6: LDS hash table of T int32 keys per row with T > row products,
// so a row never overflows its table and there is no fail/retry pass (the
// reference's bin 7 probes, fails and recomputes, symbolic.cuh:192-261).
// New keys are counted per lane and summed once per row, not with an LDS
// atomic per insert.
// Bin 7 (more than 16384 products): an LDS bitmap over up to 2^20 columns
// per window; a set bit is a column of C, counted when its atomicOr flips it.
#include <algorithm>

#include "device.hpp"
#include "launch.hpp"

namespace opsamd {

namespace {

// Consumers of the product walk (device.hpp): insert B's column into the
// row's hash table / bitmap and count the keys that were new.
template <int T, int H>
struct SymInsert {
    const int *__restrict__ bcol;
    int *tab;
    int cnt = 0;
    using Item = int;
    __device__ __forceinline__ int load(int k) const { return bcol[k]; }
    __device__ __forceinline__ void apply(int c, double, int)
    {
        bool fresh;
        hash_insert<T, H>(tab, c, fresh);
        cnt += fresh ? 1 : 0;
    }
};

// Bitmap of a row with a narrow column span (columns lo .. lo + 32*T - 1):
// sets the column's bit; the caller counts by popcount.
struct SymBitmap {
    const int *__restrict__ bcol;
    unsigned *bm;
    int lo;
    using Item = int;
    __device__ __forceinline__ int load(int k) const { return bcol[k]; }
    __device__ __forceinline__ void apply(int col, double, int)
    {
        const int c = col - lo;
        atomicOr(&bm[c >> 5], 1u << (c & 31));
    }
};

// G lanes per row (G <= 64: BS/G rows per block; G == BS > 64: one row per
// block, G/64 waves share the table).
template <int G, int T, int BS, int H, bool BITMAP>
__global__ __launch_bounds__(BS) void k_sym_hash(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                 const int *__restrict__ brpt, const int *__restrict__ bcol,
                                                 const RowList rl, int *__restrict__ row_nnz)
{
    constexpr int U = G < kWave ? G : kWave;
    constexpr int RPB = BS / G;
    constexpr int NU = G / U;
    static_assert(BS % G == 0 && (T & (T - 1)) == 0, "bad config");
    __shared__ __align__(16) int tab[RPB * T];
    __shared__ int s_cnt[RPB];
    __shared__ int s_tile;

    const int tid = threadIdx.x;
    const int slot = tid / G;
    const int unit = (tid % G) / U;
    const int *rows;
    const int nrows = row_list(rl, rows);
    const int ntiles = (nrows + RPB - 1) / RPB;
    TileSched ts;
    for (int tile = block_first_tile(rl, ntiles, ts, &s_tile); tile < ntiles;
         tile = block_next(rl, tile, ts, &s_tile)) {
        for (int i = tid; i < RPB * T / 4; i += BS) reinterpret_cast<int4 *>(tab)[i] = make_int4(-1, -1, -1, -1);
        if (NU > 1 && tid < RPB) s_cnt[tid] = 0;
        __syncthreads();

        const int ridx = tile * RPB + slot;
        int cnt = 0;
        int r = 0;
        if (ridx < nrows) {
            r = rows ? rows[ridx] : ridx;
            const int a_beg = arpt[r], a_end = arpt[r + 1];
            const FullRow range{brpt};
            int *stab = tab + slot * T;
            bool narrow = false;
            if constexpr (NU == 1 && BITMAP) {
                // A row whose products span fewer than 32*T columns sets bits
                // in the slot's T words (an OR per product, no CAS, no
                // probing) and counts them by popcount.
                const int gl = tid % G;
                int mn = INT_MAX, mx = -1;
                for (int e = a_beg + gl; e < a_end; e += U) {
                    const int k = acol[e];
                    const int bs = brpt[k], be = brpt[k + 1];
                    if (bs < be) {
                        mn = min(mn, bcol[bs]);
                        mx = max(mx, bcol[be - 1]);
                    }
                }
                mn = unit_min<U>(mn);
                mx = unit_max<U>(mx);
                narrow = mx >= 0 && mx - mn < 32 * T;
                if (narrow) {
                    const int nw = ((mx - mn) >> 5) + 1;
                    for (int i = gl; i < nw; i += U) stab[i] = 0;
                    wave_lds_sync();
                    SymBitmap op{bcol, reinterpret_cast<unsigned *>(stab), mn};
                    for_each_product<U, false, double, true>(a_beg, a_end, unit, NU, acol,
                                                             static_cast<const double *>(nullptr), range, op);
                    wave_lds_sync();
                    for (int i = gl; i < nw; i += U) cnt += __popc(static_cast<unsigned>(stab[i]));
                }
            }
            if (!narrow) {
                SymInsert<T, H> op{bcol, stab};
                for_each_product<U, false, double, true>(a_beg, a_end, unit, NU, acol,
                                                         static_cast<const double *>(nullptr), range, op);
                cnt = op.cnt;
            }
        }
        if constexpr (NU == 1) {
            cnt = unit_sum<U>(cnt);
            if ((tid % G) == 0 && ridx < nrows) row_nnz[r] = cnt;
        } else {
            cnt = unit_sum<kWave>(cnt);
            if (lane_id() == 0) atomicAdd(&s_cnt[0], cnt);
            __syncthreads();
            if (tid == 0 && ridx < nrows) row_nnz[r] = s_cnt[0];
        }
        __syncthreads();  // the next tile reinitialises the tables
    }
}

constexpr int kLongBS = 1024;
constexpr int kBitmapWords = kSymLongWindow / 32;  // 128 KB

// Products per lane in a step of the long-row bitmap walk.
constexpr int kOrPer = 4;
constexpr int kOrChunk = kOrPer * kWave;  // products per step (one chunk)

// Bitmap walk of a staged batch for the long-row symbolic kernel.  A step is
// one chunk of kOrChunk consecutive products of one B range (sorted
// columns), and lane l takes the chunk's products kOrPer*l .. kOrPer*l +
// kOrPer - 1: the bits of a lane's products that fall in one bitmap word are
// ORed in registers and the word gets one LDS OR.  With one product per lane,
// a dense stretch of columns sends the 64 lanes of a step to two or three
// words, and the LDS serialises ORs to one address (on R-MAT 2^20, 52% of the
// kernel's LDS cycles were such address conflicts); here one OR instruction
// reaches at most kOrChunk / 32 + 1 words through 64 / (32 / kOrPer) lanes
// each.  A lane past its range's end reads the range's last entry again, and
// a step past the wave's slice reads the batch's first product: ORing a bit
// that is set already changes nothing, so no product needs a validity flag.
// The chunk's owner range is held in scalar registers (chunk_walk).  NP steps
// are loaded before the previous ones are applied.  Ends with a barrier.
template <int BS>
__device__ __forceinline__ void or4_walk(const ChunkStage<BS, double> &st, int n, int chunks,
                                         const int *__restrict__ bcol, unsigned *bm, int w0)
{
    constexpr int NW = BS / kWave;
    constexpr int NP = 2;
    static_assert(BS <= kWave * kWave, "two-round owner search");
    const int lane = lane_id();
    const int wave = wave_uniform(static_cast<int>(threadIdx.x / kWave));
    chunks = wave_uniform(chunks);
    n = wave_uniform(n);
    const int per = (chunks + NW - 1) / NW;
    const int j0 = wave * per;
    const int j1 = min(chunks, j0 + per);
    if (j0 < j1) {
        const int g1 = (n + kWave - 1) / kWave;
        int q = lane * g1;
        int cnt = __popcll(__ballot(q < n && st.cex[q] <= j0));
        const int lo = (cnt - 1) * g1;
        q = lo + lane;
        cnt = __popcll(__ballot(lane < g1 && q < n && st.cex[q] <= j0));
        int qa = wave_uniform(lo + cnt - 1);
        int c_lo = 0, c_hi = 0, r_bs = 0, r_last = 0;
        auto load_q = [&]() {
            c_lo = wave_uniform(st.cex[qa]);
            c_hi = wave_uniform(st.cex[qa + 1]);
            const typename ChunkStage<BS, double>::Ent en = st.ent[qa];
            r_bs = wave_uniform(en.bs);
            r_last = r_bs + wave_uniform(en.len) - 1;
        };
        load_q();
        const int safe = wave_uniform(st.ent[0].bs);
        auto gather = [&](int jb, int (&c)[NP][kOrPer]) {
#pragma unroll
            for (int s = 0; s < NP; s++) {
                const int jj = jb + s;
                int i0 = safe, last = safe;
                if (jj < j1) {
                    if (jj == c_hi) {
                        qa++;
                        load_q();
                    }
                    i0 = r_bs + (jj - c_lo) * kOrChunk + kOrPer * lane;
                    last = r_last;
                }
#pragma unroll
                for (int k = 0; k < kOrPer; k++) c[s][k] = bcol[min(i0 + k, last)];
            }
        };
        auto apply = [&](const int (&c)[NP][kOrPer]) {
#pragma unroll
            for (int s = 0; s < NP; s++) {
                unsigned acc = 0u;
#pragma unroll
                for (int k = 0; k < kOrPer; k++) {
                    const unsigned cc = static_cast<unsigned>(c[s][k] - w0);
                    acc |= 1u << (cc & 31u);
                    if (k + 1 < kOrPer) {
                        const unsigned nx = static_cast<unsigned>(c[s][k + 1] - w0);
                        if ((nx >> 5) != (cc >> 5)) {
                            atomicOr(&bm[cc >> 5], acc);
                            acc = 0u;
                        }
                    } else {
                        atomicOr(&bm[cc >> 5], acc);
                    }
                }
            }
        };
        int cA[NP][kOrPer], cB[NP][kOrPer];
        bool haveB = false;
        for (int jb = j0;; jb += 2 * NP) {
            gather(jb, cA);
            if (haveB) apply(cB);
            if (jb + NP >= j1) {
                apply(cA);
                break;
            }
            gather(jb + NP, cB);
            haveB = true;
            apply(cA);
            if (jb + 2 * NP >= j1) {
                apply(cB);
                break;
            }
        }
    }
    __syncthreads();
}

// Sets the bits of every product of A row segment [a_beg, a_end) in the
// window's bitmap, in batches of BS A-nonzeros (chunk_stage with kOrChunk
// chunks + or4_walk).
template <int BS, typename Range>
__device__ __forceinline__ void or4_for_each_product(int a_beg, int a_end, const int *__restrict__ acol,
                                                     const Range &range, const int *__restrict__ bcol, unsigned *bm,
                                                     int w0, ChunkStage<BS, double> &st)
{
    for (int c = a_beg; c < a_end; c += BS) {
        int chunks = 0;
        const int n = chunk_stage<BS, double, Range, kOrChunk>(c, a_end, acol, static_cast<const double *>(nullptr),
                                                               range, st, chunks);
        if (chunks > 0) or4_walk<BS>(st, n, chunks, bcol, bm, w0);
    }
}

// B range of an A-nonzero restricted to columns [w0, wend).  For a single
// window covering all columns this is the whole row.  Otherwise *next
// (thread-local) takes the minimum over the calls of the first column at or
// past wend, so the caller can jump over windows no product falls in.
struct WindowRange {
    const int *__restrict__ brpt;
    const int *__restrict__ bcol;
    int w0, wend;
    bool whole;
    int *next;
    __device__ __forceinline__ void operator()(int /*e*/, int k, int &bs, int &len) const
    {
        const int s = brpt[k], t = brpt[k + 1];
        if (whole) {
            bs = s;
            len = t - s;
            return;
        }
        int lo = s, hi = t;
        while (lo < hi) {
            const int m = (lo + hi) >> 1;
            if (bcol[m] < w0)
                lo = m + 1;
            else
                hi = m;
        }
        const int first = lo;
        hi = t;
        while (lo < hi) {
            const int m = (lo + hi) >> 1;
            if (bcol[m] < wend)
                lo = m + 1;
            else
                hi = m;
        }
        bs = first;
        len = lo - first;
        if (lo < t) *next = min(*next, bcol[lo]);
    }
};

__global__ __launch_bounds__(kLongBS) void k_sym_long(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                      const int *__restrict__ brpt, const int *__restrict__ bcol,
                                                      const RowList rl, int ncols, int wcols,
                                                      int *__restrict__ row_nnz, SymSave save)
{
    extern __shared__ unsigned bm[];
    __shared__ int s_red[kLongBS / kWave];
    __shared__ int s_slot;
    __shared__ int s_tile;
    __shared__ ChunkStage<kLongBS, double> s_bst;
    static_assert(sizeof(unsigned) * kBitmapWords + sizeof(ChunkStage<kLongBS, double>) + 256 <= kMaxLds, "LDS");
    const int tid = threadIdx.x;
    const int *rows;
    const int nrows = row_list(rl, rows);
    // The bitmap is cleared once here and then by each window's count pass.
    for (int i = tid; i < (wcols >> 5); i += kLongBS) bm[i] = 0u;
    __syncthreads();
    TileSched ts;
    for (int ridx = block_first_row(rl, nrows, ts, &s_tile); ridx < nrows;
         ridx = block_next(rl, ridx, ts, &s_tile)) {
        const int r = rows ? rows[ridx] : ridx;
        const int a_beg = arpt[r], a_end = arpt[r + 1];
        int cnt = 0;
        for (int w0 = 0; w0 < ncols;) {
            const int wend = static_cast<int>(min(static_cast<long long>(ncols), static_cast<long long>(w0) + wcols));
            const int nwords = (wend - w0 + 31) >> 5;
            int next = INT_MAX;
            const WindowRange range{brpt, bcol, w0, wend, w0 == 0 && wend == ncols, &next};
            drain_vmem();
            or4_for_each_product<kLongBS>(a_beg, a_end, acol, range, bcol, bm, w0, s_bst);
            __syncthreads();
            if (save.bm && save.by_index) {
                // One window (the host keeps bitmaps only then), a slot for
                // every row of the bin: the bitmap is copied out while it is
                // counted and cleared, in one pass, and a row long enough
                // for the numeric long-row kernel records its slot (a row
                // below that leaves its copy unused).
                unsigned *dst = save.bm + static_cast<size_t>(ridx) * nwords;
                int c = 0;
                for (int i = tid; i < nwords; i += kLongBS) {
                    const unsigned x = bm[i];
                    c += __popc(x);
                    dst[i] = x;
                    bm[i] = 0u;
                }
                const int tot = block_sum<kLongBS>(c, s_red);
                if (tid == 0) {
                    if (tot > save.min_nnz) {
                        save.slot[r] = ridx;
                        atomicAdd(save.count, 1);
                    }
                    cnt = tot;
                }
            } else if (save.bm) {
                // One window (the host keeps bitmaps only then): a row long
                // enough for the numeric long-row kernel leaves its bitmap
                // in global memory, copied out while it is cleared.
                int c = 0;
                for (int i = tid; i < nwords; i += kLongBS) c += __popc(bm[i]);
                const int tot = block_sum<kLongBS>(c, s_red);
                if (tid == 0) {
                    int sl = -1;
                    if (tot > save.min_nnz) {
                        sl = atomicAdd(save.count, 1);
                        if (sl < save.cap) save.slot[r] = sl;
                        else sl = -1;
                    }
                    s_slot = sl;
                }
                __syncthreads();
                if (s_slot >= 0) {
                    unsigned *dst = save.bm + static_cast<size_t>(s_slot) * nwords;
                    for (int i = tid; i < nwords; i += kLongBS) {
                        dst[i] = bm[i];
                        bm[i] = 0u;
                    }
                } else {
                    for (int i = tid; i < nwords; i += kLongBS) bm[i] = 0u;
                }
                if (tid == 0) cnt = tot;
            } else {
                for (int i = tid; i < nwords; i += kLongBS) {
                    cnt += __popc(bm[i]);
                    bm[i] = 0u;
                }
            }
            __syncthreads();
            if (wend == ncols) break;
            // Next window holding a product of this row: windows with none
            // cost nothing, so the row's cost follows its products, not B.N.
            next = block_min<kLongBS>(next, s_red);
            if (next == INT_MAX) break;
            w0 = (next / wcols) * wcols;
        }
        const int tot = save.bm ? cnt : block_sum<kLongBS>(cnt, s_red);
        if (tid == 0) row_nnz[r] = tot;
    }
}

// Short-row bins (G <= 64) take the hash family chosen by the tuning; wide
// rows (bins 5, 6) always hash with the Fibonacci family.
template <int G, int T, int BS>
void launch_hash(const CsrIn &A, const CsrIn &B, const RowList &rl, int *row_nnz, int hash, bool bitmap,
                 hipStream_t s)
{
    constexpr int RPB = BS / G;
    const int grid = tile_grid(rl, RPB, kResidentThreads / BS);
    if constexpr (G <= kWave) {
        if (bitmap) {
            if (hash == kHashLin)
                k_sym_hash<G, T, BS, kHashLin, true>
                    <<<grid, BS, 0, s>>>(A.rpt, A.col, B.rpt, B.col, rl, row_nnz);
            else
                k_sym_hash<G, T, BS, kHashFib, true>
                    <<<grid, BS, 0, s>>>(A.rpt, A.col, B.rpt, B.col, rl, row_nnz);
            return;
        }
        if (hash == kHashLin) {
            k_sym_hash<G, T, BS, kHashLin, false>
                <<<grid, BS, 0, s>>>(A.rpt, A.col, B.rpt, B.col, rl, row_nnz);
            return;
        }
    }
    k_sym_hash<G, T, BS, kHashFib, false><<<grid, BS, 0, s>>>(A.rpt, A.col, B.rpt, B.col, rl, row_nnz);
}

}  // namespace

size_t symbolic_long_lds() { return sizeof(unsigned) * kBitmapWords; }

void symbolic_init_attributes()
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_sym_long), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(symbolic_long_lds()));
}

void launch_symbolic_bin(int bin, const CsrIn &A, const CsrIn &B, const RowList &rl, int *row_nnz, int hash,
                         bool bitmap, const SymSave &save, hipStream_t s, int shape)
{
    if (rl.n <= 0) return;
    // bins 0-4: shape 0 default, 1 twice the table, 2 other lanes per row or block size
    const int k = bin < 5 ? bin * kShapes + shape : -1;
    switch (k) {
        case 0: launch_hash<16, 32, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 1: launch_hash<16, 64, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 2: launch_hash<32, 64, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 3: launch_hash<32, 64, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 4: launch_hash<32, 128, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 5: launch_hash<64, 128, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 6: launch_hash<64, 128, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 7: launch_hash<64, 256, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 8: launch_hash<32, 128, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 9: launch_hash<64, 512, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 10: launch_hash<64, 1024, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 11: launch_hash<64, 512, 64>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 12: launch_hash<64, 1024, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 13: launch_hash<64, 2048, 256>(A, B, rl, row_nnz, hash, bitmap, s); return;
        case 14: launch_hash<64, 1024, 64>(A, B, rl, row_nnz, hash, bitmap, s); return;
        default: break;
    }
    switch (bin) {
        case 5: launch_hash<256, 4096, 256>(A, B, rl, row_nnz, hash, false, s); break;
        case 6: launch_hash<1024, 16384, 1024>(A, B, rl, row_nnz, hash, false, s); break;
        default: {
            // Bitmap sized to the column count (up to 2^20 columns per window),
            // so narrow products keep more blocks resident per CU.
            const int words = static_cast<int>(std::min<long long>(kBitmapWords, (static_cast<long long>(B.N) + 31) / 32));
            const int wcols = words * 32;
            k_sym_long<<<row_grid(rl, kNumCu), kLongBS, sizeof(unsigned) * static_cast<size_t>(words), s>>>(
                A.rpt, A.col, B.rpt, B.col, rl, B.N, wcols, row_nnz,
                B.N <= kSymLongWindow ? save : SymSave{});
            break;
        }
    }
}

}  // namespace opsamd
