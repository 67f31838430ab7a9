// Numeric phase: values and sorted columns of every row of C.
//
// Reference counterparts: k_numeric_shared_hash_pwarp,
// k_numeric_shared_hash_tb_full_occu<SH_ROW,BS>,
// k_numeric_max_shared_hash_tb_half_occu, k_numeric_global_hash_tb_full_occu
// (include/opsparse/core_algo/spgemm/kernels/numeric.cuh:9-393) and their
// dispatch h_numeric_full_occu (src/opsparse/core_algo/spgemm/spgemm_phases.cu:148-206).
//
// Rows up to 4096 entries accumulate in an open-addressing LDS hash (CAS on
// the key, atomic add of a*b on the value), with tables of >= 2n slots for
// a row of n entries, like the reference.  The ordering step is different:
// the reference ranks every entry against every other entry, O(n^2)
// (numeric.cuh:97-106, 194-204, 382-392); here rows of up to 256 entries
// are compacted and bitonic-sorted in registers across their unit
// (cross-lane ds_bpermute) and rows up to 4096 are bitonic-sorted across
// the block (registers and lane exchanges, LDS for strides across waves).
// Rows above 4096 entries use no hash and no sort: a bitmap of the row's
// columns per 2^17-column window gives every column its rank (its position
// in the sorted output), values accumulate at that rank in LDS, and the
// window is written out in order (k_num_rank).
#include <algorithm>
#include <cstdlib>

#include "device.hpp"
#include "launch.hpp"

namespace opsamd {

namespace {

// Consumer of the product walk (device.hpp): accumulate a*b into the row's
// hash table.
template <int T, int H, typename V>
struct NumInsert {
    const int *__restrict__ bcol;
    const V *__restrict__ bval;
    int *kt;
    Acc *vt;
    struct Item {
        int c;
        V b;
    };
    __device__ __forceinline__ Item load(int k) const { return Item{bcol[k], bval[k]}; }
    __device__ __forceinline__ void apply(const Item &it, V a, int)
    {
        bool fresh;
        const int h = hash_insert<T, H>(kt, it.c, fresh);
        atomicAdd(vt + h, static_cast<Acc>(a) * static_cast<Acc>(it.b));
    }
};

// Bins 0-4: G <= 64 lanes per row, BS/G rows per block, register sort.
// T table slots, at most NMAX entries per row (T >= 2 * NMAX).
template <int G, int T, int NMAX, int BS, int H, typename V>
__global__ __launch_bounds__(BS) void k_num_small(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                  const V *__restrict__ aval, const int *__restrict__ brpt,
                                                  const int *__restrict__ bcol, const V *__restrict__ bval,
                                                  const RowList rl, const long long *__restrict__ crpt,
                                                  int *__restrict__ ccol, V *__restrict__ cval, bool pack)
{
    constexpr int U = G;
    constexpr int RPB = BS / G;
    constexpr int S = T / U;     // table slots per lane
    constexpr int K = NMAX / U;  // sorted entries per lane
    static_assert(G <= kWave && BS % G == 0 && K >= 1 && T >= 2 * NMAX, "bad config");
    __shared__ __align__(16) int keys[RPB * T];
    __shared__ __align__(16) Acc vals[RPB * T];
    __shared__ int ckeys[RPB * NMAX];
    __shared__ int s_tile;

    const int tid = threadIdx.x;
    const int slot = tid / G;
    const int gl = tid % G;
    const int *rows;
    const int nrows = row_list(rl, rows);
    const int ntiles = (nrows + RPB - 1) / RPB;
    TileSched ts;
    for (int tile = block_first_tile(rl, ntiles, ts, &s_tile); tile < ntiles;
         tile = block_next(rl, tile, ts, &s_tile)) {
        for (int i = tid; i < RPB * T / 4; i += BS) reinterpret_cast<int4 *>(keys)[i] = make_int4(-1, -1, -1, -1);
        for (int i = tid; i < RPB * T * static_cast<int>(sizeof(Acc)) / 16; i += BS)
            reinterpret_cast<int4 *>(vals)[i] = make_int4(0, 0, 0, 0);
        __syncthreads();

        const int ridx = tile * RPB + slot;
        int r = 0;
        int *kt = keys + slot * T;
        Acc *vt = vals + slot * T;
        if (ridx < nrows) {
            r = rows ? rows[ridx] : ridx;
            const FullRow range{brpt};
            NumInsert<T, H, V> op{bcol, bval, kt, vt};
            for_each_product<U, true, V>(arpt[r], arpt[r + 1], 0, 1, acol, aval, range, op);
        }
        __syncthreads();
        if (ridx < nrows) {
            // Compact the row's keys (slot order) into its key list; the table stays
            // intact and supplies the values after the keys are sorted.  With pack
            // (every column below 2^(31 - log2 T)) a key carries its slot in its low
            // bits, so the sorted list gives each value's slot without a lookup.
            constexpr int LOGT = __builtin_ctz(T);
            int *ck = ckeys + slot * NMAX;
            int n = 0;
            const unsigned long long below = gl ? ((1ull << gl) - 1ull) : 0ull;
#pragma unroll
            for (int s = 0; s < S; s++) {
                const int key = kt[s * U + gl];
                const bool valid = key != -1;
                const unsigned long long m = unit_ballot<U>(valid);
                const int pos = n + __popcll(m & below);
                if (valid && pos < NMAX) ck[pos] = pack ? ((key << LOGT) | (s * U + gl)) : key;
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
                    const int key = pack ? (sk[k] >> LOGT) : sk[k];
                    const int h = pack ? (sk[k] & (T - 1)) : hash_find<T, H>(kt, sk[k]);
                    ccol[off + j] = key;
                    cval[off + j] = h >= 0 ? static_cast<V>(vt[h]) : V(0);
                }
            }
        }
        __syncthreads();  // the next tile reinitialises the tables
    }
}

// The window's bitmap words and their ranks are interleaved (x: bits,
// y: number of set bits in the words before), so the rank of a column is
// one 64-bit LDS read.
// Step 1 consumer: set the product's column bit.
// The OR returns nothing, so no product waits on an LDS result.
template <typename V>
struct RankMark {
    const int *__restrict__ bcol;
    uint2 *br;
    int w0;
    using Item = int;
    __device__ __forceinline__ int load(int k) const { return bcol[k]; }
    __device__ __forceinline__ void apply(int col, V, int)
    {
        const int c = col - w0;
        atomicOr(&br[c >> 5].x, 1u << (c & 31));
    }
};

// Step 3 consumer: add a*b at the column's rank relative to r0, for ranks
// in [r0, r0 + nr) (products of other sub-windows are skipped).
template <typename V>
struct RankAcc {
    const int *__restrict__ bcol;
    const V *__restrict__ bval;
    const uint2 *br;
    Acc *vals;
    int w0;
    int r0;
    unsigned nr;
    struct Item {
        int c;
        V b;
    };
    struct Post {
        unsigned at;  // >= nr: nothing to add
        Acc v;
    };
    __device__ __forceinline__ Item load(int k) const { return Item{bcol[k], bval[k]}; }
    __device__ __forceinline__ uint2 prep(const Item &it) const { return br[(it.c - w0) >> 5]; }
    __device__ __forceinline__ Post stage(const Item &it, const uint2 &e, V a, int, bool ok) const
    {
        const int c = it.c - w0;
        const unsigned at = e.y + __popc(e.x & ((1u << (c & 31)) - 1u)) - static_cast<unsigned>(r0);
        return Post{ok ? at : nr, static_cast<Acc>(a) * static_cast<Acc>(it.b)};
    }
    __device__ __forceinline__ void post(const Post &p)
    {
        if (p.at < nr) atomicAdd(vals + p.at, p.v);
    }
};

// Element i of an array addressed by a 32-bit byte offset from a base held in
// scalar registers (global_load with saddr): one shift per load instead of
// the sign extension and 64-bit add of a pointer index.  The caller
// guarantees i * sizeof(T) < 2^32.
template <typename T>
__device__ __forceinline__ T load_off32(const T *__restrict__ base, unsigned i)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * static_cast<unsigned>(sizeof(T)));
}

// The bitmap-rank accumulate walk of a staged batch: the same split of the
// batch's products over the block's waves as block_walk (wave w takes one
// contiguous slice, 64 consecutive products per step, NP steps gathered at
// once in two ping-pong buffers), written for the fewest vector
// instructions per step, which is what bounds this walk: every quantity
// that is the same for the whole wave (slice bounds, the current range,
// its B offset and A value) is held in scalar registers, so a step inside
// one range costs one vector add for its B indices; B entries are read with
// 32-bit offsets from scalar bases (OFF32: B holds fewer than 2^29
// entries); and the rank of a product is one bit-field extract, one bit
// count and one three-way add on the window word read from LDS.  Only a
// step that crosses range starts finds each lane's range (the start mask
// of block_walk).  Ends with a barrier.
template <int BS, bool OFF32, typename V>
__device__ __forceinline__ void rank_walk(const BlockStage<BS, V> &st, int n, int total, const RankAcc<V> &op)
{
    constexpr int NW = BS / kWave;
    constexpr int NP = 4;
    static_assert(BS <= kWave * kWave, "two-round owner search");
    const int lane = lane_id();
    const int wave = wave_uniform(static_cast<int>(threadIdx.x / kWave));
    n = wave_uniform(n);
    total = wave_uniform(total);
    const int seg = (((total + NW - 1) / NW) + kWave - 1) & ~(kWave - 1);
    const int s_beg = wave * seg;
    const int s_end = min(total, s_beg + seg);
    if (s_beg < s_end) {
        const int g1 = (n + kWave - 1) / kWave;
        int q = lane * g1;
        int cnt = __popcll(__ballot(q < n && st.excl[q] <= s_beg));
        const int lo = (cnt - 1) * g1;
        q = lo + lane;
        cnt = __popcll(__ballot(lane < g1 && q < n && st.excl[q] <= s_beg));
        int qa = wave_uniform(lo + cnt - 1);
        const unsigned long long below = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull);
        // range qa in scalar registers: where the next range starts, its B
        // index offset (B index = product position + offset) and A value
        int q_next = 0, q_off = 0;
        Acc q_av = Acc(0);
        auto load_q = [&]() {
            q_next = wave_uniform(qa + 1 <= n ? st.excl[qa + 1] : INT_MAX);
            if (qa < n) {
                const typename BlockStage<BS, V>::Ent en = st.ent[qa];
                q_off = wave_uniform(en.off);
                q_av = wave_uniform(static_cast<Acc>(en.av));
            }
        };
        load_q();
        // lanes past the slice's end read the batch's first product (a
        // valid entry inside the window) and add nothing
        const int safe = wave_uniform(st.ent[0].off);
        const int w0 = op.w0;
        const unsigned rbias = static_cast<unsigned>(-op.r0);
        const unsigned nr = op.nr;
        auto gather = [&](int b0, int (&bi)[NP], Acc (&aa)[NP], bool (&ok)[NP]) {
#pragma unroll
            for (int s = 0; s < NP; s++) {
                const int sb = b0 + s * kWave;
                ok[s] = false;
                bi[s] = safe;
                aa[s] = Acc(0);
                if (sb < s_end) {
                    const int p = sb + lane;
                    ok[s] = p < s_end;
                    if (sb + kWave <= q_next) {
                        bi[s] = ok[s] ? p + q_off : safe;
                        aa[s] = q_av;
                        if (sb + kWave == q_next) {
                            qa++;
                            load_q();
                        }
                    } else {
                        const int qc = qa + 1 + lane;
                        const int pos = qc <= n ? st.excl[qc] - sb : kWave + 1;
                        const unsigned lo32 = pos > 0 && pos < 32 ? 1u << pos : 0u;
                        const unsigned hi32 = pos >= 32 && pos < kWave ? 1u << (pos - 32) : 0u;
                        const unsigned long long m =
                            (static_cast<unsigned long long>(wave_or(hi32)) << 32) | wave_or(lo32);
                        const int qo = qa + __popcll(m & below);
                        qa = wave_uniform(static_cast<int>(qa + __popcll(m) + (__ballot(pos == kWave) != 0ull ? 1 : 0)));
                        const typename BlockStage<BS, V>::Ent en = st.ent[min(qo, n - 1)];
                        bi[s] = ok[s] ? p + en.off : safe;
                        aa[s] = static_cast<Acc>(en.av);
                        load_q();
                    }
                }
            }
        };
        auto load = [&](const int (&bi)[NP], int (&c)[NP], V (&b)[NP]) {
#pragma unroll
            for (int s = 0; s < NP; s++) {
                if constexpr (OFF32) {
                    c[s] = load_off32(op.bcol, static_cast<unsigned>(bi[s]));
                    b[s] = load_off32(op.bval, static_cast<unsigned>(bi[s]));
                } else {
                    c[s] = op.bcol[bi[s]];
                    b[s] = op.bval[bi[s]];
                }
            }
        };
        auto apply = [&](const int (&c)[NP], const V (&b)[NP], const Acc (&aa)[NP], const bool (&ok)[NP]) {
            uint2 wd[NP];
            unsigned cc[NP];
#pragma unroll
            for (int s = 0; s < NP; s++) {
                cc[s] = static_cast<unsigned>(c[s] - w0);
                wd[s] = op.br[cc[s] >> 5];
            }
#pragma unroll
            for (int s = 0; s < NP; s++) {
                const unsigned at = __popc(__builtin_amdgcn_ubfe(wd[s].x, 0u, cc[s] & 31u)) + wd[s].y + rbias;
                if (ok[s] && at < nr) atomicAdd(op.vals + at, aa[s] * static_cast<Acc>(b[s]));
            }
        };
        constexpr int stride = NP * kWave;
        int biA[NP], biB[NP], cA[NP], cB[NP];
        Acc aaA[NP], aaB[NP];
        V bA[NP], bB[NP];
        bool okA[NP], okB[NP];
#pragma unroll
        for (int s = 0; s < NP; s++) {
            okB[s] = false;
            aaB[s] = Acc(0);
            cB[s] = w0;
            bB[s] = V(0);
        }
        for (int base = s_beg;; base += 2 * stride) {
            gather(base, biA, aaA, okA);
            load(biA, cA, bA);
            apply(cB, bB, aaB, okB);
            gather(base + stride, biB, aaB, okB);
            load(biB, cB, bB);
            apply(cA, bA, aaA, okA);
            if (base + 2 * stride >= s_end) break;
        }
        apply(cB, bB, aaB, okB);
    }
    __syncthreads();
}

// Every product of A row segment [a_beg, a_end) through rank_walk, in
// batches of BS A-nonzeros.  Returns true when the row fit in one batch:
// `st` then still holds it for another walk (rank_walk with n and total).
template <int BS, bool OFF32, typename V, typename Range>
__device__ __forceinline__ bool rank_for_each_product(int a_beg, int a_end, const int *__restrict__ acol,
                                                      const V *__restrict__ aval, const Range &range,
                                                      const RankAcc<V> &op, BlockStage<BS, V> &st, int &n, int &total)
{
    n = 0;
    total = 0;
    for (int c = a_beg; c < a_end; c += BS) {
        n = block_stage<BS, V>(c, a_end, acol, aval, range, st, total);
        if (total > 0) rank_walk<BS, OFF32, V>(st, n, total, op);
    }
    return a_end - a_beg <= BS;
}

// Bins 5-6: one row per block of BS threads, at most T/2 entries per row.
// Rows whose columns span fewer than 32*T columns take the bitmap-rank path
// of the long-row kernel in one window (a bitmap word per 32 columns from
// the row's smallest column; no hash, no sort).  Other rows use a table of
// T slots in dynamic LDS, then compaction of the keys and a block bitonic
// sort of the keys (in registers and lane exchanges, through LDS only for
// strides across waves); the values are read back from the table by key.
template <int T, int BS, typename V>
__global__ __launch_bounds__(BS) void k_num_block(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                  const V *__restrict__ aval, const int *__restrict__ brpt,
                                                  const int *__restrict__ bcol, const V *__restrict__ bval,
                                                  const RowList rl, const long long *__restrict__ crpt,
                                                  int *__restrict__ ccol, V *__restrict__ cval)
{
    extern __shared__ unsigned long long smem_u64[];
    const int *rows;
    const int nrows = row_list(rl, rows);
    // Hash path: T values | T keys | T/2 sorted keys.
    Acc *vt = reinterpret_cast<Acc *>(smem_u64);
    int *kt = reinterpret_cast<int *>(vt + T);
    int *sk = kt + T;
    // Rank path over the same memory: T bitmap words with ranks | T/2 values.
    uint2 *br = reinterpret_cast<uint2 *>(smem_u64);
    Acc *rv = reinterpret_cast<Acc *>(br + T);
    static_assert(sizeof(uint2) * T + sizeof(Acc) * (T / 2) <= sizeof(Acc) * T + sizeof(int) * (T + T / 2),
                  "rank path must fit in the hash path's LDS");
    __shared__ int s_pos;
    __shared__ int s_tile;
    __shared__ int s_scan[BS / kWave + 1];
    __shared__ int s_mm[2 * (BS / kWave)];
    __shared__ BlockStage<BS, V> s_bst;
    constexpr int NU = BS / kWave;
    constexpr int KPT = T / 2 / BS;  // sorted keys per thread on the hash path
    static_assert(KPT >= 1 && KPT * BS == T / 2, "sort layout");
    constexpr int WPW = T / NU;  // bitmap words per wave in the rank scan
    constexpr int RPW = WPW / kWave;
    static_assert(T % (NU * kWave) == 0, "words per wave");
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int lane = lane_id();
    const FullRow range{brpt};

    TileSched ts;
    for (int ridx = block_first_row(rl, nrows, ts, &s_tile); ridx < nrows;
         ridx = block_next(rl, ridx, ts, &s_tile)) {
        // The previous row's output stores complete here, so the walks'
        // loads are waited for one buffer at a time.
        drain_vmem();
        const int r = rows ? rows[ridx] : ridx;
        const int a_beg = arpt[r], a_end = arpt[r + 1];
        const long long off = crpt[r];
        const int n = min(static_cast<int>(crpt[r + 1] - off), T / 2);  // rows of this bin have n <= T/2
        // Column span of the row: first and last columns of its B rows.
        int mn = INT_MAX, mxc = -1;
        for (int e = a_beg + tid; e < a_end; e += BS) {
            const int k = acol[e];
            const int b0 = brpt[k], b1 = brpt[k + 1];
            if (b0 < b1) {
                mn = min(mn, bcol[b0]);
                mxc = max(mxc, bcol[b1 - 1]);
            }
        }
        mn = unit_min<kWave>(mn);
        mxc = unit_max<kWave>(mxc);
        if (lane_id() == 0) {
            s_mm[wave] = mn;
            s_mm[NU + wave] = mxc;
        }
        __syncthreads();
        mn = INT_MAX;
        mxc = -1;
        for (int w = 0; w < NU; w++) {
            mn = min(mn, s_mm[w]);
            mxc = max(mxc, s_mm[NU + w]);
        }
        if (mxc >= 0 && mxc - mn < 32 * T) {
            const int nw = ((mxc - mn) >> 5) + 1;
            for (int i = tid; i < nw; i += BS) br[i] = make_uint2(0u, 0u);
            __syncthreads();
            int sn, stot;
            bool one;
            {
                RankMark<V> op{bcol, br, mn};
                one = block_for_each_product<BS, false, V>(a_beg, a_end, acol, aval, range, op, s_bst, sn, stot);
            }
            __syncthreads();
            // Ranks: wave w owns words [w * WPW, (w + 1) * WPW), its lanes
            // take 64 consecutive words at a time (conflict-free reads);
            // DPP scans of the word popcounts, wave totals through LDS.
            int cnt[RPW], inc[RPW];
            int wtot = 0;
#pragma unroll
            for (int i = 0; i < RPW; i++) {
                const int w = wave * WPW + i * kWave + lane;
                cnt[i] = w < nw ? __popc(br[w].x) : 0;
                inc[i] = wave_incl_scan(cnt[i]);
                wtot += __builtin_amdgcn_readlane(inc[i], kWave - 1);
            }
            if (lane == 0) s_scan[wave] = wtot;
            __syncthreads();
            int run = 0;
            for (int w = 0; w < wave; w++) run += s_scan[w];
#pragma unroll
            for (int i = 0; i < RPW; i++) {
                const int w = wave * WPW + i * kWave + lane;
                if (w < nw) br[w].y = static_cast<unsigned>(run + inc[i] - cnt[i]);
                run += __builtin_amdgcn_readlane(inc[i], kWave - 1);
            }
            for (int i = tid; i < n; i += BS) rv[i] = Acc(0);
            __syncthreads();
            {
                RankAcc<V> op{bcol, bval, br, rv, mn, 0, static_cast<unsigned>(n)};
                // A row that fit in one batch is still staged from the mark.
                if (one) {
                    if (stot > 0) block_walk<BS, true, V>(s_bst, sn, stot, op);
                } else {
                    block_for_each_product<BS, true, V>(a_beg, a_end, acol, aval, range, op, s_bst, sn, stot);
                }
            }
            __syncthreads();
            for (int i = tid; i < n; i += BS) cval[off + i] = static_cast<V>(rv[i]);
            __syncthreads();
            int *scol = reinterpret_cast<int *>(rv);
            for (int w = tid; w < nw; w += BS) {
                const uint2 ew = br[w];
                unsigned bits = ew.x;
                while (bits) {
                    const int bit = __ffs(bits) - 1;
                    bits &= bits - 1u;
                    const int rr = static_cast<int>(ew.y) + __popc(ew.x & ((1u << bit) - 1u));
                    if (rr < n) scol[rr] = mn + w * 32 + bit;
                }
            }
            __syncthreads();
            for (int i = tid; i < n; i += BS) ccol[off + i] = scol[i];
            __syncthreads();
            continue;
        }
        for (int i = tid; i < T; i += BS) {
            kt[i] = -1;
            vt[i] = Acc(0);
        }
        if (tid == 0) s_pos = 0;
        __syncthreads();
        NumInsert<T, kHashFib, V> op{bcol, bval, kt, vt};
        int sn, stot;
        block_for_each_product<BS, true, V>(a_beg, a_end, acol, aval, range, op, s_bst, sn, stot);
        __syncthreads();
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
        __syncthreads();
        // Sort the row's keys (padded to T/2 with INT_MAX) with KPT per
        // thread in registers; key j then goes to output position j.
        const int sbase = wave * kWave * KPT;
        int key[KPT];
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int j = sbase + k * kWave + lane_id();
            key[k] = j < n ? sk[j] : INT_MAX;
        }
        __syncthreads();
        block_bitonic_sort_keys<BS, KPT>(key, sk);
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int j = sbase + k * kWave + lane_id();
            if (j < n) {
                const int h = hash_find<T, kHashFib>(kt, key[k]);
                ccol[off + j] = key[k];
                cval[off + j] = h >= 0 ? static_cast<V>(vt[h]) : V(0);
            }
        }
        __syncthreads();
    }
}

// Bin 5 by wave: one row per wave, the bitmap-rank path of the block bins
// without block barriers.  Each wave has W bitmap words (a column span of
// up to 32*W from the row's smallest column) and NMAX values in LDS.  Rows
// with a wider span are appended to spill (spill[0]: count, then row ids)
// for the block kernel, which runs next on the same stream.
template <int NMAX, int W, typename V>
struct WaveRankLds {
    uint2 br[W];
    Acc vals[NMAX];
};

template <int NMAX, int W, int WPB, typename V>
__global__ __launch_bounds__(WPB * kWave) void k_num_wrank(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                           const V *__restrict__ aval, const int *__restrict__ brpt,
                                                           const int *__restrict__ bcol, const V *__restrict__ bval,
                                                           const RowList rl, const long long *__restrict__ crpt,
                                                           int *__restrict__ ccol, V *__restrict__ cval,
                                                           int *__restrict__ spill)
{
    constexpr int WPL = W / kWave;  // bitmap words per lane in the rank scan
    static_assert(W % kWave == 0 && NMAX * sizeof(Acc) >= NMAX * sizeof(int), "layout");
    __shared__ WaveRankLds<NMAX, W, V> s_lds[WPB];
    __shared__ int s_tile;
    const int wave = threadIdx.x / kWave;
    const int lane = lane_id();
    WaveRankLds<NMAX, W, V> &L = s_lds[wave];
    // Each wave takes two consecutive rows of the list.  The loads that
    // start a row (row id, A row, B row offsets, first and last B columns:
    // four dependent steps) are issued for both rows together, so the second
    // row's chain overlaps the first's; the rows are then processed in turn.
    const int *rows;
    const int nrows = row_list(rl, rows);
    const int ntiles = (nrows + 2 * WPB - 1) / (2 * WPB);
    struct Start {
        int r, a_beg, a_end, cbs, clen, mn, mx;
        V cav;
    };
    // A row of at most 64 A-nonzeros is loaded once: lane l keeps its
    // A-nonzero's B range and value for the span check and both walks.
    auto row_start = [&](int ridx, Start &st) {
        st.r = rows ? rows[ridx] : ridx;
        st.a_beg = arpt[st.r];
        st.a_end = arpt[st.r + 1];
        st.cbs = 0;
        st.clen = 0;
        st.cav = V(0);
        st.mn = INT_MAX;
        st.mx = -1;
        const int e = st.a_beg + lane;
        if (e < st.a_end) {
            const int k = acol[e];
            const int b0 = brpt[k], b1 = brpt[k + 1];
            st.cbs = b0;
            st.clen = b1 - b0;
            st.cav = aval[e];
            if (b0 < b1) {
                st.mn = bcol[b0];
                st.mx = bcol[b1 - 1];
            }
        }
    };
    auto row_span_rest = [&](Start &st) {
        for (int e = st.a_beg + kWave + lane; e < st.a_end; e += kWave) {
            const int k = acol[e];
            const int b0 = brpt[k], b1 = brpt[k + 1];
            if (b0 < b1) {
                st.mn = min(st.mn, bcol[b0]);
                st.mx = max(st.mx, bcol[b1 - 1]);
            }
        }
        st.mn = unit_min<kWave>(st.mn);
        st.mx = unit_max<kWave>(st.mx);
    };
    auto process = [&](const Start &st) {
        const int r = st.r, a_beg = st.a_beg, a_end = st.a_end;
        const int mn = st.mn, mx = st.mx;
        const bool one = a_end - a_beg <= kWave;
        const int cnt = min(kWave, a_end - a_beg);
        const int cbs = st.cbs, clen = st.clen;
        const V cav = st.cav;
        if (mx < 0) return;  // no products: the row is empty
        if (mx - mn >= 32 * W) {
            if (lane == 0) spill[2 + atomicAdd(spill, 1)] = r;
            return;
        }
        const int nw = ((mx - mn) >> 5) + 1;
        for (int i = lane; i < nw; i += kWave) L.br[i] = make_uint2(0u, 0u);
        wave_lds_sync();
        const FullRow range{brpt};
        {
            RankMark<V> op{bcol, L.br, mn};
            if (one)
                walk_chunk<kWave, false, V>(cnt, cbs, clen, cav, op, 0, 1);
            else
                for_each_product<kWave, false, V>(a_beg, a_end, 0, 1, acol, aval, range, op);
        }
        wave_lds_sync();
        // Ranks: lanes take 64 consecutive words at a time (conflict-free
        // reads), DPP scans of the word popcounts.
        int run = 0;
#pragma unroll
        for (int i = 0; i < WPL; i++) {
            const int w = i * kWave + lane;
            const int c = w < nw ? __popc(L.br[w].x) : 0;
            const int inc = wave_incl_scan(c);
            if (w < nw) L.br[w].y = static_cast<unsigned>(run + inc - c);
            run += __builtin_amdgcn_readlane(inc, kWave - 1);
        }
        const long long off = crpt[r];
        const int n = min(static_cast<int>(crpt[r + 1] - off), NMAX);
        for (int i = lane; i < n; i += kWave) L.vals[i] = Acc(0);
        wave_lds_sync();
        {
            RankAcc<V> op{bcol, bval, L.br, L.vals, mn, 0, static_cast<unsigned>(n)};
            if (one)
                walk_chunk<kWave, true, V>(cnt, cbs, clen, cav, op, 0, 1);
            else
                for_each_product<kWave, true, V>(a_beg, a_end, 0, 1, acol, aval, range, op);
        }
        wave_lds_sync();
        for (int i = lane; i < n; i += kWave) cval[off + i] = static_cast<V>(L.vals[i]);
        wave_lds_sync();
        int *scol = reinterpret_cast<int *>(L.vals);
        for (int w = lane; w < nw; w += kWave) {
            const uint2 ew = L.br[w];
            unsigned bits = ew.x;
            while (bits) {
                const int bit = __ffs(bits) - 1;
                bits &= bits - 1u;
                const int rr = static_cast<int>(ew.y) + __popc(ew.x & ((1u << bit) - 1u));
                if (rr < n) scol[rr] = mn + w * 32 + bit;
            }
        }
        wave_lds_sync();
        for (int i = lane; i < n; i += kWave) ccol[off + i] = scol[i];
        wave_lds_sync();  // the next row rewrites this wave's LDS
    };
    auto pair = [&](int ridx0) {
        const bool two = ridx0 + 1 < nrows;
        Start s0, s1;
        row_start(ridx0, s0);
        row_start(two ? ridx0 + 1 : ridx0, s1);
        row_span_rest(s0);
        row_span_rest(s1);
        process(s0);
        if (two) process(s1);
    };
    // A tile is WPB pairs of rows, one per wave; the block takes tiles from
    // the counter (device-sized lists, a barrier per tile) or by the
    // XCD-aware stride.
    TileSched ts;
    for (int tile = block_first_tile(rl, ntiles, ts, &s_tile); tile < ntiles;
         tile = block_next(rl, tile, ts, &s_tile)) {
        const int ridx0 = (tile * WPB + wave) * 2;
        if (ridx0 < nrows) pair(ridx0);
    }
}

// Clears a spill list's count and work counter.
__global__ void k_zero_count(int *p)
{
    if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

constexpr int kLongBS = 1024;

// Bin 7: rows above 4096 entries, bitmap-rank windows.  The columns are
// split into windows of WB = 2^17; per window:
//  1. mark: every product of the window sets its column's bit in an LDS
//     bitmap (read first, atomic OR only for a clear bit).  A row whose
//     column bitmap the symbolic phase kept (ls.bm) reads the window's
//     words from it instead and walks its products only in step 3;
//  2. rank: a scan of the per-word popcounts (DPP inside each wave, wave
//     totals through LDS) gives each word the number of the row's columns
//     before it, so the position of column c in the window's sorted output
//     is rk[word] + popcount(lower bits);
//  3. accumulate: the products are walked again and a*b is added to a
//     compact LDS value array at that position.  When the window holds more
//     than VCAP columns, this pass runs over column sub-windows of at most
//     VCAP columns (cut at the column of rank r0 + VCAP).  Each sub-window
//     locates its own B ranges with cursors (the end in a B row searched
//     between the cuts of the unit cut table when there is one), so every
//     product is visited once; without the unit cut table, a window of two
//     sub-windows instead walks twice over the B ranges step 1 found (still
//     staged in LDS when the A row fits one batch), skipping the ranks of
//     the other sub-window;
//  4. write: values (already in rank order) and columns (placed by rank in
//     LDS) are copied out contiguously.
// No sort is needed and the value array only holds columns that exist, so a
// row of 30K entries over 1M columns takes 8 bitmap windows instead of 64
// dense ones.  Windows with no product are skipped (the next window starts
// at the smallest column beyond the current one).  Walks are block-wide
// (block_for_each_product): each batch of A-nonzeros has its B ranges
// computed once, one per thread.
//
// A window's B ranges come from the window-cut table of B when one is given
// (ls.wcut: the first entry of every B row at every window boundary);
// otherwise from cursor[], four arrays of span ints (span = highest
// A-nonzero index + 1): two hold the window-start positions of the
// A-nonzeros' B rows and alternate between windows (the end found by
// galloping), two alternate between the sub-windows of step 3 (with or
// without the table).
// Shapes: 1024-thread blocks with 12288 values per pass (one block per CU),
// or 512-thread blocks with 4608 (two blocks per CU, for rows of at most
// kLightRowMax entries, whose windows mostly fit one pass: two rows in
// flight per CU hide each other's per-window waits).
template <typename V, int BS = kLongBS>
struct RankWin {
    static constexpr int WB = kRankWindow;                      // columns per bitmap window
    static constexpr int WORDS = WB / 32;                       // bitmap words (32 KB)
    static constexpr int VCAP = BS == kLongBS ? 12288 : 4608;   // values per accumulate pass (96 / 36 KB)
    static constexpr size_t lds = sizeof(Acc) * VCAP + sizeof(uint2) * WORDS;
};
constexpr int kLightBS = 512;

// B range of an A-nonzero recorded by an earlier walk of the same window:
// [start[e] (or the row start on the first window), end[e]).
struct StoredRange {
    const int *__restrict__ brpt;
    const int *__restrict__ start;
    const int *__restrict__ end;
    bool first;
    __device__ __forceinline__ void operator()(int e, int k, int &bs, int &len) const
    {
        const int lo = first ? brpt[k] : start[e];
        bs = lo;
        len = end[e] - lo;
    }
};

// The whole B row (a single column window).
struct RowRange {
    const int *__restrict__ brpt;
    __device__ __forceinline__ void operator()(int /*e*/, int k, int &bs, int &len) const
    {
        bs = brpt[k];
        len = brpt[k + 1] - bs;
    }
};

template <typename V, bool OFF32, int BS>
__global__ __launch_bounds__(BS, 4) void k_num_rank(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                      const V *__restrict__ aval, const int *__restrict__ brpt,
                                                      const int *__restrict__ bcol, const V *__restrict__ bval,
                                                      const RowList rl, int ncols, const long long *__restrict__ crpt,
                                                      int *__restrict__ ccol, V *__restrict__ cval, LongScratch ls,
                                                      long long light_max)
{
    using R = RankWin<V, BS>;
    constexpr int NU = BS / kWave;
    constexpr int WPT = R::WORDS / BS;  // bitmap words per thread
    constexpr int WPW = R::WORDS / NU;       // bitmap words per wave in the rank scan
    constexpr int RPW = WPW / kWave;
    static_assert(R::WORDS % BS == 0, "words per thread");
    extern __shared__ unsigned long long smem_u64[];
    Acc *vals = reinterpret_cast<Acc *>(smem_u64);
    uint2 *br = reinterpret_cast<uint2 *>(vals + R::VCAP);
    __shared__ int s_scan[BS / kWave + 1];
    __shared__ int s_min[BS / kWave];
    __shared__ int s_cut;
    __shared__ int s_tile;
    __shared__ BlockStage<BS, V> s_bst;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int lane = lane_id();
    int *const cursor = ls.cursor;
    const long long span = ls.span;
    // Cursor arrays picked by parity with selects (an indexed array of
    // pointers would live in scratch memory).
    auto cs = [&](int i) { return cursor + (i & 1) * span; };
    auto cz = [&](int i) { return cursor + (2 + (i & 1)) * span; };
    const int cstride = ls.nwin + 1;

    const int *rows;
    const int nrows = row_list(rl, rows);
    for (int i = 0; i < WPT; i++) br[tid + i * BS] = make_uint2(0u, 0u);
    __syncthreads();

    TileSched ts;
    for (int ridx = block_first_row(rl, nrows, ts, &s_tile); ridx < nrows;
         ridx = block_next(rl, ridx, ts, &s_tile)) {
        const int r = rows ? rows[ridx] : ridx;
        const int a_beg = arpt[r], a_end = arpt[r + 1];
        // A row whose column bitmap the symbolic phase kept skips step 1:
        // each window's words are read from it instead.
        const int slot = ls.bm ? ls.bm_slot[r] : -1;
        const unsigned *const sbm = slot >= 0 ? ls.bm + static_cast<long long>(slot) * ls.bm_words : nullptr;
        long long out = crpt[r];
        const long long row_end = crpt[r + 1];
        // With light_max >= 0 the bin is split between two launches: the
        // 512-thread shape takes the rows of at most light_max entries, the
        // 1024-thread shape the others.
        if (light_max >= 0 && ((row_end - out <= light_max) != (BS < kLongBS))) continue;
        // First window: the one holding the smallest column of the row.  A
        // kept row starts at window 0 instead: its windows are visited in
        // turn, an empty one costing only its rank scan, which is cheaper
        // than the three dependent loads per A-nonzero and two barriers of
        // the search (long rows of power-law inputs nearly all have columns
        // in window 0).
        int w0 = 0;
        if (!sbm) {
            int mn = INT_MAX;
            for (int e = a_beg + tid; e < a_end; e += BS) {
                const int k = acol[e];
                if (brpt[k] < brpt[k + 1]) mn = min(mn, bcol[brpt[k]]);
            }
            mn = unit_min<kWave>(mn);
            if (lane_id() == 0) s_min[wave] = mn;
            __syncthreads();
            mn = INT_MAX;
            for (int w = 0; w < NU; w++) mn = min(mn, s_min[w]);
            __syncthreads();
            w0 = mn == INT_MAX ? ncols : (mn / R::WB) * R::WB;
        }
        bool first = true;
        int p = 0;
        // Kept bitmap words of the next window, loaded during this window's
        // accumulate pass (the rank scan of a kept row otherwise waits for
        // its load).
        unsigned pf[RPW];
        bool pf_ok = false;
        while (w0 < ncols) {
            const int wend = (ncols - w0 <= R::WB) ? ncols : w0 + R::WB;
            const bool last = wend >= ncols;
            // 1. mark.  A kept row has no mark walk, so the previous window's
            // output stores are left to complete during the rank scan (the
            // drain before the accumulate walk then rarely waits).
            if (!sbm) drain_vmem();
            int nxt = INT_MAX;
            int sn = 0, stot = 0;
            bool one = false;
            const int wi = w0 / R::WB;
            if (sbm) {
                nxt = last ? INT_MAX : wend;
            } else {
                RankMark<V> op{bcol, br, w0};
                if (ls.wcut) {
                    const WinCutRange range{ls.wcut, bcol, cstride, wi, &nxt};
                    one = block_for_each_product<BS, false, V>(a_beg, a_end, acol, aval, range, op, s_bst, sn,
                                                                    stot);
                } else {
                    const CursorRange range{brpt, bcol, cs(p), cs(p ^ 1), wend, first, last, &nxt};
                    one = block_for_each_product<BS, false, V>(a_beg, a_end, acol, aval, range, op, s_bst, sn,
                                                                    stot);
                }
            }
            if (!sbm) {
                nxt = unit_min<kWave>(nxt);
                if (lane_id() == 0) s_min[wave] = nxt;
                __syncthreads();
            }
            // 2. rank: wave w owns words [w * WPW, (w + 1) * WPW) and its
            // lanes take them 64 consecutive words at a time (conflict-free
            // LDS reads); ranks come from DPP scans of the words' bit counts.
            int cnt[RPW], inc[RPW], rsum[RPW];
            int wtot = 0;
            const bool use_pf = pf_ok;
            pf_ok = false;
#pragma unroll
            for (int i = 0; i < RPW; i++) {
                const int wd = wave * WPW + i * kWave + lane;
                unsigned x;
                if (sbm) {
                    const int g = wi * R::WORDS + wd;
                    x = use_pf ? pf[i] : (g < ls.bm_words ? sbm[g] : 0u);
                    br[wd].x = x;
                } else {
                    x = br[wd].x;
                }
                cnt[i] = __popc(x);
                inc[i] = wave_incl_scan(cnt[i]);
                rsum[i] = __builtin_amdgcn_readlane(inc[i], kWave - 1);
                wtot += rsum[i];
            }
            if (lane == 0) s_scan[wave] = wtot;
            __syncthreads();
            int tot = 0, run = 0;
            for (int w = 0; w < NU; w++) {
                const int t = s_scan[w];
                run += w < wave ? t : 0;
                tot += t;
            }
#pragma unroll
            for (int i = 0; i < RPW; i++) {
                br[wave * WPW + i * kWave + lane].y = static_cast<unsigned>(run + inc[i] - cnt[i]);
                run += rsum[i];
            }
            if (!sbm) {
                nxt = INT_MAX;
                for (int w = 0; w < NU; w++) nxt = min(nxt, s_min[w]);
            }
            __syncthreads();
            // 3 + 4, per sub-window [ca, cb) of ranks [r0, r1)
            int r0 = 0, ca = w0, sub = 0;
            while (r0 < tot) {
                const int r1 = min(tot, r0 + R::VCAP);
                int cb = wend;
                if (r1 < tot) {
                    // The lane whose word holds rank r1 finds its column.
#pragma unroll
                    for (int i = 0; i < RPW; i++) {
                        const int wd = wave * WPW + i * kWave + lane;
                        const uint2 ew = br[wd];
                        const int base = static_cast<int>(ew.y);
                        if (r1 >= base && r1 < base + __popc(ew.x)) {
                            unsigned b = ew.x;
                            for (int skip = r1 - base; skip > 0; skip--) b &= b - 1u;
                            s_cut = w0 + wd * 32 + (__ffs(b) - 1);
                        }
                    }
                    __syncthreads();
                    cb = s_cut;
                }
                for (int i = tid; i < r1 - r0; i += BS) vals[i] = Acc(0);
                __syncthreads();
                drain_vmem();
                if (sbm && sub == 0 && !last) {
                    // Kept rows visit every window in turn: the next is wi + 1.
#pragma unroll
                    for (int i = 0; i < RPW; i++) {
                        const int g = (wi + 1) * R::WORDS + wave * WPW + i * kWave + lane;
                        pf[i] = g < ls.bm_words ? sbm[g] : 0u;
                    }
                    pf_ok = true;
                }
                const RankAcc<V> op{bcol, bval, br, vals, w0, r0, static_cast<unsigned>(r1 - r0)};
                if (tot <= R::VCAP || (tot <= 2 * R::VCAP && !ls.ucut)) {
                    // One sub-window, or two without the unit cut table: walk
                    // the window's products over the ranges step 1 recorded
                    // and skip the other sub-window's ranks, instead of
                    // locating a cut in every B row by galloping.  (With the
                    // unit cuts that search is short, and two sub-windows are
                    // cheaper staged each on its own than walked twice.)
                    // A row that fit in one batch is still staged from step 1.
                    if (one) {
                        if (stot > 0) rank_walk<BS, OFF32, V>(s_bst, sn, stot, op);
                    } else if (ls.wcut) {
                        const WinCutRange range{ls.wcut, bcol, cstride, wi, nullptr};
                        rank_for_each_product<BS, OFF32, V>(a_beg, a_end, acol, aval, range, op, s_bst, sn,
                                                                 stot);
                    } else if (sbm) {
                        const RowRange range{brpt};
                        rank_for_each_product<BS, OFF32, V>(a_beg, a_end, acol, aval, range, op, s_bst, sn,
                                                                 stot);
                    } else {
                        const StoredRange range{brpt, cs(p), cs(p ^ 1), first};
                        rank_for_each_product<BS, OFF32, V>(a_beg, a_end, acol, aval, range, op, s_bst, sn,
                                                                 stot);
                    }
                } else {
                    int n2, t2;
                    if (ls.wcut) {
                        const CutCursorRange range{ls.wcut,  bcol,     cstride,  wi,       cz(sub - 1),
                                                   cz(sub),  cb,       sub == 0, cb == wend, ls.ucut,
                                                   ls.nunit + 1};
                        rank_for_each_product<BS, OFF32, V>(a_beg, a_end, acol, aval, range, op, s_bst, n2, t2);
                    } else {
                        const int *cin = sub == 0 ? cs(p) : cz(sub - 1);
                        const CursorRange range{brpt,       bcol, cin, cz(sub), cb, first && sub == 0,
                                                last && cb == wend, nullptr};
                        rank_for_each_product<BS, OFF32, V>(a_beg, a_end, acol, aval, range, op, s_bst, n2, t2);
                    }
                }
                __syncthreads();
                // 4. write: the values are already in rank order, so they
                // go out as one contiguous copy; the columns are first
                // placed by rank in LDS (over the values just written) and
                // then copied the same way, so every global store is
                // coalesced.
                const int nsub = r1 - r0;
                const long long lim = min(static_cast<long long>(nsub), row_end - (out + r0));
                for (int i = tid; i < lim; i += BS) cval[out + r0 + i] = static_cast<V>(vals[i]);
                __syncthreads();
                int *scol = reinterpret_cast<int *>(vals);
                const int wa = (ca - w0) >> 5;
                const int we = (cb - w0 + 31) >> 5;
                for (int w = wa + tid; w < we; w += BS) {
                    const uint2 ew = br[w];
                    const unsigned all = ew.x;
                    unsigned bits = all;
                    const int lo = ca - w0 - w * 32;
                    const int hi = cb - w0 - w * 32;
                    if (lo > 0) bits &= ~((1u << lo) - 1u);
                    if (hi < 32) bits &= (1u << hi) - 1u;
                    while (bits) {
                        const int bit = __ffs(bits) - 1;
                        bits &= bits - 1u;
                        scol[static_cast<int>(ew.y) + __popc(all & ((1u << bit) - 1u)) - r0] = w0 + w * 32 + bit;
                    }
                }
                __syncthreads();
                for (int i = tid; i < lim; i += BS) ccol[out + r0 + i] = scol[i];
                __syncthreads();
                r0 = r1;
                ca = cb;
                sub++;
            }
            // A kept row's rank scan overwrites every word, so its bitmap is
            // cleared once, after its last window.
            if (!sbm) {
                for (int i = 0; i < WPT; i++) br[tid + i * BS] = make_uint2(0u, 0u);
                __syncthreads();
            }
            out += tot;
            first = false;
            p ^= 1;
            w0 = nxt == INT_MAX ? ncols : (nxt / R::WB) * R::WB;
        }
        if (sbm) {
            for (int i = 0; i < WPT; i++) br[tid + i * BS] = make_uint2(0u, 0u);
            __syncthreads();
        }
    }
}

template <int G, int T, int NMAX, int BS, typename V>
void launch_small(const CsrIn &A, const CsrIn &B, const RowList &rl, const long long *crpt, int *ccol, V *cval,
                  int hash, hipStream_t s)
{
    constexpr int RPB = BS / G;
    const int grid = tile_grid(rl, RPB, kResidentThreads / BS);
    // keys with their slot packed in the low log2(T) bits stay positive
    const bool pack = static_cast<long long>(B.N) <= (1ll << (31 - __builtin_ctz(T)));
    if (hash == kHashLin)
        k_num_small<G, T, NMAX, BS, kHashLin, V><<<grid, BS, 0, s>>>(A.rpt, A.col, static_cast<const V *>(A.val),
                                                                     B.rpt, B.col, static_cast<const V *>(B.val), rl,
                                                                     crpt, ccol, cval, pack);
    else
        k_num_small<G, T, NMAX, BS, kHashFib, V><<<grid, BS, 0, s>>>(A.rpt, A.col, static_cast<const V *>(A.val),
                                                                     B.rpt, B.col, static_cast<const V *>(B.val), rl,
                                                                     crpt, ccol, cval, pack);
}

template <int T, int BS, typename V>
constexpr size_t block_lds()
{
    return sizeof(Acc) * T + sizeof(int) * T + sizeof(int) * (T / 2);
}

template <int T, int BS, typename V>
void launch_block(const CsrIn &A, const CsrIn &B, const RowList &rl, const long long *crpt, int *ccol, V *cval,
                  hipStream_t s, int max_blocks)
{
    k_num_block<T, BS, V><<<row_grid(rl, max_blocks), BS, block_lds<T, BS, V>(), s>>>(
        A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col, static_cast<const V *>(B.val), rl, crpt, ccol, cval);
}

template <typename V>
void set_attrs()
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_num_block<8192, 1024, V>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(block_lds<8192, 1024, V>()));
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_num_block<2048, 256, V>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(block_lds<2048, 256, V>()));
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_num_rank<V, true, kLongBS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(RankWin<V>::lds));
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_num_rank<V, false, kLongBS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(RankWin<V>::lds));
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_num_rank<V, true, kLightBS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(RankWin<V, kLightBS>::lds));
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_num_rank<V, false, kLightBS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(RankWin<V, kLightBS>::lds));
}

// Whether the long-row bin splits its rows between its two block shapes
// (two launches).  OPSAMD_LIGHT_ROWS=0 keeps every row in the 1024-thread
// shape, for tests; a device-sized list whose work counter has no second
// counter beside it keeps one launch too.
bool split_rows(const RowList &rl)
{
    const char *env = std::getenv("OPSAMD_LIGHT_ROWS");
    return !(env && env[0] == '0') && !(rl.next && !rl.next2);
}

// The row list of the long-row bin's second launch: the same rows under the
// second work counter.
RowList second_launch(const RowList &rl)
{
    RowList r = rl;
    r.next = rl.next2;
    return r;
}

}  // namespace

void numeric_init_attributes()
{
    set_attrs<double>();
    set_attrs<float>();
}

template <typename V>
void launch_numeric_bin(int bin, const CsrIn &A, const CsrIn &B, const RowList &rl, const long long *crpt, int *ccol,
                        V *cval, const LongScratch &ls, int *spill, int hash, hipStream_t s, int shape)
{
    if (rl.n <= 0) return;
    // Bins 0-4 (G lanes per row, T slots, NMAX entries, BS threads), shape
    // 0 the default, 1 twice the table, 2 other lanes per row or block size.
    // Bin 3's default takes two rows per wave (32 lanes each), 4% faster
    // than one row per wave on the 27-point stencil (the kernel is
    // VALU-bound there, and the two rows share its per-step instructions);
    // bin 4's, one row (one wave) per block, 2-3% faster than four rows per
    // block.
    switch (bin < 5 ? bin * kShapes + shape : -1) {
        case 0: launch_small<16, 32, 16, 256, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 1: launch_small<16, 64, 16, 256, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 2: launch_small<16, 32, 16, 64, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 3: launch_small<32, 64, 32, 256, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 4: launch_small<32, 128, 32, 256, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 5: launch_small<16, 64, 32, 256, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 6: launch_small<64, 128, 64, 256, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 7: launch_small<64, 256, 64, 256, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 8: launch_small<32, 128, 64, 256, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 9: launch_small<32, 256, 128, 64, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 10: launch_small<64, 256, 128, 64, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 11: launch_small<32, 512, 128, 64, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 12: launch_small<64, 512, 256, 64, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 13: launch_small<64, 1024, 256, 64, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        case 14: launch_small<64, 512, 256, 256, V>(A, B, rl, crpt, ccol, cval, hash, s); return;
        default: break;
    }
    switch (bin) {
        case kSpillBin:
            if (spill) {
                // Wave per row; rows with a wide column span go to the block kernel.
                constexpr int WPB = 4;
                k_zero_count<<<1, kWave, 0, s>>>(spill);
                k_num_wrank<1024, 512, WPB, V><<<tile_grid(rl, 2 * WPB, kResidentThreads / (WPB * kWave)),
                                                 WPB * kWave, 0, s>>>(A.rpt, A.col, static_cast<const V *>(A.val),
                                                                      B.rpt, B.col, static_cast<const V *>(B.val), rl,
                                                                      crpt, ccol, cval, spill);
                // the spilled rows: spill[2 .. 2 + spill[0]), counted on the
                // device; the block kernel takes them from counter spill[1]
                const RowList sl{spill + 2, rl.n, spill, 0, spill + 1};
                launch_block<2048, 256, V>(A, B, sl, crpt, ccol, cval, s, 4096);
            } else {
                launch_block<2048, 256, V>(A, B, rl, crpt, ccol, cval, s, 5 * kNumCu);
            }
            break;
        case 6: launch_block<8192, 1024, V>(A, B, rl, crpt, ccol, cval, s, kNumCu); break;
        default:
        {
            // rows of at most kLightRowMax entries in 512-thread blocks, two
            // per CU, then the others in 1024-thread blocks: over the two
            // parts of a host-sized list ordered by the numeric scatter,
            // else each launch over every row, passing over the other's
            const long long lm = split_rows(rl) ? kLightRowMax : -1;
            RowList rlo = rl, rh = lm >= 0 ? second_launch(rl) : rl;
            if (lm >= 0 && ls.nheavy >= 0 && rl.rows && !rl.dsize && !rl.next && ls.nheavy <= rl.n) {
                // the list is ordered: rows of the 512-thread shape, then the others
                rlo.n = rl.n - ls.nheavy;
                rh.rows = rl.rows + rlo.n;
                rh.n = ls.nheavy;
            }
            const int grid = row_grid(rlo, kNumCu), hgrid = row_grid(rh, kNumCu);
            if (ls.off32) {
                if (lm >= 0 && rlo.n > 0)
                    k_num_rank<V, true, kLightBS><<<grid, kLightBS, RankWin<V, kLightBS>::lds, s>>>(
                        A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col, static_cast<const V *>(B.val), rlo,
                        B.N, crpt, ccol, cval, ls, lm);
                if (rh.n > 0)
                    k_num_rank<V, true, kLongBS><<<hgrid, kLongBS, RankWin<V>::lds, s>>>(
                        A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col, static_cast<const V *>(B.val), rh,
                        B.N, crpt, ccol, cval, ls, lm);
            } else {
                if (lm >= 0 && rlo.n > 0)
                    k_num_rank<V, false, kLightBS><<<grid, kLightBS, RankWin<V, kLightBS>::lds, s>>>(
                        A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col, static_cast<const V *>(B.val), rlo,
                        B.N, crpt, ccol, cval, ls, lm);
                if (rh.n > 0)
                    k_num_rank<V, false, kLongBS><<<hgrid, kLongBS, RankWin<V>::lds, s>>>(
                        A.rpt, A.col, static_cast<const V *>(A.val), B.rpt, B.col, static_cast<const V *>(B.val), rh,
                        B.N, crpt, ccol, cval, ls, lm);
            }
            break;
        }
    }
}

template void launch_numeric_bin<double>(int, const CsrIn &, const CsrIn &, const RowList &, const long long *, int *,
                                         double *, const LongScratch &, int *, int, hipStream_t, int);
template void launch_numeric_bin<float>(int, const CsrIn &, const CsrIn &, const RowList &, const long long *, int *,
                                        float *, const LongScratch &, int *, int, hipStream_t, int);

namespace {

__global__ void k_window_cuts(const int *__restrict__ brpt, const int *__restrict__ bcol, int nb, int nwin,
                              int shift, int *__restrict__ wcut)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int stride = nwin + 1;
    if (i >= static_cast<long long>(nb) * stride) return;
    const int k = static_cast<int>(i / stride);
    const int w = static_cast<int>(i % stride);
    const int lo = brpt[k], hi = brpt[k + 1];
    int v;
    if (w == 0) {
        v = lo;
    } else if (w == nwin) {
        v = hi;
    } else {
        const int bound = w << shift;
        int l = lo, h = hi;
        while (l < h) {
            const int m = (l + h) >> 1;
            if (bcol[m] < bound)
                l = m + 1;
            else
                h = m;
        }
        v = l;
    }
    wcut[i] = v;
}

}  // namespace

void launch_window_cuts(const CsrIn &B, int nwin, int *wcut, hipStream_t s, int shift)
{
    const long long n = static_cast<long long>(B.M) * (nwin + 1);
    if (n <= 0) return;
    constexpr int BS = 256;
    k_window_cuts<<<static_cast<unsigned>((n + BS - 1) / BS), BS, 0, s>>>(B.rpt, B.col, B.M, nwin, shift, wcut);
}

}  // namespace opsamd
