// Setup (products per row + symbolic bin histogram), bin scatter, and the
// row-pointer scan with numeric bin statistics.
//
// Reference counterparts: k_compute_flop (src/opsparse/numeric/compute_flop.cu:9-31),
// k_symbolic_binning / k_symbolic_binning2 (kernels/binning.cuh:11-81),
// k_numeric_binning / k_numeric_binning2 (binning.cuh:90-174) and the CUB
// exclusive scan of C.d_rpt (spgemm.cu:49, 113).
#include <algorithm>

#include "device.hpp"
#include "launch.hpp"

namespace opsamd {

namespace {

__device__ __forceinline__ int dev_sym_bin(int f, const BinBounds &bb)
{
    int b = 0;
#pragma unroll
    for (int i = 0; i < kNumBins - 1; i++) b += f > bb.sym[i] ? 1 : 0;
    return b;
}

__device__ __forceinline__ int dev_num_bin(int n, const BinBounds &bb)
{
    int b = 0;
#pragma unroll
    for (int i = 0; i < kNumBins - 1; i++) b += n > bb.num[i] ? 1 : 0;
    return b;
}

__global__ void k_zero_stats(DevStats *st)
{
    unsigned *p = reinterpret_cast<unsigned *>(st);
    for (unsigned i = threadIdx.x; i < sizeof(DevStats) / 4; i += blockDim.x) p[i] = 0;
}

// info[0] = nnz(C), info[1] = 1 when it exceeds `capacity` (0 otherwise);
// on overflow every numeric bin is emptied, so device-sized numeric kernels
// write nothing.
__global__ void k_capacity(DevStats *st, long long capacity, long long *info)
{
    if (threadIdx.x != 0) return;
    const long long nnz = static_cast<long long>(st->total_nnz);
    const bool over = nnz > capacity;
    info[0] = nnz;
    info[1] = over ? 1 : 0;
    if (over)
        for (int b = 0; b < kNumBins; b++) st->num_size[b] = 0;
}

// Products per row, grid-stride over tiles of 256 rows.  The block walks
// the tile's A-nonzeros flat (one contiguous range of A.col): in each round
// thread t takes 4 consecutive entries, so every lane has four independent
// A.col -> B.rpt load chains in flight whatever the row lengths (a row of
// 40k A-nonzeros is spread over the whole block), and adds the entries'
// B-row lengths to its row's LDS sum; the row of an entry comes from a
// binary search over the tile's row offsets in LDS.  Per-bin counts
// accumulate in LDS and reach global memory once per block.
constexpr int kSetupRows = 256;
constexpr int kSetupMaxGrid = 2048;
constexpr int kSetupUnroll = 4;

__global__ __launch_bounds__(kSetupRows) void k_setup(const int *__restrict__ arpt, const int *__restrict__ acol,
                                                      const int *__restrict__ brpt, int M, int *__restrict__ row_flop,
                                                      int *__restrict__ row_nnz, DevStats *__restrict__ st,
                                                      const BinBounds bb)
{
    __shared__ int s_hist[kNumBins];
    __shared__ unsigned long long s_bflop[kNumBins];
    __shared__ unsigned long long s_bannz[kNumBins];
    __shared__ unsigned long long s_red[kSetupRows / kWave];
    __shared__ unsigned long long s_f[kSetupRows];
    __shared__ int s_rp[kSetupRows + 1];
    __shared__ int s_max;

    const int tid = threadIdx.x;
    const bool vec = (reinterpret_cast<uintptr_t>(acol) & 15u) == 0;
    if (tid < kNumBins) {
        s_hist[tid] = 0;
        s_bflop[tid] = 0;
        s_bannz[tid] = 0;
    }
    if (tid == 0) s_max = 0;
    __syncthreads();

    unsigned long long tot = 0;
    int mx = 0;
    const int ntiles = (M + kSetupRows - 1) / kSetupRows;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int r0 = tile * kSetupRows;
        const int nr = min(kSetupRows, M - r0);
        s_rp[tid] = arpt[r0 + min(tid, nr)];
        if (tid == 0) s_rp[kSetupRows] = arpt[r0 + nr];
        s_f[tid] = 0;
        __syncthreads();
        const int e0 = s_rp[0], e1 = s_rp[nr];
        // Lane t takes 4 consecutive entries per round (one 16-byte load of
        // A.col when aligned) and keeps a running sum for its current row,
        // adding it to the row's LDS sum only when its row changes: a long
        // row costs one LDS atomic per lane, not one per entry.
        int cur_row = -1;
        unsigned long long cur = 0;
        for (int base = e0 & ~(kSetupUnroll - 1); base < e1; base += kSetupUnroll * kSetupRows) {
            const int e = base + kSetupUnroll * tid;
            int k[kSetupUnroll];
            if (vec && e >= e0 && e + kSetupUnroll <= e1) {
                const int4 v = *reinterpret_cast<const int4 *>(acol + e);
                k[0] = v.x;
                k[1] = v.y;
                k[2] = v.z;
                k[3] = v.w;
            } else {
#pragma unroll
                for (int j = 0; j < kSetupUnroll; j++) k[j] = e + j >= e0 && e + j < e1 ? acol[e + j] : -1;
            }
            int f[kSetupUnroll];
#pragma unroll
            for (int j = 0; j < kSetupUnroll; j++) f[j] = k[j] >= 0 ? brpt[k[j] + 1] - brpt[k[j]] : 0;
#pragma unroll
            for (int j = 0; j < kSetupUnroll; j++) {
                if (k[j] < 0) continue;
                const int ej = e + j;
                if (cur_row < 0 || ej >= s_rp[cur_row + 1]) {
                    if (cur) atomicAdd(&s_f[cur_row], cur);
                    cur = 0;
                    // Largest row rl < nr with s_rp[rl] <= ej.
                    int lo = cur_row < 0 ? 0 : cur_row, hi = nr - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_rp[mid] <= ej)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    cur_row = lo;
                }
                cur += static_cast<unsigned long long>(f[j]);
            }
        }
        if (cur) atomicAdd(&s_f[cur_row], cur);
        __syncthreads();
        const int r = r0 + tid;
        if (r < M) {
            const unsigned long long f = s_f[tid];
            const int fi = f > static_cast<unsigned long long>(INT_MAX) ? INT_MAX : static_cast<int>(f);
            row_flop[r] = fi;
            if (fi == 0) {
                row_nnz[r] = 0;
            } else {
                const int b = dev_sym_bin(fi, bb);
                atomicAdd(&s_hist[b], 1);
                atomicAdd(&s_bflop[b], f);
                atomicAdd(&s_bannz[b], static_cast<unsigned long long>(s_rp[tid + 1] - s_rp[tid]));
            }
            mx = max(mx, fi);
            tot += f;
        }
        __syncthreads();
    }
    atomicMax(&s_max, mx);
    tot = block_sum<kSetupRows>(tot, s_red);
    if (tid < kNumBins && s_hist[tid]) {
        atomicAdd(&st->sym_size[tid], s_hist[tid]);
        atomicAdd(&st->sym_flop[tid], s_bflop[tid]);
        atomicAdd(&st->sym_annz[tid], s_bannz[tid]);
    }
    if (tid == 0) {
        atomicAdd(&st->total_flop, tot);
        atomicMax(&st->max_flop, s_max);
    }
}

// Scatter row ids into bin lists.  Bin b occupies [sum(size[0..b)), ...).
// Tiles of 4096 rows; ranks inside a tile come from wave ballots, and each
// tile reserves its slice of every bin with one global atomic.  Order inside
// a bin depends on tile scheduling, as in the reference (binning.cuh:75);
// only the set of rows per bin is fixed.
constexpr int kScatBS = 1024;
constexpr int kScatItems = 4;

__global__ __launch_bounds__(kScatBS) void k_bin_scatter(const int *__restrict__ key, int M, int which,
                                                         DevStats *__restrict__ st, int *__restrict__ bins,
                                                         const BinBounds bb)
{
    // classes: the bins, then (numeric) the long-row bin's rows of more
    // than kLightRowMax entries, placed at the end of that bin's list
    constexpr int kClasses = kNumBins + 1;
    __shared__ int s_cnt[kClasses];
    __shared__ int s_base[kClasses];
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const unsigned long long lt = lane ? ((1ull << lane) - 1ull) : 0ull;
    if (tid < kClasses) s_cnt[tid] = 0;
    __syncthreads();
    const int r0 = blockIdx.x * (kScatBS * kScatItems);
    int bin[kScatItems], rank[kScatItems];
#pragma unroll
    for (int i = 0; i < kScatItems; i++) {
        const int r = r0 + i * kScatBS + tid;
        bin[i] = -1;
        rank[i] = 0;
        if (r < M) {
            const int v = key[r];
            if (v > 0) bin[i] = which ? (v > kLightRowMax ? kNumBins : dev_num_bin(v, bb)) : dev_sym_bin(v, bb);
        }
#pragma unroll
        for (int b = 0; b < kClasses; b++) {
            const unsigned long long m = __ballot(bin[i] == b);
            if (!m) continue;
            const int leader = __ffsll(static_cast<long long>(m)) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(&s_cnt[b], __popcll(m));
            base = __builtin_amdgcn_readlane(base, leader);
            if (bin[i] == b) rank[i] = base + __popcll(m & lt);
        }
    }
    __syncthreads();
    if (tid < kNumBins) {
        const int *size = which ? st->num_size : st->sym_size;
        int *fill = which ? st->num_fill : st->sym_fill;
        int off = 0;
        for (int i = 0; i < tid; i++) off += size[i];
        s_base[tid] = s_cnt[tid] ? off + atomicAdd(&fill[tid], s_cnt[tid]) : 0;
    }
    if (tid == kNumBins && s_cnt[tid]) {
        int off = 0;
        for (int i = 0; i < kNumBins; i++) off += st->num_size[i];
        s_base[tid] = off - st->num_heavy + atomicAdd(&st->heavy_fill, s_cnt[tid]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScatItems; i++)
        if (bin[i] >= 0) bins[s_base[bin[i]] + rank[i]] = r0 + i * kScatBS + tid;
}

// --- scan: 1024 threads x 4 rows per block ---
constexpr int kScanBS = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBS * kScanItems;

__global__ __launch_bounds__(kScanBS) void k_scan_reduce(const int *__restrict__ row_nnz,
                                                         const int *__restrict__ row_flop,
                                                         const int *__restrict__ arpt, int M,
                                                         long long *__restrict__ block_sums, DevStats *__restrict__ st,
                                                         const BinBounds bb)
{
    __shared__ unsigned long long s_red[kScanBS / kWave];
    __shared__ int s_hist[kNumBins];
    __shared__ unsigned long long s_f[kNumBins], s_n[kNumBins], s_a[kNumBins];
    __shared__ int s_max;
    const int tid = threadIdx.x;
    if (tid < kNumBins) {
        s_hist[tid] = 0;
        s_f[tid] = s_n[tid] = s_a[tid] = 0;
    }
    if (tid == 0) s_max = 0;
    __syncthreads();
    const int r0 = blockIdx.x * kScanTile + tid * kScanItems;
    unsigned long long sum = 0;
    int mx = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const int r = r0 + i;
        if (r < M) {
            const int n = row_nnz[r];
            sum += static_cast<unsigned long long>(n);
            mx = max(mx, n);
            if (n > kLightRowMax) atomicAdd(&st->num_heavy, 1);
            if (n > 0) {
                const int b = dev_num_bin(n, bb);
                atomicAdd(&s_hist[b], 1);
                atomicAdd(&s_f[b], static_cast<unsigned long long>(row_flop[r]));
                atomicAdd(&s_n[b], static_cast<unsigned long long>(n));
                atomicAdd(&s_a[b], static_cast<unsigned long long>(arpt[r + 1] - arpt[r]));
            }
        }
    }
    atomicMax(&s_max, mx);
    const unsigned long long tot = block_sum<kScanBS>(sum, s_red);
    if (tid == 0) {
        block_sums[blockIdx.x] = static_cast<long long>(tot);
        atomicMax(&st->max_nnz, s_max);
    }
    if (tid < kNumBins && s_hist[tid]) {
        atomicAdd(&st->num_size[tid], s_hist[tid]);
        atomicAdd(&st->num_flop[tid], s_f[tid]);
        atomicAdd(&st->num_nnz[tid], s_n[tid]);
        atomicAdd(&st->num_annz[tid], s_a[tid]);
    }
}

// Exclusive scan of the block sums, one block; total -> st->total_nnz and
// block_sums[nb].
__global__ __launch_bounds__(kScanBS) void k_scan_blocks(long long *__restrict__ block_sums, int nb,
                                                         DevStats *__restrict__ st)
{
    __shared__ long long s_w[kScanBS / kWave + 1];
    __shared__ long long s_carry;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wid = tid / kWave;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += kScanBS) {
        const int i = base + tid;
        const long long v = i < nb ? block_sums[i] : 0;
        long long incl = v;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const long long t = __shfl_up(incl, d, kWave);
            if (lane >= d) incl += t;
        }
        if (lane == kWave - 1) s_w[wid] = incl;
        __syncthreads();
        if (tid == 0) {
            long long run = 0;
            for (int w = 0; w < kScanBS / kWave; w++) {
                const long long t = s_w[w];
                s_w[w] = run;
                run += t;
            }
            s_w[kScanBS / kWave] = run;
        }
        __syncthreads();
        const long long carry = s_carry;
        if (i < nb) block_sums[i] = carry + s_w[wid] + incl - v;
        __syncthreads();
        if (tid == 0) s_carry = carry + s_w[kScanBS / kWave];
        __syncthreads();
    }
    if (tid == 0) {
        block_sums[nb] = s_carry;
        st->total_nnz = static_cast<unsigned long long>(s_carry);
    }
}

__global__ __launch_bounds__(kScanBS) void k_scan_write(const int *__restrict__ row_nnz, int M,
                                                        const long long *__restrict__ block_sums, int nb,
                                                        long long *__restrict__ crpt64, int *__restrict__ crpt32)
{
    __shared__ long long s_w[kScanBS / kWave + 1];
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wid = tid / kWave;
    const int r0 = blockIdx.x * kScanTile + tid * kScanItems;
    int v[kScanItems];
    long long loc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = r0 + i < M ? row_nnz[r0 + i] : 0;
        loc += v[i];
    }
    long long incl = loc;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const long long t = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += t;
    }
    if (lane == kWave - 1) s_w[wid] = incl;
    __syncthreads();
    if (tid == 0) {
        long long run = 0;
        for (int w = 0; w < kScanBS / kWave; w++) {
            const long long t = s_w[w];
            s_w[w] = run;
            run += t;
        }
    }
    __syncthreads();
    const long long total = block_sums[nb];
    const bool fits = total <= static_cast<long long>(INT_MAX);
    long long run = block_sums[blockIdx.x] + s_w[wid] + incl - loc;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const int r = r0 + i;
        if (r < M) {
            crpt64[r] = run;
            if (crpt32 && fits) crpt32[r] = static_cast<int>(run);
        }
        run += v[i];
    }
    if (blockIdx.x == 0 && tid == 0) {
        crpt64[M] = total;
        if (crpt32 && fits) crpt32[M] = static_cast<int>(total);
    }
}

__global__ void k_narrow(const long long *__restrict__ src, int n, int *__restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = static_cast<int>(src[i]);
}

__global__ void k_fill_int(int *__restrict__ p, int v, long long n)
{
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        p[i] = v;
}

__global__ void k_rownnz(const long long *__restrict__ crpt, int M, int *__restrict__ row_nnz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) row_nnz[i] = static_cast<int>(crpt[i + 1] - crpt[i]);
}

}  // namespace

void launch_rownnz_from_rpt(const long long *crpt, int M, int *row_nnz, hipStream_t s)
{
    if (M <= 0) return;
    k_rownnz<<<(M + 255) / 256, 256, 0, s>>>(crpt, M, row_nnz);
}

void launch_fill_int(int *p, int v, long long n, hipStream_t s)
{
    if (n <= 0) return;
    const int grid = static_cast<int>(std::min<long long>((n + 255) / 256, 16384));
    k_fill_int<<<grid, 256, 0, s>>>(p, v, n);
}

void launch_zero_stats(DevStats *st, hipStream_t s) { k_zero_stats<<<1, 256, 0, s>>>(st); }

void launch_capacity_check(DevStats *st, long long capacity, long long *info, hipStream_t s)
{
    k_capacity<<<1, 64, 0, s>>>(st, capacity, info);
}

void launch_setup(const CsrIn &A, const CsrIn &B, int *row_flop, int *row_nnz, DevStats *st, const BinBounds &bb,
                  hipStream_t s)
{
    if (A.M == 0) return;
    const int grid = std::min((A.M + kSetupRows - 1) / kSetupRows, kSetupMaxGrid);
    k_setup<<<grid, kSetupRows, 0, s>>>(A.rpt, A.col, B.rpt, A.M, row_flop, row_nnz, st, bb);
}

void launch_bin_scatter(const int *key, int M, int which, DevStats *st, int *bins, const BinBounds &bb,
                        hipStream_t s)
{
    if (M == 0) return;
    constexpr int tile = kScatBS * kScatItems;
    k_bin_scatter<<<(M + tile - 1) / tile, kScatBS, 0, s>>>(key, M, which, st, bins, bb);
}

int scan_blocks(int M) { return (M + kScanTile - 1) / kScanTile; }

void launch_scan(const int *row_nnz, const int *row_flop, const int *arpt, int M, long long *block_sums,
                 long long *crpt64, int *crpt32, DevStats *st, const BinBounds &bb, hipStream_t s)
{
    const int nb = scan_blocks(M);
    if (nb > 0) k_scan_reduce<<<nb, kScanBS, 0, s>>>(row_nnz, row_flop, arpt, M, block_sums, st, bb);
    k_scan_blocks<<<1, kScanBS, 0, s>>>(block_sums, nb, st);
    k_scan_write<<<nb > 0 ? nb : 1, kScanBS, 0, s>>>(row_nnz, M, block_sums, nb, crpt64, crpt32);
}

void launch_narrow_rpt(const long long *crpt64, int M, int *crpt32, hipStream_t s)
{
    const int n = M + 1;
    k_narrow<<<(n + 255) / 256, 256, 0, s>>>(crpt64, n, crpt32);
}

}  // namespace opsamd
