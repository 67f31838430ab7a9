// Host-side declarations of the SpGEMM kernel launchers and the row-bin
// configuration shared by the kernels and the pipeline driver.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

namespace opsamd {

constexpr int kNumBins = 8;

// Hash families for the LDS tables (see hash_home in device.hpp).
enum HashKind { kHashLin = 0, kHashFib = 1 };

// Symbolic bins by row products f (inclusive upper bounds); rows with f == 0
// are not binned (their nnz is 0).  Bin 7 is the long-row path.  The
// reference bins by the same quantity with bounds {26, 426, ..., 10240, inf}
// sized for 32-lane warps and 48 KB of shared memory (binning.cuh:21, 51);
// these are sized for 64-lane units and 160 KB of LDS.
//   bin:  0    1    2    3     4      5      6       7
//   G:    16   32   64   64    64     256    1024    1024 (lanes per row)
//   T:    32   64   128  512   1024   4096   16384   bitmap, 2^20 columns per window
constexpr int kSymBound[kNumBins] = {16, 32, 64, 256, 1024, 4096, 16384, INT_MAX};

// Numeric bins by exact row nnz n (inclusive upper bounds); rows with n == 0
// are not binned.  The reference uses {16, 128, ..., 4095, inf}
// (binning.cuh:105, 143).  Hash tables hold >= 2n slots.
//   bin:  0    1    2    3     4     5      6      7
//   G:    16   32   64   64    64    256    1024   1024
//   T:    32   64   128  256   512   2048   8192   dense column windows
//   sort: unit bitonic in registers (bins 0-4), LDS bitonic (5-6), none (7)
constexpr int kNumBound[kNumBins] = {16, 32, 64, 128, 256, 1024, 4096, INT_MAX};

// Largest allowed bound per bin: a symbolic bin-b kernel's table holds T_b
// keys, so rows with at most T_b products can never overflow it; a numeric
// bin-b kernel holds at most NMAX_b entries per row.
constexpr int kSymCap[kNumBins - 1] = {32, 64, 128, 512, 1024, 4096, 16384};
constexpr int kNumCap[kNumBins - 1] = {16, 32, 64, 128, 256, 1024, 4096};

// Run-time bin boundaries (inclusive upper bounds of bins 0..6; bin 7 is
// unbounded).  Defaults are kSymBound / kNumBound; the autotuner may lower
// them (moving rows to wider kernels with larger tables) within the caps.
struct BinBounds {
    int sym[kNumBins - 1];
    int num[kNumBins - 1];
};

inline BinBounds default_bounds()
{
    BinBounds b{};
    for (int i = 0; i < kNumBins - 1; i++) {
        b.sym[i] = kSymBound[i];
        b.num[i] = kNumBound[i];
    }
    return b;
}

// Device-side counters of one call (zeroed by k_zero_stats).
struct DevStats {
    unsigned long long total_flop;
    unsigned long long total_nnz;
    int max_flop;
    int max_nnz;
    int sym_size[kNumBins];
    int sym_fill[kNumBins];
    int num_size[kNumBins];
    int num_fill[kNumBins];
    unsigned long long sym_flop[kNumBins];
    unsigned long long sym_annz[kNumBins];
    unsigned long long num_flop[kNumBins];
    unsigned long long num_nnz[kNumBins];
    unsigned long long num_annz[kNumBins];
    int kept_rows;  // long rows that asked for a kept column bitmap (SymSave::count)
    int pad_;
    // work counters of device-sized launches (RowList::next), per bin
    int next_sym[kNumBins];
    int next_num[kNumBins];
    // work counters of the second launch of a numeric bin that splits its
    // rows between two block shapes (RowList::next2)
    int next_num2[kNumBins];
    // numeric long-row bin: rows of more than kLightRowMax entries (counted
    // by the scan) and the numeric scatter's counter for them
    int num_heavy;
    int heavy_fill;
};

struct CsrIn {
    const int *rpt;
    const int *col;
    const void *val;
    int M, N;
};

// Rows one per-bin kernel processes.  Host-sized (dsize == nullptr): the
// list rows[0..n) (rows == nullptr: rows 0..n-1).  Device-sized: the list
// starts at rows + dsize[0] + ... + dsize[bin - 1] and holds dsize[bin]
// entries, read on the device (the bin counts a setup or scan kernel left in
// DevStats, or a counter an earlier kernel appended rows under); n is then
// only an upper bound that sizes the grid, and every kernel loops over its
// tiles, so a launch recorded in a graph serves any bin sizes.  With
// `next` (a counter zeroed before the launch), blocks take their tiles from
// it one at a time instead of by a fixed stride, so a grid narrower than
// the tile count stays balanced whatever the rows cost.  A bin that runs
// its rows in two launches (each taking the rows of one block shape and
// passing over the others) gives the second launch `next2`, a counter of its
// own; a device-sized list with `next` and no `next2` runs in one launch.
struct RowList {
    const int *rows = nullptr;
    int n = 0;
    const int *dsize = nullptr;
    int bin = 0;
    int *next = nullptr;
    int *next2 = nullptr;
};

// Grid of a per-bin kernel over tiles of `per_tile` rows.  Host-sized: one
// block per tile (a multiple of 8, for the XCD-aware tile order).
// Device-sized: at most `max_blocks` blocks, each looping over tiles.
constexpr int kNumCu = 256;
constexpr int kResidentThreads = kNumCu * 2048;  // threads resident on the device at full occupancy
inline int tile_grid(const RowList &rl, int per_tile, int max_blocks)
{
    long long tiles = (static_cast<long long>(rl.n) + per_tile - 1) / per_tile;
    if (rl.dsize && tiles > max_blocks) tiles = max_blocks;
    return static_cast<int>((tiles + 7) / 8 * 8);
}
// Grid of a kernel whose blocks take rows blockIdx.x, blockIdx.x + gridDim.x, ...
inline int row_grid(const RowList &rl, int max_blocks)
{
    return rl.dsize ? (rl.n < max_blocks ? rl.n : max_blocks) : rl.n;
}

// Rows of the numeric long-row bin with at most this many entries run in
// the kernel's 512-thread shape, the others in its 1024-thread shape.
constexpr long long kLightRowMax = 131072;

// --- setup / binning / scan (k_setup_scan.hip) ---
void launch_zero_stats(DevStats *st, hipStream_t s);
// row_flop[M] (int32, saturated), per-bin counts; rows with f == 0 get row_nnz = 0.
void launch_setup(const CsrIn &A, const CsrIn &B, int *row_flop, int *row_nnz, DevStats *st, const BinBounds &bb,
                  hipStream_t s);
// Scatters row ids into bins (which = 0: symbolic by row_flop, 1: numeric by
// row_nnz).  Numeric: the long-row bin's list holds its rows of at most
// kLightRowMax entries first and the DevStats::num_heavy others after them
// (the scan that counted them runs first).
void launch_bin_scatter(const int *key, int M, int which, DevStats *st, int *bins, const BinBounds &bb,
                        hipStream_t s);
// Exclusive scan of row_nnz into crpt64 (and crpt32 when nnz(C) fits int32
// and crpt32 != nullptr); numeric bin statistics.  block_sums needs
// scan_blocks(M) + 1 int64 entries.
int scan_blocks(int M);
void launch_scan(const int *row_nnz, const int *row_flop, const int *arpt, int M, long long *block_sums,
                 long long *crpt64, int *crpt32, DevStats *st, const BinBounds &bb, hipStream_t s);
void launch_narrow_rpt(const long long *crpt64, int M, int *crpt32, hipStream_t s);
// After the scan and the numeric binning of a device-sized call: info[0] =
// nnz(C), info[1] = 1 if nnz(C) > capacity (then every numeric bin size in
// st is set to 0, so no numeric kernel writes), else 0.
void launch_capacity_check(DevStats *st, long long capacity, long long *info, hipStream_t s);
// p[0..n) = v with a kernel (no memset node when captured in a graph).
void launch_fill_int(int *p, int v, long long n, hipStream_t s);

// --- symbolic (k_symbolic.hip) ---
// Columns per bitmap window of the long-row symbolic kernel (128 KB of LDS).
constexpr int kSymLongWindow = 1 << 20;

// Column bitmaps the long-row symbolic kernel keeps for the numeric phase
// (B's columns in one window): a row whose count exceeds min_nnz takes slot
// atomicAdd(count, 1) of bm (words per slot: ceil(B.N / 32)) and, if that
// slot is below cap, records it in slot[row] (left at -1 otherwise);
// bm == nullptr keeps none.
struct SymSave {
    unsigned *bm = nullptr;
    int *slot = nullptr;
    int *count = nullptr;
    int min_nnz = 0;
    int cap = 0;
    // cap covers every row of the bin: a row's slot is its place in the
    // bin's row list (no slot counter to wait for)
    bool by_index = false;
};
// Computes row_nnz for the rows of one bin.
// hash: family for the short-row bins 0-4 (kHashLin or kHashFib in device.hpp).
// bitmap: bins 0-4 count rows with a narrow column span in a bitmap.
// save: bitmaps of bin 7 kept for the numeric phase (only when B.N <= kSymLongWindow).
// shape: kernel shape of bins 0-4 (kShapes alternatives, 0 = default; the
// table is in opsparse_amd.h, ops_tuning).
constexpr int kShapes = 3;
void launch_symbolic_bin(int bin, const CsrIn &A, const CsrIn &B, const RowList &rl, int *row_nnz, int hash,
                         bool bitmap, const SymSave &save, hipStream_t s, int shape = 0);
// Bytes of dynamic LDS the long-row symbolic kernel asks for.
size_t symbolic_long_lds();

// --- numeric (k_numeric.hip) ---
// Numeric bin run one row per wave first (rows too wide for its bitmap are
// passed on to the block kernel through `spill`).
constexpr int kSpillBin = 5;

// Column window of the long-row bitmap kernels (numeric and backward).
constexpr int kRankWindow = 1 << 17;
// Largest window-cut table built (bytes); beyond it the long-row kernels
// locate window boundaries by galloping in each B row instead.
constexpr long long kMaxCutBytes = 1ll << 30;

// Columns between the cuts of the unit cut table (LongScratch::ucut).
constexpr int kUnitCutShift = 14;

// Scratch of the long-row bin.
struct LongScratch {
    int *cursor = nullptr;      // four ints per A-nonzero index
    int span = 0;               // highest A-nonzero index + 1
    const int *wcut = nullptr;  // window cuts of every B row, or nullptr
    int nwin = 0;               // column windows of kRankWindow (wcut has nwin + 1 per B row)
    // Column bitmaps kept by the symbolic phase (SymSave): row r's bitmap is
    // bm + bm_slot[r] * bm_words when bm_slot[r] >= 0.  Only set when every
    // window's B ranges come without a mark walk (wcut, or a single window).
    const unsigned *bm = nullptr;
    const int *bm_slot = nullptr;
    int bm_words = 0;
    // B's entries are addressed by 32-bit byte offsets: every entry index of
    // B (up to its rpt[M]) is below 2^29.
    bool off32 = false;
    // With wcut, optionally the cut table at every 2^kUnitCutShift columns
    // (nunit + 1 per B row), which bounds the search for a value
    // sub-window's end in each B row.
    const int *ucut = nullptr;
    int nunit = 0;
    // With a host-sized row list: its last nheavy rows hold more than
    // kLightRowMax entries and the others at most that (the numeric
    // scatter's order), so each block shape launches over its own rows;
    // -1: unknown (each launch passes over the other's rows).
    int nheavy = -1;
};

// wcut[k * (nwin + 1) + w] = first entry of B row k with column >= w *
// kRankWindow (w = 0: the row start, w = nwin: the row end); with shift,
// at every multiple of 2^shift columns instead.
void launch_window_cuts(const CsrIn &B, int nwin, int *wcut, hipStream_t s, int shift = 17);

// spill: scratch of (rows of bin kSpillBin + 2) ints (a count, a work
// counter, the rows), or nullptr (block kernel only).
template <typename V>
void launch_numeric_bin(int bin, const CsrIn &A, const CsrIn &B, const RowList &rl, const long long *crpt, int *ccol,
                        V *cval, const LongScratch &ls, int *spill, int hash, hipStream_t s, int shape = 0);
// One-time kernel attribute setup (dynamic LDS limits).
void numeric_init_attributes();
void symbolic_init_attributes();

// --- backward (k_backward.hip) ---
// da[e - a0] receives the gradient of A entry e (a0 = A.rpt[0]), db[e - b0]
// that of B entry e (b0 = B.rpt[0]).  ucut: B's unit cut table (nunit + 1
// ints per B row) for the long-row bin's sub-window ends, or nullptr.
template <typename V>
void launch_backward_bin(int bin, const CsrIn &A, const CsrIn &B, const int *rows, int nrows, const long long *crpt,
                         const int *ccol, const V *dc, V *da, V *db, int a0, int b0, int *cursor, int span,
                         const int *ucut, int nunit, hipStream_t s);
void backward_init_attributes();
void launch_zero_values(void *p, long long n, bool f32, hipStream_t s);
// row_nnz[i] = crpt[i+1] - crpt[i] (numeric binning key for the backward).
void launch_rownnz_from_rpt(const long long *crpt, int M, int *row_nnz, hipStream_t s);


}  // namespace opsamd
