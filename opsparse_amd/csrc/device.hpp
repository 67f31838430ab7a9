// Device-side building blocks for the gfx950 SpGEMM kernels.
//
// Design notes (CDNA4, wave64):
//  * A "unit" is an aligned group of U lanes (U = 8..64, power of two) that
//    cooperates on one A-row chunk.  All cross-lane traffic inside a unit goes
//    through ds_bpermute with explicit source lanes, so units of one wave
//    never see each other and no LDS memory or barrier is needed for it.
//  * Products of an A row are flattened: the unit loads U A-nonzeros, scans
//    their B-row lengths, and lane l then takes products l, l+U, ... of the
//    chunk, finding its A-nonzero by a binary search over the scanned
//    offsets.  Every lane is busy whatever the B-row lengths are, which is
//    what a 64-wide wavefront needs (the reference gives each A-nonzero a
//    32-lane warp and strides over its B row, symbolic.cuh:94-96, which
//    idles most lanes on short B rows).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <type_traits>

#include "launch.hpp"

namespace opsamd {

constexpr int kWave = 64;

// Type of every LDS value accumulator, for FP64 and FP32 values alike: an
// LDS FP32 atomic add (ds_add_f32) runs about 2.8x slower than the FP64 one
// on gfx950 (the stencil's FP32 hash kernel: 4.6 ms against 1.7 ms), and
// FP32 results are then rounded once, from exact products summed in FP64.
using Acc = double;


__device__ __forceinline__ int lane_id() { return static_cast<int>(__lane_id()); }

// Value of `v` held by absolute lane `src` of this wave.
__device__ __forceinline__ int bperm(int v, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, v); }
__device__ __forceinline__ unsigned bperm(unsigned v, int src)
{
    return static_cast<unsigned>(__builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(v)));
}
__device__ __forceinline__ float bperm(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ unsigned long long bperm(unsigned long long v, int src)
{
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(v));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(v >> 32));
    return (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) | static_cast<unsigned>(lo);
}
__device__ __forceinline__ long long bperm(long long v, int src)
{
    return static_cast<long long>(bperm(static_cast<unsigned long long>(v), src));
}
__device__ __forceinline__ double bperm(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(b));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Inclusive prefix sum over aligned groups of U lanes.
template <int U>
__device__ __forceinline__ int unit_incl_scan(int v)
{
    const int lane = lane_id();
    const int gl = lane & (U - 1);
#pragma unroll
    for (int d = 1; d < U; d <<= 1) {
        const int t = bperm(v, (lane - d) & (kWave - 1));
        if (gl >= d) v += t;
    }
    return v;
}

// Sum over aligned groups of U lanes (every lane gets the total).
template <int U, typename T>
__device__ __forceinline__ T unit_sum(T v)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = U >> 1; d >= 1; d >>= 1) v += bperm(v, lane ^ d);
    return v;
}

template <int U>
__device__ __forceinline__ int unit_min(int v)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = U >> 1; d >= 1; d >>= 1) v = min(v, bperm(v, lane ^ d));
    return v;
}

// Ballot restricted to this lane's unit, shifted down to bit 0.
template <int U>
__device__ __forceinline__ unsigned long long unit_ballot(bool p)
{
    const unsigned long long m = __ballot(p);
    if constexpr (U == 64) {
        return m;
    } else {
        const int base = lane_id() & ~(U - 1);
        return (m >> base) & ((1ull << U) - 1ull);
    }
}

// Wavefront-scope ordering point for LDS traffic between lanes of one wave.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Default B-row range of an A-nonzero: the whole B row.
struct FullRow {
    const int *__restrict__ brpt;
    __device__ __forceinline__ void operator()(int /*e*/, int k, int &bs, int &len) const
    {
        bs = brpt[k];
        len = brpt[k + 1] - bs;
    }
};

template <int U>
__device__ __forceinline__ int unit_max(int v)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = U >> 1; d >= 1; d >>= 1) v = max(v, bperm(v, lane ^ d));
    return v;
}

// Product consumers have two steps so the walkers below can issue the
// global reads of two products before either LDS update:
//   Item load(int b_index) const;          // global reads of one B entry
//   void apply(const Item &, V a, int q);  // update with A's value; q is the
//                                          // product's A-nonzero within the
//                                          // chunk (0..U-1)
// A consumer may also define
//   void chunk_end(int e, bool valid, bool shared);
// which the walk calls once per chunk after all of the chunk's products
// (e: this lane's A-nonzero index, valid: e is inside the row segment,
// shared: other units walked slices of the same chunk).

template <typename Op, typename = void>
struct has_chunk_end : std::false_type {
};
template <typename Op>
struct has_chunk_end<Op, std::void_t<decltype(&Op::chunk_end)>> : std::true_type {
};

// Instead of apply, a consumer may split its update into
//   Slot prep(const Item &);                          // LDS reads only
//   void commit(const Item &, const Slot &, V a, int q);  // the updates
// The walkers then issue the reads of every product of a step before any
// update.  LDS operations of a wave complete in order, so with a plain apply
// the read of each product waits for the atomics of the products before it;
// batched, the reads and atomics of a step overlap.
template <typename Op, typename = void>
struct has_prep : std::false_type {
};
template <typename Op>
struct has_prep<Op, std::void_t<decltype(&Op::prep)>> : std::true_type {
};

// A prep consumer may split commit further into
//   Post stage(const Item &, const Slot &, V a, int q, bool ok) const;  // arithmetic only
//   void post(const Post &);                                           // the updates
// stage runs for every product of a step (also for !ok ones, whose Slot
// is unset and whose Post must then do nothing), so the step's LDS reads
// are waited for once and its updates issue back to back; with commit,
// the branch around each product makes the compiler wait for all earlier
// LDS operations (the previous products' atomics included) before each one.
template <typename Op, typename = void>
struct has_post : std::false_type {
};
template <typename Op>
struct has_post<Op, std::void_t<decltype(&Op::post)>> : std::true_type {
};

template <int NP, typename Op, typename V>
__device__ __forceinline__ void apply_batch(Op &op, const typename Op::Item (&it)[NP], const V (&a)[NP],
                                            const int (&q)[NP], const bool (&ok)[NP])
{
    if constexpr (has_post<Op>::value) {
        decltype(op.prep(it[0])) sl[NP];
#pragma unroll
        for (int n = 0; n < NP; n++)
            if (ok[n]) sl[n] = op.prep(it[n]);
        decltype(op.stage(it[0], sl[0], a[0], q[0], ok[0])) ps[NP];
#pragma unroll
        for (int n = 0; n < NP; n++) ps[n] = op.stage(it[n], sl[n], a[n], q[n], ok[n]);
#pragma unroll
        for (int n = 0; n < NP; n++) op.post(ps[n]);
    } else if constexpr (has_prep<Op>::value) {
        decltype(op.prep(it[0])) sl[NP];
#pragma unroll
        for (int n = 0; n < NP; n++)
            if (ok[n]) sl[n] = op.prep(it[n]);
#pragma unroll
        for (int n = 0; n < NP; n++)
            if (ok[n]) op.commit(it[n], sl[n], a[n], q[n]);
    } else {
#pragma unroll
        for (int n = 0; n < NP; n++)
            if (ok[n]) op.apply(it[n], a[n], q[n]);
    }
}

// Grouped walk of one loaded chunk: GS lanes per A-nonzero, each group
// strides over its B row NPG entries per lane at a time (8 for groups of 4
// lanes, so a B row of up to 32 entries takes one step; 4 otherwise).  No
// per-product cross-lane search; used when the chunk's B rows have similar
// lengths.
template <int U, int GS, bool WITH_A, typename V, bool SLICE, typename Op>
__device__ __forceinline__ void chunk_grouped(int cnt, int bs, int len, V av, Op &op, int part, int nparts)
{
    static_assert(GS <= U, "group wider than unit");
    const int lane = lane_id();
    const int gl = lane & (U - 1);
    const int gbase = lane & ~(U - 1);
    constexpr int PER = U / GS;
    const int sub = gl & (GS - 1);
    const int grp = gl / GS;
    // Several parts (units sharing the chunk) split it by A-nonzero groups
    // (part p takes rounds p, p+nparts, ...) or, with SLICE, by slicing
    // every B row (part p takes steps p, p+nparts, ... of each row).
    // Slicing keeps all parts busy on rows with few A-nonzeros and long B
    // rows; group distribution suits many A-nonzeros with short B rows.
    // Each lane takes NPG entries of its group's B row per step (their loads
    // are issued together).
    constexpr int NPG = GS == 4 ? 8 : 4;
    const int t0 = SLICE ? sub + part * NPG * GS : sub;
    const int tstep = SLICE ? nparts * NPG * GS : NPG * GS;
    const int qstart = SLICE ? 0 : part * PER;
    const int qstep = SLICE ? PER : nparts * PER;
    for (int q0 = qstart; q0 < cnt; q0 += qstep) {
        const int q = q0 + grp;
        const int src = gbase + (q < U ? q : 0);
        // Every bpermute runs with the whole unit active: a source lane that
        // is inactive in the instruction would read back as 0.
        const int qb = bperm(bs, src);
        const int lq = bperm(len, src);
        const int ql = q < cnt ? lq : 0;
        V qa = V(0);
        if constexpr (WITH_A) qa = bperm(av, src);
        V aa[NPG];
        int qq[NPG];
#pragma unroll
        for (int n = 0; n < NPG; n++) {
            aa[n] = qa;
            qq[n] = q;
        }
        for (int t = t0; t < ql; t += tstep) {
            typename Op::Item it[NPG];
            bool ok[NPG];
#pragma unroll
            for (int n = 0; n < NPG; n++) {
                ok[n] = t + n * GS < ql;
                if (ok[n]) it[n] = op.load(qb + t + n * GS);
            }
            apply_batch<NPG>(op, it, aa, qq, ok);
        }
    }
}

// Largest q with excl[q] <= p over the unit (binary search, all lanes).
template <int U>
__device__ __forceinline__ void unit_locate(int excl, int gbase, int p0, int p1, int &q0, int &v0, int &q1, int &v1)
{
    q0 = 0;
    v0 = 0;
    q1 = 0;
    v1 = 0;
#pragma unroll
    for (int s = U >> 1; s >= 1; s >>= 1) {
        const int c0 = q0 + s, c1 = q1 + s;
        const int e0 = bperm(excl, gbase + c0);
        const int e1 = bperm(excl, gbase + c1);
        if (e0 <= p0) {
            q0 = c0;
            v0 = e0;
        }
        if (e1 <= p1) {
            q1 = c1;
            v1 = e1;
        }
    }
}

// Flattened walk of one loaded chunk: lane l takes products l, l+U, ... of
// the chunk (two per step) and finds each one's A-nonzero by binary search
// over the scanned B-row lengths.  Balanced for any length distribution.
template <int U, bool WITH_A, typename V, typename Op>
__device__ __forceinline__ void chunk_flat(int bs, int len, V av, Op &op, int part, int nparts)
{
    const int lane = lane_id();
    const int gl = lane & (U - 1);
    const int gbase = lane & ~(U - 1);
    const int incl = unit_incl_scan<U>(len);
    const int excl = incl - len;
    const int total = bperm(incl, gbase + U - 1);
    for (int base = part * 2 * U; base < total; base += nparts * 2 * U) {
        const int p0 = base + gl, p1 = p0 + U;
        int q0, v0, q1, v1;
        unit_locate<U>(excl, gbase, p0, p1, q0, v0, q1, v1);
        const int b0 = bperm(bs, gbase + q0);
        const int b1 = bperm(bs, gbase + q1);
        V a0 = V(0), a1 = V(0);
        if constexpr (WITH_A) {
            a0 = bperm(av, gbase + q0);
            a1 = bperm(av, gbase + q1);
        }
        const bool ok[2] = {p0 < total, p1 < total};
        typename Op::Item it[2];
        if (ok[0]) it[0] = op.load(b0 + (p0 - v0));
        if (ok[1]) it[1] = op.load(b1 + (p1 - v1));
        const V aa[2] = {a0, a1};
        const int qq[2] = {q0, q1};
        apply_batch<2>(op, it, aa, qq, ok);
    }
}

// Walks one loaded chunk of cnt A-nonzeros (lane l of the unit holds
// A-nonzero l's B range [bs, bs + len) and A value av; returns false when
// the chunk has no product) with the walk its
// length profile calls for: if the longest B range is at most twice the
// mean, lanes are grouped per A-nonzero (group width about half the mean,
// 4..U); otherwise products are flattened across the unit.
template <int U, bool WITH_A, typename V, bool SLICE = false, typename Op>
__device__ __forceinline__ bool walk_chunk(int cnt, int bs, int len, V av, Op &op, int part, int nparts)
{
    const int total = unit_sum<U>(len);
    const int mx = unit_max<U>(len);
    if (total == 0) return false;
    const int mean = (total + cnt - 1) / cnt;
    if (mx <= 2 * mean) {
        // Groups of about half the mean length: several products per lane per
        // round amortise the per-round cross-lane reads.  Below 32 lanes'
        // worth, groups of 4 lanes taking 8 entries each: fewer rounds than
        // groups of 8 taking 4.
        const int half = mean >> 1;
        const int gs = half >= 64 ? 64 : (half >= 32 ? 32 : (half >= 16 ? 16 : 4));
        if (gs <= 4) {
            chunk_grouped<U, 4, WITH_A, V, SLICE>(cnt, bs, len, av, op, part, nparts);
        } else if (gs == 8 || U == 8) {
            chunk_grouped<U, 8, WITH_A, V, SLICE>(cnt, bs, len, av, op, part, nparts);
        } else if (gs == 16 || U == 16) {
            if constexpr (U >= 16) chunk_grouped<U, 16, WITH_A, V, SLICE>(cnt, bs, len, av, op, part, nparts);
        } else if (gs == 32 || U == 32) {
            if constexpr (U >= 32) chunk_grouped<U, 32, WITH_A, V, SLICE>(cnt, bs, len, av, op, part, nparts);
        } else {
            if constexpr (U >= 64) chunk_grouped<U, 64, WITH_A, V, SLICE>(cnt, bs, len, av, op, part, nparts);
        }
    } else {
        chunk_flat<U, WITH_A, V>(bs, len, av, op, part, nparts);
    }
    return true;
}

// Calls the consumer for every product of A row segment [a_beg, a_end).
// `unit` of `nunits` units (all lanes of every unit must call this with the
// same arguments) share the work in one of two ways:
//  * rows with at least nunits*U A-nonzeros: unit u takes A chunks u,
//    u+nunits, ...;
//  * shorter rows (few A-nonzeros, possibly long B rows): every unit loads
//    every chunk and takes an interleaved slice of its products, so all
//    waves of a block stay busy.
// `range(e, k, bs, len)` gives the B entries [bs, bs+len) an A-nonzero e
// (column k) contributes.
//
// Each chunk of U A-nonzeros is walked by walk_chunk.
template <int U, bool WITH_A, typename V, bool SLICE = false, typename Range, typename Op>
__device__ __forceinline__ void for_each_product(int a_beg, int a_end, int unit, int nunits,
                                                 const int *__restrict__ acol, const V *__restrict__ aval,
                                                 const Range &range, Op &op)
{
    const int lane = lane_id();
    const int gl = lane & (U - 1);
    const bool shared = nunits > 1 && (a_end - a_beg) < nunits * U;
    const int c0 = shared ? a_beg : a_beg + unit * U;
    const int cstep = shared ? U : nunits * U;
    const int part = shared ? unit : 0;
    const int nparts = shared ? nunits : 1;
    for (int c = c0; c < a_end; c += cstep) {
        const int e = c + gl;
        int bs = 0, len = 0;
        V av = V(0);
        if (e < a_end) {
            const int k = acol[e];
            range(e, k, bs, len);
            if constexpr (WITH_A) av = aval[e];
        }
        const int cnt = min(U, a_end - c);
        if (!walk_chunk<U, WITH_A, V, SLICE>(cnt, bs, len, av, op, part, nparts)) continue;
        if constexpr (has_chunk_end<Op>::value) op.chunk_end(e, e < a_end, shared);
    }
}

// OR of `v` over the 64 lanes of the wave, returned to every lane.  DPP
// row shifts reduce each row of 16 lanes into its lane 15, the row
// broadcasts carry rows 0-2 into lane 63, and lane 63 is read back as a
// scalar: no LDS instruction, so nothing waits behind outstanding LDS work.
__device__ __forceinline__ unsigned wave_or(unsigned v)
{
    int x = static_cast<int>(v);
    x |= __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x |= __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x |= __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x |= __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
    x |= __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false); // row_bcast:15
    x |= __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false); // row_bcast:31
    return static_cast<unsigned>(__builtin_amdgcn_readlane(x, kWave - 1));
}

// Waits until every global store (and load) of this wave has completed.
// The compiler's wait counting treats loads and stores as one counter that
// may complete out of order when both are in flight, and then waits for
// everything before each use of a loaded value; a kernel that writes
// output and then walks products calls this between the two, once, so
// the walk's loads are waited for one buffer at a time.
__device__ __forceinline__ void drain_vmem()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt unconstrained
}

// Lane 0's value in every lane, held in scalar registers.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float wave_uniform(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double wave_uniform(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(b));
    const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Inclusive prefix sum over the 64 lanes of the wave: Hillis-Steele steps
// by DPP row shifts inside each row of 16 lanes, then the row broadcasts
// add row 0's total to row 1 and row 2's to row 3, and rows 0-1's total to
// rows 2-3.  No LDS instruction.
__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Block-wide stage of one batch of up to BS A-nonzeros for the block flat
// walk: the A-nonzeros with a non-empty B range, in order, with the
// exclusive scan of their range lengths (excl[n] = total products) and per
// entry the B index offset (B index = product position + off), the
// A-nonzero index and its value.
template <int BS, typename V>
struct BlockStage {
    static constexpr int NW = BS / kWave;
    struct alignas(16) Ent {
        int off;
        int e;
        V av;
    };
    int excl[BS + 1];
    Ent ent[BS];
    int wsum[2][NW];
};

// Stages one batch from per-thread ranges (thread t holds A-nonzero e with
// B range [bs, bs + len) and A value av; len = 0 for none) into `st`: the
// non-empty ranges in thread order.  Returns their number; `total` gets the
// batch's product count.  Every thread of the block calls it; it ends with a
// barrier.  The caller's previous walk of `st` must be complete (a barrier)
// before.
template <int BS, typename V>
__device__ __forceinline__ int block_stage_ranges(int e, int bs, int len, V av, BlockStage<BS, V> &st, int &total)
{
    constexpr int NW = BS / kWave;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid / kWave;
    const int f = len > 0 ? 1 : 0;
    const int il = wave_incl_scan(len);
    const int iff = wave_incl_scan(f);
    if (lane == kWave - 1) {
        st.wsum[0][wave] = il;
        st.wsum[1][wave] = iff;
    }
    __syncthreads();
    int pl = 0, pf = 0, tl = 0, tf = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const int a = st.wsum[0][w], b = st.wsum[1][w];
        if (w < wave) {
            pl += a;
            pf += b;
        }
        tl += a;
        tf += b;
    }
    if (f) {
        const int ex = pl + il - len;
        const int ix = pf + iff - f;
        st.excl[ix] = ex;
        st.ent[ix] = typename BlockStage<BS, V>::Ent{bs - ex, e, av};
    }
    if (tid == 0) st.excl[tf] = tl;
    __syncthreads();
    total = tl;
    return tf;
}

// Stages A-nonzeros [c, min(c + BS, a_end)) (one per thread) into `st`
// (aval may be nullptr: the staged values are then 0): each thread computes
// its own A-nonzero's range, so a range is computed once per batch whatever
// the number of waves.
template <int BS, typename V, typename Range>
__device__ __forceinline__ int block_stage(int c, int a_end, const int *__restrict__ acol, const V *__restrict__ aval,
                                           const Range &range, BlockStage<BS, V> &st, int &total)
{
    const int e = c + static_cast<int>(threadIdx.x);
    int bs = 0, len = 0;
    V av = V(0);
    if (e < a_end) {
        range(e, acol[e], bs, len);
        if (aval) av = aval[e];
    }
    return block_stage_ranges<BS, V>(e, bs, len, av, st, total);
}

// Walks every product of a staged batch with all waves of the block: wave
// w takes one contiguous slice of about total/NW products, and its lanes
// take 64 consecutive products per step (coalesced B reads whatever the
// range lengths).  The owner of each product comes from the step's range
// starts: lane l reads the start of range qa+1+l (qa owns the step's first
// product), the starts inside the step form a 64-bit mask (wave_or), and a
// lane's owner is qa plus the starts at or before it.  This costs the same
// few instructions per 64 products for any mix of long and short ranges and
// needs no per-product search.  NP steps are gathered at once and their B
// loads stay in flight while the previous steps' updates run.  Ends with a
// barrier (the stage may then be overwritten).
template <int BS, bool WITH_A, typename V, typename Op>
__device__ __forceinline__ void block_walk(const BlockStage<BS, V> &st, int n, int total, Op &op)
{
    constexpr int NW = BS / kWave;
    constexpr int NP = 4;
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const int seg = (((total + NW - 1) / NW) + kWave - 1) & ~(kWave - 1);
    const int s_beg = wave * seg;
    const int s_end = min(total, s_beg + seg);
    if (s_beg < s_end) {
        // Owner of s_beg: largest q < n with excl[q] <= s_beg, by a
        // 64-ary search in two rounds (n <= 64 * 64).
        static_assert(BS <= kWave * kWave, "two-round owner search");
        const int g1 = (n + kWave - 1) / kWave;
        int q = lane * g1;
        int cnt = __popcll(__ballot(q < n && st.excl[q] <= s_beg));
        const int lo = (cnt - 1) * g1;
        q = lo + lane;
        cnt = __popcll(__ballot(lane < g1 && q < n && st.excl[q] <= s_beg));
        int qa = lo + cnt - 1;
        const unsigned long long below = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull);
        // Wave-uniform copy of range qa: where the next range starts, and
        // qa's B offset, A-nonzero and A value.  A step that ends at or
        // before q_next lies in range qa and needs no LDS read at all.
        int q_next = 0, q_off = 0, q_e = 0;
        V q_av = V(0);
        auto load_q = [&]() {
            q_next = __builtin_amdgcn_readfirstlane(qa + 1 <= n ? st.excl[qa + 1] : INT_MAX);
            if (qa < n) {
                const typename BlockStage<BS, V>::Ent en = st.ent[qa];
                q_off = __builtin_amdgcn_readfirstlane(en.off);
                q_e = __builtin_amdgcn_readfirstlane(en.e);
                q_av = wave_uniform(en.av);
            }
        };
        load_q();
        // Steps past the slice, and lanes past its end, get ok = false and
        // the B index of the batch's first product (a valid entry), so
        // every load below is issued unconditionally.
        const int safe = st.ent[0].off;
        auto gather = [&](int b0, int (&bi)[NP], V (&aa)[NP], int (&qq)[NP], bool (&ok)[NP]) {
#pragma unroll
            for (int s = 0; s < NP; s++) {
                const int sb = b0 + s * kWave;
                ok[s] = false;
                bi[s] = safe;
                aa[s] = V(0);
                qq[s] = 0;
                if (sb < s_end) {
                    const int p = sb + lane;
                    ok[s] = p < s_end;
                    if (sb + kWave <= q_next) {
                        if (ok[s]) bi[s] = p + q_off;
                        aa[s] = q_av;
                        qq[s] = q_e;
                        if (sb + kWave == q_next) {
                            qa++;
                            load_q();
                        }
                    } else {
                        // Lane l reads the start of range qa+1+l; the starts
                        // inside the step form a mask, and a lane's owner is qa
                        // plus the starts at or before it.
                        const int qc = qa + 1 + lane;
                        const int pos = qc <= n ? st.excl[qc] - sb : kWave + 1;
                        const unsigned lo32 = pos > 0 && pos < 32 ? 1u << pos : 0u;
                        const unsigned hi32 = pos >= 32 && pos < kWave ? 1u << (pos - 32) : 0u;
                        const unsigned long long m =
                            (static_cast<unsigned long long>(wave_or(hi32)) << 32) | wave_or(lo32);
                        const int qo = qa + __popcll(m & below);
                        qa += __popcll(m) + (__ballot(pos == kWave) != 0ull ? 1 : 0);
                        if (ok[s]) {
                            const typename BlockStage<BS, V>::Ent en = st.ent[qo];
                            bi[s] = p + en.off;
                            aa[s] = en.av;
                            qq[s] = en.e;
                        }
                        load_q();
                    }
                }
            }
        };
        // Two buffers of NP steps in ping-pong, so no loaded value is copied
        // (a copy would wait for the load) and the loads of one buffer are in
        // flight while the other's updates run.  All loads are unconditional,
        // which lets the compiler wait for exactly the buffer it consumes.
        // Every load is issued inside the loop (none before it), so the
        // loop's entry holds no load in flight and the compiler's waits at
        // the loop head stay exact.  Buffer B starts empty (ok = false).
        const int stride = NP * kWave;
        int bbA[NP], qqA[NP], bbB[NP], qqB[NP];
        V aaA[NP], aaB[NP];
        bool okA[NP], okB[NP];
        typename Op::Item itA[NP], itB[NP];
#pragma unroll
        for (int s = 0; s < NP; s++) {
            okB[s] = false;
            qqB[s] = 0;
            aaB[s] = V(0);
            itB[s] = typename Op::Item{};
        }
        for (int base = s_beg;; base += 2 * stride) {
            gather(base, bbA, aaA, qqA, okA);
#pragma unroll
            for (int s = 0; s < NP; s++) itA[s] = op.load(bbA[s]);
            apply_batch<NP>(op, itB, aaB, qqB, okB);
            gather(base + stride, bbB, aaB, qqB, okB);
#pragma unroll
            for (int s = 0; s < NP; s++) itB[s] = op.load(bbB[s]);
            apply_batch<NP>(op, itA, aaA, qqA, okA);
            if (base + 2 * stride >= s_end) break;
        }
        apply_batch<NP>(op, itB, aaB, qqB, okB);
    }
    __syncthreads();
}

// Calls the consumer for every product of A row segment [a_beg, a_end) with
// the whole block, in batches of BS A-nonzeros (block_stage + block_walk).
// The consumer's q argument is the A-nonzero index.  Returns true when the
// row fit in one batch: `st` then still holds it for another walk over the
// same ranges (block_walk with the returned n and total).
template <int BS, bool WITH_A, typename V, typename Range, typename Op>
__device__ __forceinline__ bool block_for_each_product(int a_beg, int a_end, const int *__restrict__ acol,
                                                       const V *__restrict__ aval, const Range &range, Op &op,
                                                       BlockStage<BS, V> &st, int &n, int &total)
{
    n = 0;
    total = 0;
    for (int c = a_beg; c < a_end; c += BS) {
        n = block_stage<BS, V>(c, a_end, acol, aval, range, st, total);
        if (total > 0) block_walk<BS, WITH_A, V>(st, n, total, op);
    }
    return a_end - a_beg <= BS;
}

// ---- Chunk walk: every 64-lane step inside one B range ---------------------
//
// The block walk above flattens a batch's products, so a step may straddle
// several ranges and each such step pays an owner search.  For long-row
// inputs most products sit in long ranges (R-MAT 2^20, rows above 4096
// entries: 84% of a window's products come from ranges of 64 or more), and
// there a step that never leaves its range is much cheaper: a range of len
// products is cut into ceil(len / 64) chunks of one step each, the block's
// waves take contiguous runs of chunks, and a chunk's B indices, A value
// and A-nonzero are uniform values kept in scalar registers.  Lanes past a
// range's end idle (about 20% of lanes on those windows), which costs far
// less than the per-step ownership work it replaces.

template <int BS, typename V>
struct ChunkStage {
    static constexpr int NW = BS / kWave;
    struct Ent {
        int bs;   // B index of the range's first product
        int len;  // products in the range (> 0)
        int e;    // A-nonzero index
    };
    V av[BS];         // A value of each range (0 when the walk has none)
    int cex[BS + 1];  // exclusive scan of the ranges' chunk counts; cex[n] = chunks in the batch
    Ent ent[BS];
    int wsum[2][NW];
};

// LDS a kernel may hold per block on gfx950 (static + dynamic).
constexpr size_t kMaxLds = 160 * 1024;

// Stages one batch from per-thread ranges (thread t holds A-nonzero e with B
// range [bs, bs + len) and A value av; len = 0 for none): the non-empty
// ranges in thread order with their chunk offsets (chunks of CH products).
// Returns their number; `chunks` gets the batch's chunk count.  Every thread
// of the block calls it; it ends with a barrier.  The previous walk of `st`
// must be complete (a barrier) before.
template <int BS, typename V, int CH = kWave>
__device__ __forceinline__ int chunk_stage_ranges(int e, int bs, int len, V av, ChunkStage<BS, V> &st, int &chunks)
{
    constexpr int NW = BS / kWave;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid / kWave;
    const int f = len > 0 ? 1 : 0;
    const int nch = (len + CH - 1) / CH;
    const int ic = wave_incl_scan(nch);
    const int iff = wave_incl_scan(f);
    if (lane == kWave - 1) {
        st.wsum[0][wave] = ic;
        st.wsum[1][wave] = iff;
    }
    __syncthreads();
    int pc = 0, pf = 0, tc = 0, tf = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const int a = st.wsum[0][w], b = st.wsum[1][w];
        if (w < wave) {
            pc += a;
            pf += b;
        }
        tc += a;
        tf += b;
    }
    if (f) {
        const int ix = pf + iff - 1;
        st.cex[ix] = pc + ic - nch;
        st.ent[ix] = typename ChunkStage<BS, V>::Ent{bs, len, e};
        st.av[ix] = av;
    }
    if (tid == 0) st.cex[tf] = tc;
    __syncthreads();
    chunks = tc;
    return tf;
}

template <int BS, typename V, typename Range, int CH = kWave>
__device__ __forceinline__ int chunk_stage(int c, int a_end, const int *__restrict__ acol, const V *__restrict__ aval,
                                           const Range &range, ChunkStage<BS, V> &st, int &chunks)
{
    const int e = c + static_cast<int>(threadIdx.x);
    int bs = 0, len = 0;
    V av = V(0);
    if (e < a_end) {
        range(e, acol[e], bs, len);
        if (aval) av = aval[e];
    }
    return chunk_stage_ranges<BS, V, CH>(e, bs, len, av, st, chunks);
}

// Walks every product of a staged batch (n ranges, `chunks` chunks) with all
// waves of the block: wave w takes chunks [w * per, (w + 1) * per), per =
// ceil(chunks / NW), one 64-lane step per chunk, NP steps gathered at once
// in two ping-pong buffers (every load issued inside the loop, so the
// compiler waits for one buffer at a time).  The current range is held in
// scalar registers and advances by one when a step reaches the next range's
// first chunk (every staged range has at least one chunk).  Lanes past the
// range's end get ok = false and the batch's first B index, so every load is
// issued unconditionally and stays in bounds.  Ends with a barrier.
template <int BS, bool WITH_A, typename V, typename Op>
__device__ __forceinline__ void chunk_walk(const ChunkStage<BS, V> &st, int n, int chunks, Op &op)
{
    constexpr int NW = BS / kWave;
    constexpr int NP = 4;
    static_assert(BS <= kWave * kWave, "two-round owner search");
    const int lane = lane_id();
    // wave-uniform in scalar registers, so every branch on chunk indices
    // below is a scalar branch
    const int wave = wave_uniform(static_cast<int>(threadIdx.x / kWave));
    chunks = wave_uniform(chunks);
    n = wave_uniform(n);
    const int per = (chunks + NW - 1) / NW;
    const int j0 = wave * per;
    const int j1 = min(chunks, j0 + per);
    if (j0 < j1) {
        // Owner of chunk j0: largest q < n with cex[q] <= j0 (64-ary search in two rounds).
        const int g1 = (n + kWave - 1) / kWave;
        int q = lane * g1;
        int cnt = __popcll(__ballot(q < n && st.cex[q] <= j0));
        const int lo = (cnt - 1) * g1;
        q = lo + lane;
        cnt = __popcll(__ballot(lane < g1 && q < n && st.cex[q] <= j0));
        int qa = wave_uniform(lo + cnt - 1);
        int c_lo = 0, c_hi = 0, r_bs = 0, r_len = 0, r_e = 0;
        V r_av = V(0);
        auto load_q = [&]() {
            c_lo = wave_uniform(st.cex[qa]);
            c_hi = wave_uniform(st.cex[qa + 1]);
            const typename ChunkStage<BS, V>::Ent en = st.ent[qa];
            r_bs = wave_uniform(en.bs);
            r_len = wave_uniform(en.len);
            r_e = wave_uniform(en.e);
            if constexpr (WITH_A) r_av = wave_uniform(st.av[qa]);
        };
        load_q();
        const int safe = wave_uniform(st.ent[0].bs);
        auto gather = [&](int jb, int (&bi)[NP], V (&aa)[NP], int (&qq)[NP], bool (&ok)[NP]) {
#pragma unroll
            for (int s = 0; s < NP; s++) {
                const int jj = jb + s;
                ok[s] = false;
                bi[s] = safe;
                aa[s] = V(0);
                qq[s] = 0;
                if (jj < j1) {
                    if (jj == c_hi) {
                        qa++;
                        load_q();
                    }
                    const int pos = (jj - c_lo) * kWave + lane;
                    ok[s] = pos < r_len;
                    bi[s] = ok[s] ? r_bs + pos : safe;
                    aa[s] = r_av;
                    qq[s] = r_e;
                }
            }
        };
        int bbA[NP], qqA[NP], bbB[NP], qqB[NP];
        V aaA[NP], aaB[NP];
        bool okA[NP], okB[NP];
        typename Op::Item itA[NP], itB[NP];
#pragma unroll
        for (int s = 0; s < NP; s++) {
            okB[s] = false;
            qqB[s] = 0;
            aaB[s] = V(0);
            itB[s] = typename Op::Item{};
        }
        for (int jb = j0;; jb += 2 * NP) {
            gather(jb, bbA, aaA, qqA, okA);
#pragma unroll
            for (int s = 0; s < NP; s++) itA[s] = op.load(bbA[s]);
            apply_batch<NP>(op, itB, aaB, qqB, okB);
            gather(jb + NP, bbB, aaB, qqB, okB);
#pragma unroll
            for (int s = 0; s < NP; s++) itB[s] = op.load(bbB[s]);
            apply_batch<NP>(op, itA, aaA, qqA, okA);
            if (jb + 2 * NP >= j1) break;
        }
        apply_batch<NP>(op, itB, aaB, qqB, okB);
    }
    __syncthreads();
}

// Every product of A row segment [a_beg, a_end) with the whole block, in
// batches of BS A-nonzeros (chunk_stage + chunk_walk).  The consumer's q
// argument is the A-nonzero index.  Returns true when the row fit in one
// batch: `st` then still holds it for another walk over the same ranges
// (chunk_walk with the returned n and chunks).
template <int BS, bool WITH_A, typename V, typename Range, typename Op>
__device__ __forceinline__ bool chunk_for_each_product(int a_beg, int a_end, const int *__restrict__ acol,
                                                       const V *__restrict__ aval, const Range &range, Op &op,
                                                       ChunkStage<BS, V> &st, int &n, int &chunks)
{
    n = 0;
    chunks = 0;
    for (int c = a_beg; c < a_end; c += BS) {
        n = chunk_stage<BS, V>(c, a_end, acol, aval, range, st, chunks);
        if (chunks > 0) chunk_walk<BS, WITH_A, V>(st, n, chunks, op);
    }
    return a_end - a_beg <= BS;
}

// First index in [lo, t) of the sorted B row segment whose column is >=
// bound (t if none): galloping from lo, then a binary search, so the cost
// grows with the distance from lo rather than with the row length.
__device__ __forceinline__ int gallop_lower(const int *__restrict__ bcol, int lo, int t, int bound)
{
    if (lo >= t || bcol[lo] >= bound) return lo;
    int prev = lo, step = 1;
    while (prev + step < t && bcol[prev + step] < bound) {
        prev += step;
        step <<= 1;
    }
    int l = prev + 1, h = min(prev + step, t);
    while (l < h) {
        const int m = (l + h) >> 1;
        if (bcol[m] < bound)
            l = m + 1;
        else
            h = m;
    }
    return l;
}

// First index in [lo, t) of a B row whose column is >= bound (t if none),
// between the cuts of bound's unit in the unit cut table (u: the row's cuts
// at bound's unit and the next): entries before the unit's cut are below
// bound, entries from the next unit's cut on are above it, and a bound on a
// unit boundary is the cut itself.
__device__ __forceinline__ int unit_lower(const int *__restrict__ bcol, const int *__restrict__ u, int lo, int t,
                                          int bound)
{
    const int l = max(lo, u[0]);
    if ((bound & ((1 << kUnitCutShift) - 1)) == 0) return min(t, l);
    int h = max(l, min(t, u[1]));
    int x = l;
    while (x < h) {
        const int m = (x + h) >> 1;
        if (bcol[m] < bound)
            x = m + 1;
        else
            h = m;
    }
    return min(t, x);
}

// Cursor-carrying B range for window [w0, wend): lo is the first entry of
// the B row not consumed by earlier windows; hi the first entry >= wend,
// found between the cuts of the unit cut table when one is given (ucut,
// ustride ints per B row), else by galloping from lo.  With next_col, tracks
// the smallest column beyond the window so empty windows can be skipped.
struct CursorRange {
    const int *__restrict__ brpt;
    const int *__restrict__ bcol;
    const int *__restrict__ cur_in;  // position after the previous window (unused on the first)
    int *__restrict__ cur_out;       // position after this window
    int wend;
    bool first;
    bool last;  // wend covers all columns
    int *next_col;
    const int *__restrict__ ucut = nullptr;
    int ustride = 0;
    __device__ __forceinline__ void operator()(int e, int k, int &bs, int &len) const
    {
        const int t = brpt[k + 1];
        const int lo = first ? brpt[k] : cur_in[e];
        int hi;
        if (last)
            hi = t;
        else if (ucut)
            hi = unit_lower(bcol, ucut + static_cast<long long>(k) * ustride + (wend >> kUnitCutShift), lo, t, wend);
        else
            hi = gallop_lower(bcol, lo, t, wend);
        cur_out[e] = hi;
        if (next_col && hi < t) *next_col = min(*next_col, bcol[hi]);
        bs = lo;
        len = hi - lo;
    }
};

// B range of an A-nonzero inside column window wi, from the window-cut
// table (entries [cut[k][wi], cut[k][wi + 1]) of B row k; cut[k][0] is the
// row start and cut[k][nwin] the row end): two independent loads instead of
// a search.  Tracks the smallest column beyond the window when next_col is
// given.
struct WinCutRange {
    const int *__restrict__ wcut;
    const int *__restrict__ bcol;
    int stride;  // windows + 1
    int wi;
    int *next_col;
    __device__ __forceinline__ void operator()(int /*e*/, int k, int &bs, int &len) const
    {
        const int *c = wcut + static_cast<long long>(k) * stride;
        const int lo = c[wi], hi = c[wi + 1];
        if (next_col) {
            const int t = c[stride - 1];
            if (hi < t) *next_col = min(*next_col, bcol[hi]);
        }
        bs = lo;
        len = hi - lo;
    }
};

// Sub-window [.., cb) of column window wi: the range starts at the window's
// cut (first sub-window) or where the previous sub-window stopped, and ends
// at the first column >= cb.  With the unit cut table (ucut: the first
// entry of every B row at each multiple of 2^kUnitCutShift columns, ustride
// entries per row) that entry lies between the cuts of cb's unit, a short
// binary search loaded alongside the window cut; without it, it is found by
// galloping from the start, no further than the window's end cut.
struct CutCursorRange {
    const int *__restrict__ wcut;
    const int *__restrict__ bcol;
    int stride;
    int wi;
    const int *__restrict__ cur_in;
    int *__restrict__ cur_out;
    int cb;
    bool first;
    bool to_end;  // cb is the window end
    const int *__restrict__ ucut = nullptr;
    int ustride = 0;
    __device__ __forceinline__ void operator()(int e, int k, int &bs, int &len) const
    {
        const int *c = wcut + static_cast<long long>(k) * stride;
        const int t = c[wi + 1];
        const int lo = first ? c[wi] : cur_in[e];
        int hi;
        if (to_end) {
            hi = t;
        } else if (ucut) {
            hi = unit_lower(bcol, ucut + static_cast<long long>(k) * ustride + (cb >> kUnitCutShift), lo, t, cb);
        } else {
            hi = gallop_lower(bcol, lo, t, cb);
        }
        cur_out[e] = hi;
        bs = lo;
        len = hi - lo;
    }
};

// Home slot of a column key, two families:
//  * LIN: the reference's multiplicative hash col * 107 mod T
//    (symbolic.cuh:98) on the full-rate 24-bit multiply.  Consecutive
//    columns land 107 slots (11 LDS banks) apart, which makes it nearly
//    collision- and conflict-free on rows made of runs of adjacent columns
//    (stencils, banded and uniformly random patterns): used for the short-row
//    bins.
//  * FIB: multiplicative (Fibonacci) hashing in 24-bit words,
//    h = ((col * 0x9E3779) mod 2^24) >> (24 - log2 T), 0x9E3779 = 2^24 /
//    golden ratio.  Its top bits depend on every one of the key's low 24
//    bits; power-law (R-MAT) column indices have strongly skewed low bits,
//    which piles keys onto a few home slots under LIN.  Used for the long-row
//    bins, where power-law rows put most of their products.
constexpr unsigned kLinMul = 107u;
constexpr unsigned kFibMul = 0x9E3779u;
template <int T, int H>
__device__ __forceinline__ unsigned hash_home(int key)
{
    static_assert(T >= 2 && T <= (1 << 24) && (T & (T - 1)) == 0, "table size");
    if constexpr (H == kHashLin) {
        return __umul24(static_cast<unsigned>(key), kLinMul) & (T - 1);
    } else {
        constexpr int kLog = __builtin_ctz(T);
        return (__umul24(static_cast<unsigned>(key), kFibMul) & 0xFFFFFFu) >> (24 - kLog);
    }
}

// Linear-probing insert of `key` (>= 0) into an int table of T (power of 2)
// slots initialised to -1.  Returns the slot; `fresh` tells whether this
// call created the key.  A plain LDS read comes first: most products of a
// row repeat a key that is already in the table, and ds_read is cheaper than
// ds_cmpst; the CAS is only issued on an empty slot.
template <int T, int H>
__device__ __forceinline__ int hash_insert(int *tab, int key, bool &fresh)
{
    unsigned h = hash_home<T, H>(key);
    while (true) {
        int cur = tab[h];
        if (cur == -1) cur = atomicCAS(tab + h, -1, key);
        if (cur == -1) {
            fresh = true;
            return static_cast<int>(h);
        }
        if (cur == key) {
            fresh = false;
            return static_cast<int>(h);
        }
        h = (h + 1) & (T - 1);
    }
}

// Resolves a RowList on the device: returns the row count and sets `rows`
// to the list (nullptr: rows 0..n-1).  The values are uniform across the
// grid (scalar loads).
__device__ __forceinline__ int row_list(const RowList &rl, const int *&rows)
{
    rows = rl.rows;
    if (!rl.dsize) return rl.n;
    int off = 0;
    for (int b = 0; b < rl.bin; b++) off += rl.dsize[b];
    rows += off;
    return rl.dsize[rl.bin];
}

// Tile loops of per-bin kernels over a RowList: the first tile of this block
// and the one after `cur`.  Host-sized lists (no counter) take tiles by a
// fixed stride (XCD-aware for tiles, by block index for rows).  With a
// counter, thread 0 takes `chunk` consecutive tiles at a time from it (about
// eight takes per block: one atomic per tile would serialise the grid on a
// single address) and `slot` (a __shared__ int) hands the first to the
// block.  Every thread of the block calls these, with the block's LDS use
// of the previous tile complete.
struct TileSched {
    int end = 0;    // end of the chunk being worked
    int chunk = 1;  // tiles per take
};
__device__ __forceinline__ int xcd_tile();
__device__ __forceinline__ int block_take(const RowList &rl, TileSched &ts, int *slot)
{
    __syncthreads();  // every thread has read the previous take from slot
    if (threadIdx.x == 0) *slot = atomicAdd(rl.next, ts.chunk);
    __syncthreads();
    const int t = *slot;
    ts.end = t + ts.chunk;
    return t;
}
__device__ __forceinline__ int block_first(const RowList &rl, int ntiles, TileSched &ts, int *slot, int fixed)
{
    if (!rl.next) return fixed;
    ts.chunk = max(1, ntiles / (static_cast<int>(gridDim.x) * 8));
    return block_take(rl, ts, slot);
}
__device__ __forceinline__ int block_next(const RowList &rl, int cur, TileSched &ts, int *slot)
{
    if (!rl.next) return cur + static_cast<int>(gridDim.x);
    if (cur + 1 < ts.end) return cur + 1;
    return block_take(rl, ts, slot);
}
__device__ __forceinline__ int block_first_tile(const RowList &rl, int ntiles, TileSched &ts, int *slot)
{
    return block_first(rl, ntiles, ts, slot, rl.next ? 0 : xcd_tile());
}
__device__ __forceinline__ int block_first_row(const RowList &rl, int nrows, TileSched &ts, int *slot)
{
    return block_first(rl, nrows, ts, slot, static_cast<int>(blockIdx.x));
}

// Blocks are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b % 8).  With gridDim.x a multiple of 8, this returns a logical tile index
// such that each XCD walks one contiguous range of tiles, so rows that share
// B rows meet in the same L2.
constexpr int kNumXcd = 8;
__device__ __forceinline__ int xcd_tile()
{
    const int per = gridDim.x / kNumXcd;
    return (static_cast<int>(blockIdx.x) % kNumXcd) * per + static_cast<int>(blockIdx.x) / kNumXcd;
}
__host__ __device__ constexpr int xcd_grid(int tiles) { return (tiles + kNumXcd - 1) / kNumXcd * kNumXcd; }

// Lookup; returns -1 when the key is absent (the probe stops at an empty
// slot or after T probes, so a mismatched pattern can never spin).
template <int T, int H>
__device__ __forceinline__ int hash_find(const int *tab, int key)
{
    unsigned h = hash_home<T, H>(key);
    for (int i = 0; i < T; i++) {
        const int v = tab[h];
        if (v == key) return static_cast<int>(h);
        if (v == -1) return -1;
        h = (h + 1) & (T - 1);
    }
    return -1;
}

// Bitonic sort of K*U int keys held K per lane across a unit of U lanes;
// element j = k*U + (lane % U) ends up holding the j-th smallest key.
template <int U, int K>
__device__ __forceinline__ void unit_bitonic_sort_keys(int (&key)[K])
{
    const int lane = lane_id();
    const int gl = lane & (U - 1);
    constexpr int NEL = K * U;
#pragma unroll
    for (int size = 2; size <= NEL; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= U) {
                const int ks = stride / U;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int pk = k ^ ks;
                    if (pk > k) {
                        const bool asc = ((k * U + gl) & size) == 0;
                        const int a = key[k], b = key[pk];
                        key[k] = asc ? min(a, b) : max(a, b);
                        key[pk] = asc ? max(a, b) : min(a, b);
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int j = k * U + gl;
                    const bool asc = (j & size) == 0;
                    const bool lower = (j & stride) == 0;
                    const int o = bperm(key[k], lane ^ stride);
                    key[k] = (asc == lower) ? min(key[k], o) : max(key[k], o);
                }
            }
        }
    }
}

// Bitonic sort of BS*K int keys over a block.  Thread (wave w, lane l)
// holds keys k = 0..K-1 at positions w*64*K + k*64 + l; on return position j
// holds the j-th smallest key.  Strides below 64*K run in registers (key
// swaps) and lane exchanges; only longer strides go through `lds` (BS*K
// ints), between barriers: 18 barriers for 4096 keys instead of one per
// network stage.  Every thread of the block calls it.
template <int BS, int K>
__device__ __forceinline__ void block_bitonic_sort_keys(int (&key)[K], int *lds)
{
    constexpr int WE = kWave * K;
    constexpr int NEL = BS * K;
    static_assert((K & (K - 1)) == 0 && (NEL & (NEL - 1)) == 0, "powers of two");
    const int lane = lane_id();
    const int base = static_cast<int>(threadIdx.x / kWave) * WE;
#pragma unroll 1
    for (int size = 2; size <= NEL; size <<= 1) {
        if (size > WE) {
            // strides size/2 .. WE cross waves
#pragma unroll
            for (int k = 0; k < K; k++) lds[base + k * kWave + lane] = key[k];
            __syncthreads();
#pragma unroll 1
            for (int stride = size >> 1; stride >= WE; stride >>= 1) {
#pragma unroll
                for (int i0 = 0; i0 < NEL / 2; i0 += BS) {
                    const int i = i0 + static_cast<int>(threadIdx.x);
                    const int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));
                    const int hi = lo + stride;
                    const bool asc = (lo & size) == 0;
                    const int a = lds[lo], b = lds[hi];
                    if ((a > b) == asc) {
                        lds[lo] = b;
                        lds[hi] = a;
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int k = 0; k < K; k++) key[k] = lds[base + k * kWave + lane];
            __syncthreads();
        }
        // strides below WE and at least 64: swaps between a lane's keys
#pragma unroll 1
        for (int stride = min(size >> 1, WE >> 1); stride >= kWave; stride >>= 1) {
            const int ks = stride / kWave;
#pragma unroll
            for (int m = 1; m < K; m <<= 1) {
                if (ks != m) continue;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if ((k ^ m) > k) {
                        const bool asc = ((base + k * kWave) & size) == 0;
                        const int a = key[k], b = key[k ^ m];
                        key[k] = asc ? min(a, b) : max(a, b);
                        key[k ^ m] = asc ? max(a, b) : min(a, b);
                    }
                }
            }
        }
        // strides below 64: lane exchanges
#pragma unroll 1
        for (int stride = min(size >> 1, kWave >> 1); stride > 0; stride >>= 1) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int j = base + k * kWave + lane;
                const bool asc = (j & size) == 0;
                const bool lower = (j & stride) == 0;
                const int o = bperm(key[k], lane ^ stride);
                key[k] = (asc == lower) ? min(key[k], o) : max(key[k], o);
            }
        }
    }
}

// Block-wide exclusive scan of one int per thread; returns the exclusive
// prefix and writes the block total.  `scratch` needs BS/64 + 1 ints of LDS.
template <int BS>
__device__ __forceinline__ int block_excl_scan(int v, int *scratch, int &total)
{
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    const int incl = unit_incl_scan<kWave>(v);
    if (lane == kWave - 1) scratch[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < BS / kWave; w++) {
            const int t = scratch[w];
            scratch[w] = run;
            run += t;
        }
        scratch[BS / kWave] = run;
    }
    __syncthreads();
    total = scratch[BS / kWave];
    const int r = scratch[wid] + incl - v;
    __syncthreads();
    return r;
}

template <int BS, typename T>
__device__ __forceinline__ T block_sum(T v, T *scratch)
{
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    v = unit_sum<kWave>(v);
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x == 0) {
        for (int w = 0; w < BS / kWave; w++) s += scratch[w];
        scratch[0] = s;
    }
    __syncthreads();
    s = scratch[0];
    __syncthreads();
    return s;
}

template <int BS>
__device__ __forceinline__ int block_min(int v, int *scratch)
{
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    v = unit_min<kWave>(v);
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    int m = scratch[0];
#pragma unroll
    for (int w = 1; w < BS / kWave; w++) m = min(m, scratch[w]);
    __syncthreads();
    return m;
}

}  // namespace opsamd
