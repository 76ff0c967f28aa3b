// hj_fp64.hip -- the plain FP64 distance-count kernel (the drop-in path).
//
// Replaces the reference's 8x8 AVX-512 tile (EGO_LOOP_TRAN egojoin.h:311-392,
// EGO_LOOP_TRAN_NOSELF :482-553) and its count epilogue (egojoin.cpp:800-815,
// :868-884).  Differences by design:
//   * one wavefront owns a 64 x 64 tile pair: lane = column point, the row
//     points stream through scalar registers (uniform loads), so no LDS and no
//     register transposes are needed;
//   * the distance is accumulated in the fixed-order difference form
//     s = fma(a_k - b_k, a_k - b_k, s), whose partial sums never decrease, so
//     the KBLOCK early exit (egojoin.h:349-359) is exact and the count equals
//     the FP64 oracle bit for bit (the reference's expanded form
//     eps^2/2 - |a|^2/2 - |b|^2/2 + a.b cancels and can flip pairs within
//     rounding of eps);
//   * the early exit is a wavefront vote over R rows x 64 columns;
//   * counts are ballot + popcount per row, one 64-bit atomic per wave at the
//     end (the reference's sign-bit sum + OpenMP reduction);
//   * the diagonal rule (egojoin.h:383-392) is the predicate i < j on sorted
//     indices inside diagonal tiles;
//   * a persistent, occupancy-sized grid drains the tile-pair list in
//     interleaved chunks (hj_work.h), replacing the static load partition of
//     i-blocks over threads (egojoin.h:207-222).
#include <algorithm>

#include "hj_internal.h"
#include "hj_queue.h"
#include "hj_work.h"

namespace hj {

namespace {

// Row values are read through the constant address space: the rows are never
// written while a join kernel runs and the row address is wavefront-uniform,
// so every row value comes in through scalar loads (merged into 64-byte
// loads) even though the kernel also issues atomics inside its loop.
typedef const __attribute__((address_space(4))) double const_f64;

// Evaluates rows [i0, i0+R) of tile I against the 64 column points of the
// calling lanes.  Rows come from the tile layout too ([dim][64] per tile):
// ar points at row i0 of dimension 0, so row r, dimension k is ar[k * 64 + r]
// and the R rows of one dimension are one contiguous scalar load.  Returns
// the number of dimensions evaluated (a multiple of KB) and leaves in acc[r]
// the final fixed-order sums, or sums already above eps2 for every unmasked
// pair when the wave exited early (then *alive = false).
template <int KB, int R>
__device__ __forceinline__ int row_group(const_f64 *__restrict__ ar, int dpad,
                                         const double *__restrict__ cb, uint32_t lane,
                                         const double (&b0)[KB], const bool (&m)[R],
                                         double eps2, double (&acc)[R], bool *alive) {
    // a masked pair starts at +inf: it never counts and never keeps the group
    // alive, so the vote is one minimum over the rows (vector ops) instead of
    // a masked test per row (scalar mask arithmetic)
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = m[r] ? 0.0 : INFINITY;
    int c = 0;
    while (c < dpad) {
        // the first KB column values are loaded once per tile pair (every row
        // group evaluates them)
        double b[KB];
#pragma unroll
        for (int k = 0; k < KB; k++) b[k] = c == 0 ? b0[k] : cb[(uint32_t)(c + k) * 64 + lane];
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
            for (int k = 0; k < KB; k++) {
                double t = ar[(uint32_t)(c + k) * 64 + r] - b[k];
                acc[r] = __builtin_fma(t, t, acc[r]);
            }
        }
        c += KB;
        double mn = acc[0];
#pragma unroll
        for (int r = 1; r < R; r++) mn = fmin(mn, acc[r]);
        if (!__any(mn <= eps2)) {
            *alive = false;
            return c;
        }
    }
    *alive = true;
    return c;
}

// Counts of consecutive work items usually belong to one segment; they are
// flushed to that segment's counter when the segment changes.
struct SegCount {
    uint32_t seg = 0xFFFFFFFFu;
    uint64_t cnt = 0;
    __device__ __forceinline__ void set(uint32_t s, uint64_t *seg_counts) {
        if (s != seg) {
            flush(seg_counts);
            seg = s;
        }
    }
    __device__ __forceinline__ void flush(uint64_t *seg_counts) {
        if (seg != 0xFFFFFFFFu && cnt && (threadIdx.x & 63) == 0)
            atomicAdd((unsigned long long *)&seg_counts[seg], (unsigned long long)cnt);
        cnt = 0;
    }
};

template <bool SELF, int KB, int R>
__global__ __launch_bounds__(256) void fp64_count_kernel(
    const double *__restrict__ rows, const double *__restrict__ cols,
    const uint64_t *__restrict__ pairs, uint64_t npairs, const SegDev *__restrict__ segs,
    const uint32_t *__restrict__ tile_seg, int dpad, uint64_t *__restrict__ ctr,
    uint64_t *__restrict__ seg_counts) {
    const uint32_t lane = threadIdx.x & 63;
    uint64_t cnt = 0, work = 0, cand = 0;
    SegCount sc;
    WorkCursor<4, false> cur(npairs);
    // segment of the current query tile (work items of one query tile are
    // consecutive in the list, so the lookup is rare)
    uint32_t curI = 0xFFFFFFFFu, nvi = 0, ttile0 = 0;
    uint64_t tn = 0;
    double eps2 = 0.0;
    uint64_t w;
    while (cur.next(w)) {
        const uint64_t pr = pairs[w];
        const uint32_t I = (uint32_t)pr, J = (uint32_t)(pr >> 32);
        if (I != curI) {
            const uint32_t sg = tile_seg[I];
            const SegDev &S = segs[sg];
            sc.set(sg, seg_counts);
            eps2 = S.eps2;
            nvi = tile_points(S.qn - 64ull * (I - S.qtile0));
            tn = S.tn;
            ttile0 = S.ttile0;
            curI = I;
        }
        const uint32_t nvj = tile_points(tn - 64ull * (J - ttile0));
        const uint32_t j = J * 64 + lane;  // layout column
        const bool vj = lane < nvj;
        const uint32_t i_beg = I * 64;
        const uint32_t i_end = i_beg + nvi;
        const bool diag = SELF && I == J;
        cand += diag ? (uint64_t)nvj * (nvj - 1) / 2 : (uint64_t)nvi * nvj;
        const double *cb = cols + (size_t)J * dpad * 64;  // wave-uniform tile base
        double b0[KB];
#pragma unroll
        for (int k = 0; k < KB; k++) b0[k] = cb[(uint32_t)k * 64 + lane];
        for (uint32_t i0 = i_beg; i0 < i_end; i0 += R) {
            bool m[R];
            bool anym = false;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t i = i0 + r;
                m[r] = vj && i < i_end && (!diag || i < j);
                anym |= m[r];
            }
            if (!__any(anym)) continue;
            double acc[R];
            bool alive;
            const_f64 *ar = (const_f64 *)(rows + (size_t)I * dpad * 64 + (i0 - i_beg));
            work += (uint64_t)row_group<KB, R>(ar, dpad, cb, lane, b0, m, eps2, acc, &alive) * (R * 64);
            if (alive) {
                uint64_t c = 0;
#pragma unroll
                for (int r = 0; r < R; r++) c += __popcll(__ballot(acc[r] <= eps2));
                cnt += c;
                sc.cnt += c;
            }
        }
    }
    sc.flush(seg_counts);
    if (lane == 0) {
        atomicAdd((unsigned long long *)&ctr[1], (unsigned long long)cnt);
        atomicAdd((unsigned long long *)&ctr[2], (unsigned long long)work);
        atomicAdd((unsigned long long *)&ctr[3], (unsigned long long)cand);
    }
}

// Pair output: same traversal; every accepted (row, column) pair of layout
// positions is appended as (row | column << 32) through the wavefront's LDS
// queue (cursor ctr[6]).  Appends past out_cap are counted but not written.
template <bool SELF, int KB, int R>
__global__ __launch_bounds__(256) void fp64_pairs_kernel(
    const double *__restrict__ rows, const double *__restrict__ cols,
    const uint64_t *__restrict__ pairs, uint64_t npairs, const SegDev *__restrict__ segs,
    const uint32_t *__restrict__ tile_seg, int dpad, uint64_t *__restrict__ ctr,
    uint64_t *__restrict__ out, uint64_t out_cap) {
    const uint32_t lane = threadIdx.x & 63;
    __shared__ QueueLds<1> qs;
    queue_clear_fills(qs);
    if (threadIdx.x == 0) queue_setup(qs, 0, out, out_cap, &ctr[6]);
    __syncthreads();
    WorkCursor<4, false> cur(npairs);
    uint32_t curI = 0xFFFFFFFFu, nvi = 0, ttile0 = 0;
    uint64_t tn = 0;
    double eps2 = 0.0;
    uint64_t w;
    while (cur.next(w)) {
        const uint64_t pr = pairs[w];
        const uint32_t I = (uint32_t)pr, J = (uint32_t)(pr >> 32);
        if (I != curI) {
            const SegDev &S = segs[tile_seg[I]];
            eps2 = S.eps2;
            nvi = tile_points(S.qn - 64ull * (I - S.qtile0));
            tn = S.tn;
            ttile0 = S.ttile0;
            curI = I;
        }
        const uint32_t nvj = tile_points(tn - 64ull * (J - ttile0));
        const uint32_t j = J * 64 + lane;
        const bool vj = lane < nvj;
        const uint32_t i_beg = I * 64;
        const uint32_t i_end = i_beg + nvi;
        const bool diag = SELF && I == J;
        const double *cb = cols + (size_t)J * dpad * 64;  // wave-uniform tile base
        double b0[KB];
#pragma unroll
        for (int k = 0; k < KB; k++) b0[k] = cb[(uint32_t)k * 64 + lane];
        for (uint32_t i0 = i_beg; i0 < i_end; i0 += R) {
            bool m[R];
            bool anym = false;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t i = i0 + r;
                m[r] = vj && i < i_end && (!diag || i < j);
                anym |= m[r];
            }
            if (!__any(anym)) continue;
            double acc[R];
            bool alive;
            const_f64 *ar = (const_f64 *)(rows + (size_t)I * dpad * 64 + (i0 - i_beg));
            (void)row_group<KB, R>(ar, dpad, cb, lane, b0, m, eps2, acc, &alive);
            if (!alive) continue;
#pragma unroll
            for (int r = 0; r < R; r++)  // masked pairs hold +inf
                queue_push<1, 0>(qs, acc[r] <= eps2, (uint64_t)(i0 + r) | ((uint64_t)j << 32));
        }
    }
    queue_drain<1, 0>(qs);
}

}  // namespace

// Workgroups of the FP64 kernels: by default about FP64_CHUNKS_PER_WAVE chunks
// per wavefront (at least the resident count), so the hardware dispatcher
// hands out the list in order and the resident wavefronts work on one window
// of it; measured 2x faster than one persistent wave per resident slot at
// n = 1M, d = 64.
constexpr uint64_t FP64_CHUNKS_PER_WAVE = 4;

template <class K> int fp64_grid(K kernel, const JoinArgs &ja) {
    if (ja.grid > 0) return ja.grid;
    const int resident = persistent_grid(kernel, 256, ja.num_cus, 0);
    const uint64_t chunks = (ja.npairs + 3) / 4;
    const uint64_t want = (chunks + 4 * FP64_CHUNKS_PER_WAVE - 1) / (4 * FP64_CHUNKS_PER_WAVE);
    return (int)std::min<uint64_t>(std::max<uint64_t>(want, (uint64_t)resident), 1u << 30);
}

template <bool SELF, int KB, int R>
void fp64_count_launch(const JoinArgs &ja, const double *rows, const double *cols, int dpad,
                       hipStream_t s) {
    auto k = fp64_count_kernel<SELF, KB, R>;
    const int g = fp64_grid(k, ja);
    k<<<g, 256, 0, s>>>(rows, cols, ja.pairs, ja.npairs, ja.segs, ja.tile_seg, dpad, ja.ctr,
                        ja.seg_counts);
}

template <bool SELF, int KB>
void fp64_count_rows(const JoinArgs &ja, const double *rows, const double *cols, int dpad,
                     hipStream_t s) {
    if (ja.tile_rows == 8)
        fp64_count_launch<SELF, KB, 8>(ja, rows, cols, dpad, s);
    else
        fp64_count_launch<SELF, KB, 16>(ja, rows, cols, dpad, s);
}

void launch_fp64_count(const JoinArgs &ja, const double *rows, const double *cols, int dpad,
                       bool self, int kblock, hipStream_t s) {
    if (ja.npairs == 0) return;
    if (!work_list_fits(ja.npairs, 4)) throw InvalidArg{"work list too long"};
    if (dpad % kblock != 0) throw InvalidArg{"internal: tile rows not whole early-exit blocks"};
    if (kblock == 4) {
        if (self)
            fp64_count_rows<true, 4>(ja, rows, cols, dpad, s);
        else
            fp64_count_rows<false, 4>(ja, rows, cols, dpad, s);
    } else if (kblock == 8) {
        if (self)
            fp64_count_rows<true, 8>(ja, rows, cols, dpad, s);
        else
            fp64_count_rows<false, 8>(ja, rows, cols, dpad, s);
    } else {
        // KBLOCK = 16, the setting recommended for uniform data (FAQ.md:14-22);
        // the caller lays tiles out with dpad a multiple of 16
        if (self)
            fp64_count_rows<true, 16>(ja, rows, cols, dpad, s);
        else
            fp64_count_rows<false, 16>(ja, rows, cols, dpad, s);
    }
    HJ_HIP(hipGetLastError());
}

void launch_fp64_pairs(const JoinArgs &ja, const double *rows, const double *cols, int dpad,
                       bool self, uint64_t *out, uint64_t out_cap, hipStream_t s) {
    if (ja.npairs == 0) return;
    constexpr int R = 8;
    if (self) {
        auto k = fp64_pairs_kernel<true, 4, R>;
        k<<<fp64_grid(k, ja), 256, 0, s>>>(
            rows, cols, ja.pairs, ja.npairs, ja.segs, ja.tile_seg, dpad, ja.ctr, out, out_cap);
    } else {
        auto k = fp64_pairs_kernel<false, 4, R>;
        k<<<fp64_grid(k, ja), 256, 0, s>>>(
            rows, cols, ja.pairs, ja.npairs, ja.segs, ja.tile_seg, dpad, ja.ctr, out, out_cap);
    }
    HJ_HIP(hipGetLastError());
}

}  // namespace hj
