// hj_fp64.hip -- the plain FP64 distance-count kernels (the drop-in path).
//
// Replace the reference's 8x8 AVX-512 tile (EGO_LOOP_TRAN egojoin.h:311-392,
// EGO_LOOP_TRAN_NOSELF :482-553) and its count epilogue (egojoin.cpp:800-815,
// :868-884).  Differences by design:
//   * one wavefront owns a 64 x 64 tile pair: lane = column point, the row
//     points stream through scalar registers (uniform loads), so no LDS and no
//     register transposes are needed;
//   * the count kernel accumulates, as the reference does, the expanded form
//     |a|^2 + |b|^2 - 2 a.b with one FMA per pair and dimension and
//     precomputed prefix norms (the reference's block norms, egojoin.h:182-195),
//     but against thresholds that carry a rigorous FP64 rounding bound
//     (fp64_thresholds): pairs outside the thin band around eps^2 are decided
//     by it, and pairs inside it by the fixed-order difference form
//     s = fma(a_k - b_k, a_k - b_k, s) of the oracle, so the count equals the
//     FP64 oracle bit for bit (the reference's own expanded form can flip pairs
//     within rounding of eps).  The KBLOCK early exit (egojoin.h:349-359)
//     stays exact: a partial sum above the band rules every longer sum out.
//     Data far from the origin relative to eps (a band that would not be thin)
//     and the pair-list kernel use the difference form throughout;
//   * the early exit is a wavefront vote over R rows x 64 columns;
//   * counts are ballot + popcount per row, one 64-bit atomic per wave at the
//     end (the reference's sign-bit sum + OpenMP reduction);
//   * the diagonal rule (egojoin.h:383-392) is the predicate i < j on sorted
//     indices inside diagonal tiles;
//   * the list is drained in per-XCD segments (the expanded kernel) or
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

// Expanded form.  Evaluates rows [i0, i0+R) of tile I against the 64 column
// points of the calling lanes with one FMA per pair and dimension: dot_r =
// sum_k a_rk b_k.  The layout holds, after every 4 dimensions, the points'
// halved negative prefix norms h = -P/2 (for rows through scalar loads,
// [block][64] per tile, so the R rows of one block are one contiguous scalar
// load).  At the end of every KB dimensions v_r = dot_r + hA_r is compared
// with w = -hB - x/2:  s_r = Pa_r + Pb - 2 dot_r <= x  <=>  v_r >= w.  A group
// whose every pair has s > x_out exits (partial distances never decrease).
// Pads carry h = -inf (v_r = -inf, or w = +inf), so they never count; DIAG (a
// self-join's diagonal tile) also masks i >= j.  The first block's column
// values come preloaded once per tile pair (every group evaluates them);
// loading the second block ahead of the first vote, per group or per tile
// pair, measured 10-15% slower (the registers cost occupancy, and the kernel
// hides latency by its resident wavefronts).  Returns the dimensions
// evaluated; alive = some pair may be within eps, and then v[r] and *hb hold
// the full-d v_r and hB.
template <int KB, int R, bool DIAG>
__device__ __forceinline__ int row_group_x(const_f64 *__restrict__ ar, const_f64 *__restrict__ ha,
                                           int dpad, const double *__restrict__ cb,
                                           const double *__restrict__ hbp, uint32_t lane,
                                           const double (&b0)[KB], double hb0, uint32_t i0,
                                           uint32_t j, double xh_out, double (&v)[R], double *hb,
                                           bool *alive) {
    double dot[R];
#pragma unroll
    for (int r = 0; r < R; r++) dot[r] = 0.0;
    int c = 0;
    while (c < dpad) {
        double b[KB];
        double hbc;
        const uint32_t nb = (uint32_t)((c + KB) / 4 - 1);  // prefix norms after c + KB dims
        if (c == 0) {
#pragma unroll
            for (int k = 0; k < KB; k++) b[k] = b0[k];
            hbc = hb0;
        } else {
#pragma unroll
            for (int k = 0; k < KB; k++) b[k] = cb[(uint32_t)(c + k) * 64 + lane];
            hbc = hbp[nb * 64 + lane];
        }
        // two dimensions at a time: the empty asm ties the first pair's FMAs
        // ahead of the second pair's row loads, so the R x KB row values are
        // not all held in scalar registers at once (with them all live the
        // kernel spilled scalar registers to vector lanes; measured 9.8
        // against 10.2 ms at n = 1M, d = 64, eps 0.2)
#pragma unroll
        for (int h = 0; h < KB; h += 2) {
#pragma unroll
            for (int r = 0; r < R; r++) {
#pragma unroll
                for (int k = h; k < h + 2; k++)
                    dot[r] = __builtin_fma(ar[(uint32_t)(c + k) * 64 + r], b[k], dot[r]);
            }
            if (h + 2 < KB) {
#pragma unroll
                for (int r = 0; r < R; r++) asm volatile("" : "+v"(dot[r]));
            }
        }
        c += KB;
        double mx = -INFINITY;
#pragma unroll
        for (int r = 0; r < R; r++) {
            double vr = dot[r] + ha[nb * 64 + r];
            if (DIAG) vr = i0 + (uint32_t)r < j ? vr : -INFINITY;
            v[r] = vr;
            mx = fmax(mx, vr);
        }
        *hb = hbc;
        if (!__any(mx >= -hbc - xh_out)) {
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
    WorkCursor<4> cur(npairs);
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

// Expanded-form count kernel: the traversal of fp64_count_kernel with
// row_group_x.  Pairs whose final s lies within the rounding band (x_in,
// x_out] are decided by the fixed-order predicate itself, from the same
// layout.  Work is cut into per-XCD segments of the list: workgroup b runs on
// XCD b % 8 and takes, in dispatch order, consecutive slices of that XCD's
// segment, so tile pairs that share column tiles meet in one L2.
template <bool SELF, int KB, int R>
__global__ __launch_bounds__(256) void fp64x_count_kernel(
    const double *__restrict__ rows, const double *__restrict__ cols,
    const double *__restrict__ rnorm, const double *__restrict__ cnorm,
    const uint64_t *__restrict__ pairs, uint64_t npairs, const SegDev *__restrict__ segs,
    const uint32_t *__restrict__ tile_seg, int dpad, uint64_t *__restrict__ ctr,
    uint64_t *__restrict__ seg_counts) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // this wavefront's slice [w, w_end) of the list
    // (a grid smaller than the XCD count keeps one segment, so every pair has
    // a workgroup)
    const uint32_t nx = gridDim.x < XCD_COUNT ? 1u : XCD_COUNT;
    const uint32_t xcd = blockIdx.x % nx, kb = blockIdx.x / nx;
    const uint32_t blocks_here = (gridDim.x - xcd + nx - 1) / nx;
    const uint64_t seg_lo = npairs * xcd / nx, seg_hi = npairs * (xcd + 1) / nx;
    const uint64_t slices = (uint64_t)blocks_here * 4;
    const uint64_t sl = (uint64_t)kb * 4 + wave;
    uint64_t w = seg_lo + (seg_hi - seg_lo) * sl / slices;
    const uint64_t w_end = seg_lo + (seg_hi - seg_lo) * (sl + 1) / slices;
    const int nblk = dpad / 4;
    uint64_t cnt = 0, work = 0, cand = 0;
    SegCount sc;
    uint32_t curI = 0xFFFFFFFFu, nvi = 0, ttile0 = 0;
    uint64_t tn = 0;
    double eps2 = 0.0, xh_in = 0.0, xh_out = 0.0;
    for (; w < w_end; w++) {
        const uint64_t pr = pairs[w];
        const uint32_t I = (uint32_t)pr, J = (uint32_t)(pr >> 32);
        if (I != curI) {
            const uint32_t sg = tile_seg[I];
            const SegDev &S = segs[sg];
            sc.set(sg, seg_counts);
            eps2 = S.eps2;
            xh_in = 0.5 * S.x_in;
            xh_out = 0.5 * S.x_out;
            nvi = tile_points(S.qn - 64ull * (I - S.qtile0));
            tn = S.tn;
            ttile0 = S.ttile0;
            curI = I;
        }
        const uint32_t nvj = tile_points(tn - 64ull * (J - ttile0));
        const uint32_t j = J * 64 + lane;  // layout column
        const uint32_t i_beg = I * 64;
        const uint32_t i_end = i_beg + nvi;
        const bool diag = SELF && I == J;
        cand += diag ? (uint64_t)nvj * (nvj - 1) / 2 : (uint64_t)nvi * nvj;
        const double *cb = cols + (size_t)J * dpad * 64;  // wave-uniform tile base
        const double *pb = cnorm + (size_t)J * nblk * 64;
        double b0[KB];
        const double hb0 = pb[(uint32_t)(KB / 4 - 1) * 64 + lane];
#pragma unroll
        for (int k = 0; k < KB; k++) b0[k] = cb[(uint32_t)k * 64 + lane];
        for (uint32_t i0 = i_beg; i0 < i_end; i0 += R) {
            if (diag && i0 >= i_beg + nvj - 1) break;  // no column j > i left
            const_f64 *ar = (const_f64 *)(rows + (size_t)I * dpad * 64 + (i0 - i_beg));
            const_f64 *ha = (const_f64 *)(rnorm + (size_t)I * nblk * 64 + (i0 - i_beg));
            double v[R], hb;
            bool alive;
            const int dims =
                diag ? row_group_x<KB, R, true>(ar, ha, dpad, cb, pb, lane, b0, hb0, i0, j,
                                                xh_out, v, &hb, &alive)
                     : row_group_x<KB, R, false>(ar, ha, dpad, cb, pb, lane, b0, hb0, i0, j,
                                                 xh_out, v, &hb, &alive);
            work += (uint64_t)dims * (R * 64);
            if (!alive) continue;
            const double w_in = -hb - xh_in, w_out = -hb - xh_out;
            uint64_t c = 0;
#pragma unroll
            for (int r = 0; r < R; r++) {
                c += __popcll(__ballot(v[r] >= w_in));
                const bool amb = v[r] >= w_out && v[r] < w_in;
                if (__any(amb)) {
                    // within rounding of eps: the fixed-order predicate
                    double s = 0.0;
                    for (int k = 0; k < dpad; k++) {
                        const double t = ar[(uint32_t)k * 64 + r] - cb[(uint32_t)k * 64 + lane];
                        s = __builtin_fma(t, t, s);
                    }
                    c += __popcll(__ballot(amb && s <= eps2));
                }
            }
            cnt += c;
            sc.cnt += c;
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
    WorkCursor<4> cur(npairs);
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

// Workgroups of the expanded kernel: about FP64X_PAIRS_PER_WAVE tile pairs per
// wavefront, at least the resident count.
constexpr uint64_t FP64X_PAIRS_PER_WAVE = 16;

template <bool SELF, int KB, int R>
void fp64x_count_launch(const JoinArgs &ja, const double *rows, const double *cols,
                        const double *rnorm, const double *cnorm, int dpad, hipStream_t s) {
    auto k = fp64x_count_kernel<SELF, KB, R>;
    int g = ja.grid;
    if (g <= 0) {
        const int resident = persistent_grid(k, 256, ja.num_cus, 0);
        const uint64_t want = (ja.npairs + 4 * FP64X_PAIRS_PER_WAVE - 1) / (4 * FP64X_PAIRS_PER_WAVE);
        g = (int)std::min<uint64_t>(std::max<uint64_t>(want, (uint64_t)resident), 1u << 30);
    }
    k<<<g, 256, 0, s>>>(rows, cols, rnorm, cnorm, ja.pairs, ja.npairs, ja.segs, ja.tile_seg, dpad,
                        ja.ctr, ja.seg_counts);
}

template <bool SELF, int KB>
void fp64_count_rows(const JoinArgs &ja, const double *rows, const double *cols,
                     const double *rnorm, const double *cnorm, int dpad, bool expanded,
                     hipStream_t s) {
    if (expanded) {
        if (ja.tile_rows == 8)
            fp64x_count_launch<SELF, KB, 8>(ja, rows, cols, rnorm, cnorm, dpad, s);
        else
            fp64x_count_launch<SELF, KB, 16>(ja, rows, cols, rnorm, cnorm, dpad, s);
    } else {
        if (ja.tile_rows == 8)
            fp64_count_launch<SELF, KB, 8>(ja, rows, cols, dpad, s);
        else
            fp64_count_launch<SELF, KB, 16>(ja, rows, cols, dpad, s);
    }
}

void launch_fp64_count(const JoinArgs &ja, const double *rows, const double *cols,
                       const double *rnorm, const double *cnorm, int dpad, bool self, int kblock,
                       bool expanded, hipStream_t s) {
    if (ja.npairs == 0) return;
    if (!work_list_fits(ja.npairs, 4)) throw InvalidArg{"work list too long"};
    if (dpad % kblock != 0) throw InvalidArg{"internal: tile rows not whole early-exit blocks"};
    if (expanded && (!rnorm || !cnorm)) throw InvalidArg{"internal: prefix norms missing"};
    if (kblock == 4) {
        if (self)
            fp64_count_rows<true, 4>(ja, rows, cols, rnorm, cnorm, dpad, expanded, s);
        else
            fp64_count_rows<false, 4>(ja, rows, cols, rnorm, cnorm, dpad, expanded, s);
    } else if (kblock == 8) {
        if (self)
            fp64_count_rows<true, 8>(ja, rows, cols, rnorm, cnorm, dpad, expanded, s);
        else
            fp64_count_rows<false, 8>(ja, rows, cols, rnorm, cnorm, dpad, expanded, s);
    } else {
        // KBLOCK = 16, the setting recommended for uniform data (FAQ.md:14-22);
        // the caller lays tiles out with dpad a multiple of 16
        if (self)
            fp64_count_rows<true, 16>(ja, rows, cols, rnorm, cnorm, dpad, expanded, s);
        else
            fp64_count_rows<false, 16>(ja, rows, cols, rnorm, cnorm, dpad, expanded, s);
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
