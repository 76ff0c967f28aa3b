// hj_prep.hip -- grid keys, gathers, tile bounds and the tile-pair work list.
//
// Reference correspondence:
//   launch_minmax / launch_keys   grid cells floor(x * (1/eps)) of
//                                 epsilonGridCompare (egojoin.cpp:9-17), packed
//                                 into one 64-bit key per point (dimension 0 in
//                                 the most significant field) so that unsigned
//                                 key order is the lexicographic cell order.
//   launch_gather                 the sorted copy the reference obtains by
//                                 sorting rows in place, plus the transposed
//                                 tile layout the reference builds with
//                                 transpose_8xd (egojoin.cpp:234-267), written
//                                 in one pass.
//   launch_tile_index /           stripe bounds: prepareStripes
//   launch_tile_ranges            (egojoin.cpp:270-464) and prepareStripesNoSelf
//                                 (:1126-1240) find, per point and stripe, the
//                                 sorted range of candidate partners with
//                                 galloping searches, then take block min/max
//                                 (:437-447).  Here a whole 64-point tile is
//                                 bounded at once (see launch_tile_ranges),
//                                 with two-level lockstep searches (a per-tile
//                                 first-key table, then 64 keys).
//   launch_choose_active          the active-dimension count A of
//                                 prepareStripes (2 S <= 3^A) becomes the A
//                                 that minimizes a work-list cost estimated on
//                                 sampled query tiles.
//   launch_emit_pairs /           the Hilbert traversal FGF_HILBERT_FOR
//   launch_hilbert_keys           (hilloop.h:9-82) with its cond/cond2 pruning
//                                 (egojoin.h:325-328) becomes an explicit list
//                                 of candidate tile pairs, ordered by the
//                                 textbook Hilbert index of (I, J).
#include <algorithm>
#include <vector>

#include "hj_device.h"
#include "hj_internal.h"

namespace hj {

namespace {

constexpr int MM_ROWS = 4;
constexpr int MM_UNROLL = 8;
constexpr uint64_t MM_CHUNK_ROWS = 1024;
constexpr uint64_t PC_SLICE = 4096;  // points per block of prefix_classes_kernel

// The rows of chunk c: from the chunk table, or with no table (one segment)
// rows [row0 + c * rows_per, row0 + (c + 1) * rows_per) clipped to row_end;
// of those every step-th from the first (a sample; step 1: all of them).
struct MinmaxChunks {
    const uint64_t *begin, *end;  // [nchunks] each, or null
    const uint32_t *step;         // [nchunks] with the table
    uint64_t row0, row_end, rows_per;
    uint32_t step1;               // without the table
    uint32_t nchunks;             // chunks of the table (stride of the flags)
};

// Per-segment min / max of dimensions [k0, k0 + nd) of the rows, stage 1: one
// block per chunk of at most MM_CHUNK_ROWS rows of one segment (and per 64
// dimensions of the range).  A wavefront covers 64 / lpr rows at a time, lane
// = (row, dimension) with lpr lanes per row (the power of two >= nd, at least
// 8, at most 64), so a narrow range still keeps every lane loading; each
// thread keeps MM_UNROLL independent loads in flight.  Every block writes its
// ranges and whether it met a non-finite coordinate (cbad[y][chunk]), so
// nothing needs clearing first.
__global__ __launch_bounds__(256) void minmax_partial(const double *__restrict__ pts, int d,
                                                      int k0, int nd, int lpr, MinmaxChunks ch,
                                                      double *__restrict__ pmin,
                                                      double *__restrict__ pmax,
                                                      uint32_t *__restrict__ cbad) {
    __shared__ double smin[MM_ROWS][64], smax[MM_ROWS][64];
    const int lane = threadIdx.x;
    const int j = blockIdx.y * 64 + (lane & (lpr - 1));  // dimension k0 + j of the range
    const int rsub = lane / lpr, rpw = 64 / lpr;        // row within the wavefront's group
    uint64_t r_begin, r_end, step;
    if (ch.begin) {
        r_begin = ch.begin[blockIdx.x];
        r_end = ch.end[blockIdx.x];
        step = ch.step[blockIdx.x];
    } else {
        r_begin = ch.row0 + (uint64_t)blockIdx.x * ch.rows_per;
        r_end = min(r_begin + ch.rows_per, ch.row_end);
        step = ch.step1;
    }
    const uint64_t nvis = (r_end - r_begin + step - 1) / step;  // rows visited
    const uint64_t rstep = (uint64_t)MM_ROWS * rpw;
    double lo = INFINITY, hi = -INFINITY;
    int nonfinite = 0;
    if (j < nd) {
        const double *col = pts + k0 + j;
        for (uint64_t i0 = threadIdx.y * rpw + rsub; i0 < nvis; i0 += rstep * MM_UNROLL) {
            double v[MM_UNROLL];
#pragma unroll
            for (int u = 0; u < MM_UNROLL; u++) {
                const uint64_t i = i0 + u * rstep;
                v[u] = i < nvis ? col[(r_begin + i * step) * d] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < MM_UNROLL; u++) {
                if (i0 + u * rstep < nvis) {
                    if (!isfinite(v[u])) nonfinite = 1;
                    lo = fmin(lo, v[u]);
                    hi = fmax(hi, v[u]);
                }
            }
        }
    }
    smin[threadIdx.y][lane] = lo;
    smax[threadIdx.y][lane] = hi;
    const int any_bad = __syncthreads_or(nonfinite);
    if (threadIdx.y == 0 && lane == 0)
        cbad[(size_t)blockIdx.y * ch.nchunks + blockIdx.x] = any_bad ? 1u : 0u;
    // lanes 0 .. lpr-1 of wavefront 0 combine the rows and wavefronts of their dimension
    if (threadIdx.y == 0 && lane < lpr && j < nd) {
        for (int y = 0; y < MM_ROWS; y++)
            for (int q = 0; q < rpw; q++) {
                lo = fmin(lo, smin[y][q * lpr + lane]);
                hi = fmax(hi, smax[y][q * lpr + lane]);
            }
        pmin[(size_t)blockIdx.x * nd + j] = lo;
        pmax[(size_t)blockIdx.x * nd + j] = hi;
    }
}

// Stage 2: one block per (dimension, segment) over the segment's chunks
// (seg_chunk0[seg] .. seg_chunk0[seg + 1], or with no table chunks 0 ..
// nchunks_one of the one segment).  The dimension-0 block of a segment also
// reports whether any of its chunks met a non-finite coordinate (segbad), and
// block (0, 0) clears clear[0 .. 3] (flags of the kernels that follow).
__global__ __launch_bounds__(256) void minmax_final(const double *__restrict__ pmin,
                                                    const double *__restrict__ pmax,
                                                    const uint32_t *__restrict__ cbad,
                                                    uint32_t nchunks, uint32_t ny,
                                                    const uint32_t *__restrict__ seg_chunk0,
                                                    uint32_t nchunks_one, int nd,
                                                    double *__restrict__ mins,
                                                    double *__restrict__ maxs,
                                                    uint32_t *__restrict__ segbad,
                                                    uint32_t *__restrict__ clear) {
    __shared__ double slo[256], shi[256];
    const int dim = blockIdx.x;
    const uint32_t seg = blockIdx.y;
    const uint32_t c0 = seg_chunk0 ? seg_chunk0[seg] : 0u;
    const uint32_t c1 = seg_chunk0 ? seg_chunk0[seg + 1] : nchunks_one;
    double lo = INFINITY, hi = -INFINITY;
    for (uint32_t c = c0 + threadIdx.x; c < c1; c += 256) {
        lo = fmin(lo, pmin[(size_t)c * nd + dim]);
        hi = fmax(hi, pmax[(size_t)c * nd + dim]);
    }
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + w]);
            shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        mins[(size_t)seg * nd + dim] = slo[0];
        maxs[(size_t)seg * nd + dim] = shi[0];
    }
    if (dim == 0) {  // uniform over the block
        int bad = 0;
        for (uint32_t y = 0; y < ny; y++)
            for (uint32_t c = c0 + threadIdx.x; c < c1; c += 256)
                bad |= cbad[(size_t)y * nchunks + c] != 0u;
        bad = __syncthreads_or(bad);
        if (threadIdx.x == 0) {
            segbad[seg] = bad ? 1u : 0u;
            if (clear && seg == 0) clear[0] = clear[1] = clear[2] = clear[3] = 0u;
        }
    }
}

// One wavefront per 64 consecutive input rows: the leading kd_max coordinates
// of the 64 rows are staged in LDS with row-contiguous (coalesced) loads, then
// each lane packs its row's key: the prefix of its segment and the cells of
// the segment's grid.
__global__ __launch_bounds__(256) void keys_kernel(const double *__restrict__ pts, size_t n, int d,
                                                   const SegDev *__restrict__ segs, uint32_t nseg,
                                                   const uint64_t *__restrict__ seg_off,
                                                   const GridDev *__restrict__ grids,
                                                   const int64_t *__restrict__ minc, int kd_max,
                                                   int query_prefix, uint64_t *__restrict__ keys,
                                                   uint32_t *__restrict__ vals,
                                                   uint64_t *__restrict__ plo,
                                                   uint64_t *__restrict__ phi,
                                                   uint32_t *__restrict__ widened) {
    extern __shared__ double kx[];  // [4 waves][64][kd_max + 1]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kd = kd_max, stride = kd + 1;
    double *x = kx + (size_t)wave * 64 * stride;
    const size_t base = ((size_t)blockIdx.x * 4 + wave) * 64;
    for (int e = lane; e < 64 * kd; e += 64) {
        const int p = e / kd, k = e - p * kd;
        const size_t i = base + p;
        x[p * stride + k] = (i < n && k < d) ? pts[i * d + k] : 0.0;
    }
    __syncthreads();
    const size_t i = base + lane;
    bool wide = false;  // the point's key bounds were widened
    if (i < n) {
        const uint32_t sg = find_seg(seg_off, nseg, i);
        const SegDev &S = segs[sg];
        const GridDev g = grids[S.grid];
        keys[i] = (query_prefix ? S.qprefix : S.tprefix) |
                  point_key(x + lane * stride, 1, g, minc + g.minc_off);
        vals[i] = (uint32_t)i;
        if (plo) {
            // the point's widened key bounds (point_key_bounds), for the tile
            // bounds after the sort
            uint64_t lo, hi;
            point_key_bounds(x + lane * stride, 1, g, minc + g.minc_off, &lo, &hi);
            plo[i] = lo;
            phi[i] = hi;
            wide = lo != hi;
        }
    }
    // one flag for the whole set: no point near a cell boundary lets the tile
    // bounds come from the sorted keys alone (tile_index_kernel)
    // (read first: on data clipped to a cell boundary most wavefronts have
    // such a point, and one atomic each on one word would serialize)
    if (widened && __ballot(wide) != 0 && lane == 0 && *(volatile uint32_t *)widened == 0u)
        atomicOr(widened, 1u);
}

// One block per layout tile: rows[i][k] (row-major, padded to dpad
// dimensions) and cols[t][k][p] (dimension-major inside the tile).
__global__ __launch_bounds__(256) void gather_kernel(const double *__restrict__ pts,
                                                     const uint32_t *__restrict__ perm,
                                                     const SegDev *__restrict__ segs,
                                                     const uint32_t *__restrict__ tile_seg,
                                                     int target, int d, int dpad,
                                                     double *__restrict__ cols,
                                                     double *__restrict__ pnorm) {
    __shared__ double lds[64][65];
    const size_t T = blockIdx.x;
    const int t = threadIdx.x;
    const SegDev &S = segs[tile_seg[T]];
    uint64_t start;
    uint32_t nv;
    tile_span(S, (uint32_t)T, target != 0, &start, &nv);
    const int nblk = dpad / 4;
    __shared__ double bsum[16][64];  // squared norm of each 4-dimension block of a chunk
    __shared__ double carry[64];     // squared norm of the chunks before
    if (t < 64) carry[t] = 0.0;
    for (int c0 = 0; c0 < dpad; c0 += 64) {
        // all 16 loads of a thread in flight before the first LDS store
        double v[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int e = t + 256 * i, p = e >> 6, k = c0 + (e & 63);
            if ((uint32_t)p < nv)
                v[i] = k < d ? pts[(size_t)perm[start + p] * d + k] : 0.0;
            else
                v[i] = k < d ? -PAD : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int e = t + 256 * i, p = e >> 6, kk = e & 63;
            lds[p][kk] = v[i];
        }
        __syncthreads();
        for (int e = t; e < 64 * 64; e += 256) {
            int kk = e >> 6, p = e & 63, k = c0 + kk;
            if (k < dpad) cols[(T * dpad + k) * 64 + p] = lds[p][kk];
        }
        // halved negative prefix norms -P/2 after every 4 dimensions: wave w
        // sums the squares of blocks 4w .. 4w+3 of the chunk for lane = point
        // p, then adds the blocks before (any summation order of the squares
        // stays within the rounding bound the FP64 kernel assumes)
        const int w = t >> 6, p = t & 63;
        const int nbc = (dpad - c0 < 64 ? dpad - c0 : 64) / 4;  // blocks in this chunk
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int b = 4 * w + j;
            double q = 0.0;
            if (b < nbc) {
#pragma unroll
                for (int k = 0; k < 4; k++) q = __builtin_fma(lds[p][4 * b + k], lds[p][4 * b + k], q);
            }
            bsum[b][p] = q;
        }
        __syncthreads();
        double run = carry[p];
        for (int b = 0; b < 4 * w; b++) run += bsum[b][p];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int b = 4 * w + j;
            run += bsum[b][p];
            if (b < nbc)
                pnorm[(T * nblk + c0 / 4 + b) * 64 + p] = (uint32_t)p < nv ? -0.5 * run : -INFINITY;
        }
        __syncthreads();
        if (w == 3) carry[p] = run;  // all 16 blocks (empty ones add 0)
        __syncthreads();
    }
}

// Per layout tile: tlo = prefix | min over its points of (key - Wlo), thi =
// prefix | max of (key + Whi) (point_key_bounds), from the per-point bounds the
// keys kernel wrote (indexed by input row, read through the EGO permutation).
// One wavefront per layout tile: the tile's first sorted key (the level-1
// table of lockstep_bounds), and with tlo / thi its widened key bounds: the
// minimum of its points' lower and the maximum of their upper widened key
// bounds (read through the permutation), which are the tile's first and last
// sorted keys when no point of the set had its bounds widened (*widened ==
// 0, keys_kernel); `target` selects the target or query span of the tile's
// segment.
__global__ __launch_bounds__(256) void tile_index_kernel(const uint64_t *__restrict__ keys,
                                                         const uint32_t *__restrict__ perm,
                                                         const uint64_t *__restrict__ plo,
                                                         const uint64_t *__restrict__ phi,
                                                         const SegDev *__restrict__ segs,
                                                         const uint32_t *__restrict__ tile_seg,
                                                         size_t ntiles, int target,
                                                         const uint32_t *__restrict__ widened,
                                                         uint64_t *__restrict__ tfirst,
                                                         uint64_t *__restrict__ tlo,
                                                         uint64_t *__restrict__ thi) {
    const size_t T = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (T >= ntiles) return;
    const SegDev &S = segs[tile_seg[T]];
    uint64_t start;
    uint32_t nv;
    tile_span(S, (uint32_t)T, target != 0, &start, &nv);
    const uint64_t first = keys[start];
    if (lane == 0) tfirst[T] = first;
    if (!tlo) return;
    if (!widened || *widened == 0u) {
        if (lane == 0) {
            tlo[T] = first;
            thi[T] = keys[start + nv - 1];
        }
        return;
    }
    uint64_t lo = ~0ull, hi = 0ull;
    if ((uint32_t)lane < nv) {
        const uint32_t i = perm[start + lane];
        lo = plo[i];
        hi = phi[i];
    }
    wave_min_max_u64(lo, hi);
    if (lane == 0) {
        const uint64_t prefix = target ? S.tprefix : S.qprefix;
        tlo[T] = prefix | lo;
        thi[T] = prefix | hi;
    }
}

constexpr int MAX_CODES = 729;  // 3^6
constexpr int CODE_SLOTS = 4;   // stripe codes resolved per lane at a time

__device__ __forceinline__ uint64_t min64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// The coarse level of the stripe searches: every cs-th first key of the
// target tiles (cs a power of two, at most COARSE_MAX entries), staged in LDS
// by the whole workgroup.
constexpr uint32_t COARSE_MAX = 256;

struct Coarse {
    const uint64_t *keys;  // LDS
    uint32_t n;            // entries
    uint32_t cs;           // tiles per entry
};

__device__ __forceinline__ Coarse stage_coarse(const uint64_t *__restrict__ tf, uint32_t ntt,
                                               uint64_t *s_coarse, int nthreads) {
    uint32_t cs = 1;
    while ((uint64_t)cs * COARSE_MAX < ntt) cs <<= 1;
    const uint32_t nc = (ntt + cs - 1) / cs;
    for (uint32_t i = threadIdx.x; i < nc; i += (uint32_t)nthreads) s_coarse[i] = tf[(size_t)i * cs];
    __syncthreads();
    return Coarse{s_coarse, nc, cs};
}

// Sorted positions of N keys x[j] in keys[0, n): b[j] = count of keys < x[j]
// (even j) or <= x[j] (odd j).  Three levels: the coarse first keys in LDS,
// the first keys tf of the tiles between two coarse entries, then the 64 keys
// of one tile.  Every level is branch-free with the same number of steps for
// every search, so each step issues all N loads before it waits.  Level-1
// invariant (after the coarse level): the count of search j in tf lies in
// [t[j], t[j] + cs]; in level 2 a position past the end reads the last key,
// and the result is clamped to n.
template <int N>
__device__ __forceinline__ void lockstep_bounds(const uint64_t *__restrict__ tf, uint32_t ntt,
                                                const Coarse &co,
                                                const uint64_t *__restrict__ keys, size_t n,
                                                const uint64_t *x, size_t *b) {
#pragma unroll
    for (int j = 0; j < N; j++) b[j] = 0;
    if (n == 0) return;
    uint32_t t[N];
    // coarse level (LDS): c = count of coarse keys before x; the tiles' count
    // then lies in [(c - 1) cs + 1, c cs] (c >= 1), or is 0 (c == 0)
    uint32_t c[N];
#pragma unroll
    for (int j = 0; j < N; j++) c[j] = 0;
    {
        uint32_t len = co.n;
        while (len > 1) {
            const uint32_t half = len >> 1;
#pragma unroll
            for (int j = 0; j < N; j++) {
                const uint64_t v = co.keys[c[j] + half];
                c[j] += ((j & 1) ? v <= x[j] : v < x[j]) ? half : 0;
            }
            len -= half;
        }
#pragma unroll
        for (int j = 0; j < N; j++) {
            const uint64_t v = co.keys[c[j]];
            c[j] += ((j & 1) ? v <= x[j] : v < x[j]) ? 1 : 0;
        }
    }
    // tiles between coarse entries: t = the last tile whose first key is
    // before x (tf[t] is, by the coarse level), then the count
#pragma unroll
    for (int j = 0; j < N; j++) t[j] = c[j] ? (c[j] - 1) * co.cs : 0;
    uint32_t len = co.cs;
    while (len > 1) {
        const uint32_t half = len >> 1;
        uint64_t v[N];
#pragma unroll
        for (int j = 0; j < N; j++) {
            const uint32_t q = t[j] + half;
            v[j] = q < ntt ? tf[q] : ~0ull;
        }
#pragma unroll
        for (int j = 0; j < N; j++) {
            const uint32_t q = t[j] + half;
            t[j] += (q < ntt && ((j & 1) ? v[j] <= x[j] : v[j] < x[j])) ? half : 0;
        }
        len -= half;
    }
    {
        uint64_t v[N];
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = tf[t[j]];
#pragma unroll
        for (int j = 0; j < N; j++) {
            t[j] += (c[j] && ((j & 1) ? v[j] <= x[j] : v[j] < x[j])) ? 1 : 0;
            b[j] = t[j] ? (size_t)64 * (t[j] - 1) : 0;
        }
    }
    const size_t last = n - 1;
    uint32_t len2 = 64;
    while (len2 > 1) {
        const uint32_t half = len2 >> 1;
        uint64_t v[N];
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = keys[min64(b[j] + half, last)];
#pragma unroll
        for (int j = 0; j < N; j++) b[j] += ((j & 1) ? v[j] <= x[j] : v[j] < x[j]) ? half : 0;
        len2 -= half;
    }
    uint64_t v[N];
#pragma unroll
    for (int j = 0; j < N; j++) v[j] = keys[min64(b[j], last)];
#pragma unroll
    for (int j = 0; j < N; j++) {
        b[j] += ((j & 1) ? v[j] <= x[j] : v[j] < x[j]) ? 1 : 0;
        b[j] = min64(b[j], n);
    }
}

// Stripe intervals of one query tile (one wavefront).  Code c (base 3, most
// significant digit = dimension 0) encodes the offsets delta_k = digit_k - 1
// of the first `active` key fields.  For code c the candidates are the target
// points of the tile's segment with local keys in
//   [lo + D_lo(c), hi + D_hi(c)],
//   D_lo(c) = sum_k<A delta_k 2^s_k - sum_A<=k<D' 2^s_k,
//   D_hi(c) = sum_k<A delta_k 2^s_k + sum_A<=k<D' 2^s_k   (s_k = bits (D'-1-k)),
// (lo, hi: the tile's widened local key bounds) -- the stripes of the
// reference with exact cell arithmetic instead of displaced coordinates.  A
// self segment needs only codes >=lex 0 (partners later in the EGO order) and
// tiles J >= I.  Bounds grow with c, so the tile intervals arrive sorted.
// The NT threads of the block write the target-tile interval of every code
// to s_start / s_end; returns the number of codes.
template <int NT>
__device__ __forceinline__ int stripe_intervals(uint32_t I, const SegDev &S, const GridDev &g,
                                                uint32_t active, uint64_t lo0, uint64_t hi0,
                                                const uint64_t *__restrict__ tkeys,
                                                const uint64_t *__restrict__ tfirst,
                                                const Coarse &co, uint32_t *s_start,
                                                uint32_t *s_end) {
    const int lane = threadIdx.x;
    int64_t rest = 0;
    for (uint32_t k = active; k < g.key_dims; k++)
        rest += (int64_t)1 << (g.bits * (g.key_dims - 1 - k));
    int total = 1;
    for (uint32_t k = 0; k < active; k++) total *= 3;
    const int code_begin = S.self ? (total - 1) / 2 : 0;
    const int ncodes = total - code_begin;
    const uint32_t tiles_t = (uint32_t)((S.tn + 63) / 64);
    // Each lane resolves up to CODE_SLOTS codes at a time (lockstep_bounds).
    const uint64_t *keys = tkeys + S.toff;
    const uint64_t *tf = tfirst + S.ttile0;
    const size_t nt = S.tn;
    const int slots = min(CODE_SLOTS, (ncodes + NT - 1) / NT);
    for (int c0 = 0; c0 < ncodes; c0 += NT * slots) {
        uint64_t x[2 * CODE_SLOTS];
        size_t b[2 * CODE_SLOTS];
#pragma unroll
        for (int j = 0; j < CODE_SLOTS; j++) {
            x[2 * j] = x[2 * j + 1] = 0;
            const int ci = c0 + NT * j + lane;
            if (j < slots && ci < ncodes) {
                int cc = code_begin + ci;
                int64_t dsum = 0;
                for (int k = (int)active - 1; k >= 0; k--) {
                    const int digit = cc % 3;
                    cc /= 3;
                    dsum += (int64_t)(digit - 1) * ((int64_t)1 << (g.bits * (g.key_dims - 1 - k)));
                }
                x[2 * j] = S.tprefix | (lo0 + (uint64_t)(dsum - rest));
                x[2 * j + 1] = S.tprefix | (hi0 + (uint64_t)(dsum + rest));
            }
        }
        if (slots == 1)
            lockstep_bounds<2>(tf, tiles_t, co, keys, nt, x, b);
        else if (slots == 2)
            lockstep_bounds<4>(tf, tiles_t, co, keys, nt, x, b);
        else
            lockstep_bounds<2 * CODE_SLOTS>(tf, tiles_t, co, keys, nt, x, b);
#pragma unroll
        for (int j = 0; j < CODE_SLOTS; j++) {
            const int ci = c0 + NT * j + lane;
            if (j < slots && ci < ncodes) {
                const size_t st = b[2 * j], en = b[2 * j + 1];
                uint32_t ts = (uint32_t)(st / 64), te = (uint32_t)((en + 63) / 64);
                if (st >= en) ts = te = 0;
                if (te > tiles_t) te = tiles_t;
                ts += S.ttile0;
                te += S.ttile0;
                if (S.self && ts < I) ts = I;
                s_start[ci] = ts;
                s_end[ci] = te;
            }
        }
    }
    return ncodes;
}

struct RunsSummary {
    uint32_t runs;      // number of merged runs
    uint64_t tiles;     // target tiles in the runs
    uint64_t cand;      // candidate point pairs of the query tile
};

// Merges the intervals of stripe_intervals into runs, 64 codes at a time, and
// writes them to out (when not null).  Nonempty intervals have
// non-decreasing starts and ends (the bounds grow with the code), so an
// interval opens a new run exactly when its start lies beyond the end of the
// previous nonempty interval, and that end closes the previous run.  The
// result is valid in every lane.
__device__ __forceinline__ RunsSummary merge_intervals(uint32_t I, const SegDev &S, int ncodes,
                                                       const uint32_t *s_start,
                                                       const uint32_t *s_end, uint32_t *out) {
    const int lane = threadIdx.x & 63;
    const uint32_t tiles_t = (uint32_t)((S.tn + 63) / 64);
    const uint32_t last = S.ttile0 + tiles_t - 1;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t m = 0, carry_end = 0;
    bool any = false, cover_last = false, cover_diag = false;
    uint64_t tot = 0;  // this lane's share of the sum of run lengths (mod 2^64)
    for (int c0 = 0; c0 < ncodes; c0 += 64) {
        const int ci = c0 + lane;
        const uint32_t s = ci < ncodes ? s_start[ci] : 0u;
        const uint32_t e = ci < ncodes ? s_end[ci] : 0u;
        const bool ne = ci < ncodes && s < e;
        const uint64_t nem = __ballot(ne);
        uint32_t pm = ne ? e : 0u;  // inclusive prefix maximum of the ends
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(pm, o);
            if (lane >= o) pm = u > pm ? u : pm;
        }
        uint32_t prev = __shfl_up(pm, 1);
        if (lane == 0) prev = 0;
        prev = prev > carry_end ? prev : carry_end;
        const bool opens = ne && ((!any && (nem & lt) == 0) || s > prev);
        const uint64_t om = __ballot(opens);
        if (opens) {
            const uint32_t r = m + (uint32_t)__popcll(om & lt);
            if (out) out[2 * r] = s;
            tot -= s;
            if (r > 0) {
                if (out) out[2 * (r - 1) + 1] = prev;
                tot += prev;
            }
        }
        cover_last = cover_last || __ballot(ne && s <= last && last < e) != 0;
        cover_diag = cover_diag || __ballot(ne && s <= I && I < e) != 0;
        m += (uint32_t)__popcll(om);
        any = any || nem != 0;
        const uint32_t cmax = __shfl(pm, 63);
        carry_end = cmax > carry_end ? cmax : carry_end;
    }
    for (int o = 32; o >= 1; o >>= 1) tot += shfl_xor_u64(tot, o);
    if (m > 0) {
        if (out && lane == 0) out[2 * (m - 1) + 1] = carry_end;
        tot += carry_end;
    }
    // candidate pairs: the query tile's real points times those of every
    // target tile in the runs; the diagonal tile of a self segment counts its
    // i < j pairs
    const uint64_t nvi = min64(S.qn - 64ull * (I - S.qtile0), 64);
    const uint64_t nv_last = tiles_t ? S.tn - 64ull * (tiles_t - 1) : 0;
    uint64_t cand = nvi * (64ull * tot - (cover_last ? 64 - nv_last : 0));
    if (S.self && cover_diag) cand -= nvi * nvi - nvi * (nvi - 1) / 2;
    return RunsSummary{m, tot, cand};
}

// One wavefront per query tile: its stripe intervals merged into runs of
// target tiles (ivals), their number (icount) and total length (jcount).
__global__ __launch_bounds__(64) void tile_ranges_kernel(
    const uint64_t *__restrict__ tlo, const uint64_t *__restrict__ thi,
    const uint32_t *__restrict__ tile_seg, size_t ntq, const uint64_t *__restrict__ tkeys,
    const uint64_t *__restrict__ tfirst, const SegDev *__restrict__ segs,
    const GridDev *__restrict__ grids, uint32_t *__restrict__ ivals,
    uint32_t *__restrict__ icount, uint64_t *__restrict__ jcount, int max_iv,
    uint64_t *__restrict__ cand_total) {
    __shared__ uint32_t s_start[MAX_CODES], s_end[MAX_CODES];
    __shared__ uint64_t s_coarse[COARSE_MAX];
    const uint32_t I = blockIdx.x;
    if (I >= ntq) return;
    const SegDev S = segs[tile_seg[I]];
    const GridDev g = grids[S.grid];
    const uint64_t lo0 = tlo[I] & S.local_mask, hi0 = thi[I] & S.local_mask;
    // (an unresolved ACTIVE_AUTO never reaches here; the clamp keeps the code
    // enumeration bounded regardless)
    const uint32_t active = g.active < 6u ? g.active : 6u;
    const Coarse co =
        stage_coarse(tfirst + S.ttile0, (uint32_t)((S.tn + 63) / 64), s_coarse, 64);
    const int ncodes =
        stripe_intervals<64>(I, S, g, active, lo0, hi0, tkeys, tfirst, co, s_start, s_end);
    __syncthreads();
    const RunsSummary R =
        merge_intervals(I, S, ncodes, s_start, s_end, ivals + (size_t)I * max_iv * 2);
    if (threadIdx.x == 0) {
        // spread over CAND_SLOTS counters to keep the atomics uncontended
        if (R.cand)
            atomicAdd((unsigned long long *)&cand_total[I % CAND_SLOTS],
                      (unsigned long long)R.cand);
        icount[I] = R.runs;
        jcount[I] = R.tiles;
    }
}

// A choice of a active dimensions whose stripe codes alone cost more than
// listing every target tile of the segment is never the cheapest: it is
// neither sampled nor taken.
__device__ __forceinline__ bool active_useful(const SegDev &S, uint32_t a, float code_cost) {
    uint32_t total = 1;
    for (uint32_t q = 0; q < a; q++) total *= 3;
    const uint32_t codes = S.self ? total - (total - 1) / 2 : total;
    return a == 0 || code_cost * (float)codes < (float)((S.tn + 63) / 64);
}

// Work-list size estimate for every choice of active dimensions, from
// `samples` query tiles spread evenly over each segment (block x = sample,
// y = segment, z = a): est[seg * 8 + a] (zeroed by the caller) sums the target
// tiles in the runs of the samples with a active dimensions,
// a = 0 .. min(6, key_dims).  SAMPLE_THREADS resolve the codes (one round of
// lockstep searches for up to 3^6 codes), the first wavefront merges them.
constexpr int SAMPLE_THREADS = 256;
__global__ __launch_bounds__(SAMPLE_THREADS) void sample_ranges_kernel(
    const uint64_t *__restrict__ tlo, const uint64_t *__restrict__ thi,
    const uint64_t *__restrict__ tkeys, const uint64_t *__restrict__ tfirst,
    const SegDev *__restrict__ segs, const GridDev *__restrict__ grids, uint32_t samples,
    float code_cost, uint64_t *__restrict__ est) {
    __shared__ uint32_t s_start[MAX_CODES], s_end[MAX_CODES];
    __shared__ uint64_t s_coarse[COARSE_MAX];
    const uint32_t k = blockIdx.y, sm = blockIdx.x, a = blockIdx.z;
    const SegDev S = segs[k];
    if (S.grid != k) return;
    const GridDev g = grids[k];
    const uint32_t ntq = (uint32_t)((S.qn + 63) / 64);
    const uint32_t amax = g.key_dims < 6u ? g.key_dims : 6u;
    uint64_t tiles = 0;
    if (sm < ntq && g.active == ACTIVE_AUTO && a <= amax && active_useful(S, a, code_cost)) {
        const uint32_t I = S.qtile0 + (uint32_t)(((uint64_t)sm * ntq) / min64(samples, ntq));
        const uint64_t lo0 = tlo[I] & S.local_mask, hi0 = thi[I] & S.local_mask;
        const Coarse co = stage_coarse(tfirst + S.ttile0, (uint32_t)((S.tn + 63) / 64), s_coarse,
                                       SAMPLE_THREADS);
        const int ncodes = stripe_intervals<SAMPLE_THREADS>(I, S, g, a, lo0, hi0, tkeys, tfirst, co,
                                                            s_start, s_end);
        __syncthreads();
        if (threadIdx.x < 64) tiles = merge_intervals(I, S, ncodes, s_start, s_end, nullptr).tiles;
    }
    if (threadIdx.x == 0 && tiles)
        atomicAdd((unsigned long long *)&est[(size_t)k * 8 + a], (unsigned long long)tiles);
}

// Active dimensions of each segment left open by the caller (one thread per
// segment): the a that minimizes the estimated cost
//   tile pairs(a) + code_cost * query tiles * codes(a),
// tile pairs(a) extrapolated from the samples of sample_ranges_kernel and
// code_cost the cost of resolving one stripe code of one query tile relative
// to filtering one tile pair.
__global__ void choose_active_kernel(const uint64_t *__restrict__ est, uint32_t samples,
                                     const SegDev *__restrict__ segs, uint32_t nseg,
                                     GridDev *__restrict__ grids, float code_cost) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nseg) return;
    const SegDev S = segs[k];
    if (S.grid != k) return;
    GridDev &g = grids[k];
    if (g.active != ACTIVE_AUTO) return;
    const uint64_t ntq = (S.qn + 63) / 64;
    const uint64_t ns = min64(samples, ntq);
    const uint32_t amax = g.key_dims < 6u ? g.key_dims : 6u;
    uint32_t best = 0;
    double best_cost = 0.0;
    uint64_t total = 1;
    for (uint32_t a = 0; a <= amax; a++, total *= 3) {
        if (!active_useful(S, a, code_cost)) continue;
        const uint64_t sum = est[(size_t)k * 8 + a];
        const uint64_t codes = S.self ? total - (total - 1) / 2 : total;
        const double pairs = ns ? (double)sum * (double)ntq / (double)ns : 0.0;
        const double cost = pairs + (double)code_cost * (double)ntq * (double)codes;
        if (a == 0 || cost < best_cost) {
            best = a;
            best_cost = cost;
        }
    }
    g.active = best;
}

// One wavefront per query tile: the tile's runs [s, e) of target tiles become
// consecutive list entries (I | J << 32) from off[I] on.  The runs' lengths
// are scanned across the lanes into LDS (64 runs at a time), then lane l
// writes entries l, l + 64, ... of that group of runs (a binary search over
// the run starts in LDS finds each entry's run), so the stores are coalesced
// however short the runs are.
__global__ __launch_bounds__(64) void emit_pairs_kernel(const uint32_t *__restrict__ ivals,
                                                        const uint32_t *__restrict__ icount,
                                                        const uint64_t *__restrict__ off,
                                                        size_t ntq, int max_iv,
                                                        uint64_t *__restrict__ pairs,
                                                        uint64_t cap) {
    __shared__ uint32_t s_beg[64], s_pos[65];  // run start tile, run's first entry
    const size_t I = blockIdx.x;
    if (I >= ntq) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t *iv = ivals + I * (size_t)max_iv * 2;
    uint64_t o = off[I];
    const uint32_t m = icount[I];
    for (uint32_t q0 = 0; q0 < m; q0 += 64) {
        const uint32_t q = q0 + lane;
        uint32_t st = 0, len = 0;
        if (q < m) {
            st = iv[2 * q];
            len = iv[2 * q + 1] - st;
        }
        uint32_t incl = len;  // inclusive scan of the run lengths
        for (int k = 1; k < 64; k <<= 1) {
            const uint32_t y = __shfl_up(incl, k);
            if (lane >= (uint32_t)k) incl += y;
        }
        s_beg[lane] = st;
        s_pos[lane] = incl - len;
        const uint32_t total = __shfl(incl, 63);
        if (lane == 0) s_pos[64] = total;
        __syncthreads();
        const uint32_t nr = m - q0 < 64 ? m - q0 : 64;
        for (uint32_t e = lane; e < total; e += 64) {
            // the run holding entry e: the last r with s_pos[r] <= e
            uint32_t r = 0, span = nr;
            while (span > 1) {
                const uint32_t half = span >> 1;
                r = s_pos[r + half] <= e ? r + half : r;
                span -= half;
            }
            const uint32_t J = s_beg[r] + (e - s_pos[r]);
            if (o + e < cap) pairs[o + e] = (uint64_t)I | ((uint64_t)J << 32);
        }
        o += total;
        __syncthreads();
    }
}

__device__ __forceinline__ uint64_t hilbert_xy2d(int order, uint32_t x, uint32_t y) {
    uint64_t h = 0;
    const uint64_t nside = (uint64_t)1 << order;
    for (uint64_t s = nside >> 1; s > 0; s >>= 1) {
        uint32_t rx = (x & s) ? 1u : 0u;
        uint32_t ry = (y & s) ? 1u : 0u;
        h += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = (uint32_t)(nside - 1 - x);
                y = (uint32_t)(nside - 1 - y);
            }
            uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return h;
}

__global__ void hilbert_pairs_kernel(const uint64_t *__restrict__ pairs, size_t np, int order,
                                     uint64_t *__restrict__ out) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    uint64_t p = pairs[k];
    out[k] = hilbert_xy2d(order, (uint32_t)p, (uint32_t)(p >> 32));
}

__global__ void hilbert_xy_kernel(const uint32_t *__restrict__ x, const uint32_t *__restrict__ y,
                                  size_t n, int order, uint64_t *__restrict__ out) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = hilbert_xy2d(order, x[k], y[k]);
}

__global__ __launch_bounds__(256) void copy16_kernel(const uint4 *__restrict__ src,
                                                     uint4 *__restrict__ dst, size_t n16) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16;
         k += (size_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

// Distinct key prefixes over the first a fields (a = 1..6) of each segment's
// target points: counts[seg * 8 + a].  Sorted keys make equal prefixes
// contiguous; one block per (segment, slice of PC_SLICE points).
__global__ __launch_bounds__(256) void prefix_classes_kernel(const uint64_t *__restrict__ keys,
                                                             const SegDev *__restrict__ segs,
                                                             const GridDev *__restrict__ grids,
                                                             unsigned long long *__restrict__ counts) {
    __shared__ unsigned int sc[7];
    const uint32_t sg = blockIdx.y;
    const SegDev S = segs[sg];
    const GridDev g = grids[S.grid];
    if (threadIdx.x < 7) sc[threadIdx.x] = 0;
    __syncthreads();
    unsigned int c[7] = {0, 0, 0, 0, 0, 0, 0};
    const uint64_t slice = PC_SLICE;
    const uint64_t b = (uint64_t)blockIdx.x * slice;
    for (uint64_t i = b + threadIdx.x; i < S.tn && i < b + slice; i += 256) {
        const uint64_t k = keys[S.toff + i] & S.local_mask;
        const uint64_t p = i ? (keys[S.toff + i - 1] & S.local_mask) : ~k;
        for (uint32_t a = 1; a <= 6 && a <= g.key_dims; a++) {
            const int sh = (int)(g.bits * (g.key_dims - a));
            c[a] += (k >> sh) != (p >> sh) ? 1u : 0u;
        }
    }
    for (int a = 1; a <= 6; a++)
        if (c[a]) atomicAdd(&sc[a], c[a]);
    __syncthreads();
    if (threadIdx.x >= 1 && threadIdx.x < 7 && sc[threadIdx.x])
        atomicAdd(&counts[(size_t)sg * 8 + threadIdx.x], (unsigned long long)sc[threadIdx.x]);
}

__global__ void zero_u64_kernel(uint64_t *p, size_t n) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

__global__ void store_u64_kernel(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
                                 size_t n) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

// Up to READBACK_RANGES ranges of words in one launch (a grid-stride loop per
// range).
__global__ void store_ranges_kernel(U64Ranges rs) {
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (int i = 0; i < rs.nr; i++)
        for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < rs.r[i].n; k += step)
            rs.r[i].dst[k] = rs.r[i].src[k];
}

// Per-dimension cell histograms for the dimension statistics: element (row,
// k) of the row-major input adds one to hist[hoff[k] + floor(x * r) - cmin[k]].
// With LDS the block keeps all histograms in LDS (nbins_total bins) and adds
// them to the global ones at the end; otherwise it adds to global memory.
template <bool LDS>
__global__ __launch_bounds__(256) void dim_hist_kernel(const double *__restrict__ pts, size_t n,
                                                       int d, double r,
                                                       const int64_t *__restrict__ cmin,
                                                       const uint32_t *__restrict__ hoff,
                                                       uint32_t nbins_total,
                                                       uint32_t *__restrict__ hist) {
    extern __shared__ uint32_t lh[];
    if (LDS) {
        for (uint32_t i = threadIdx.x; i < nbins_total; i += 256) lh[i] = 0;
        __syncthreads();
    }
    const size_t total = n * (size_t)d;
    const size_t per_block = (total + gridDim.x - 1) / gridDim.x;
    const size_t e0 = (size_t)blockIdx.x * per_block;
    const size_t e1 = e0 + per_block < total ? e0 + per_block : total;
    // column of element e, advanced by 256 mod d per step (no division in the loop)
    const int step = 256 % d;
    int k = (int)((e0 + threadIdx.x) % (size_t)d);
    for (size_t e = e0 + threadIdx.x; e < e1; e += 256) {
        const int64_t c = (int64_t)floor(pts[e] * r) - cmin[k];
        const uint32_t bin = hoff[k] + (uint32_t)c;
        if (LDS)
            atomicAdd(&lh[bin], 1u);
        else
            atomicAdd(&hist[bin], 1u);
        k += step;
        if (k >= d) k -= d;
    }
    if (LDS) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nbins_total; i += 256)
            if (lh[i]) atomicAdd(&hist[i], lh[i]);
    }
}

// out[i][perm[k]] = in[i][k]
__global__ __launch_bounds__(256) void permute_cols_kernel(const double *__restrict__ in,
                                                           double *__restrict__ out, size_t n,
                                                           int d,
                                                           const int32_t *__restrict__ perm) {
    const size_t total = n * (size_t)d;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (size_t)gridDim.x * 256) {
        const size_t i = e / (size_t)d;
        const int k = (int)(e - i * d);
        out[i * d + perm[k]] = in[e];
    }
}

inline unsigned blocks_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

void launch_seg_minmax(const double *pts, const uint64_t *seg_off, uint32_t nseg, int d,
                       int k0, int k1, double *mins, double *maxs, uint32_t *segbad,
                       uint32_t *clear, DevBuf &tmp, PinnedBuf &stage, hipStream_t s,
                       uint32_t sample_step, uint64_t sample_min_rows) {
    const int nd = k1 - k0;
    if (k0 < 0 || nd <= 0 || k1 > d || nseg == 0 || sample_step == 0)
        throw InvalidArg{"internal: coordinate range dimensions"};
    // chunks of at least MM_CHUNK_ROWS rows (fewer than 60000 chunks), never
    // crossing a segment (a sampled segment's chunks visit 1 / step of
    // them); seg_off is a host array of nseg + 1 offsets
    const uint64_t rows_per = std::max<uint64_t>(MM_CHUNK_ROWS, seg_off[nseg] / 60000 + 1);
    auto seg_step = [&](uint32_t g) -> uint64_t {
        return seg_off[g + 1] - seg_off[g] >= sample_min_rows ? sample_step : 1;
    };
    size_t nch = 0;
    for (uint32_t g = 0; g < nseg; g++) nch += (seg_off[g + 1] - seg_off[g] + rows_per - 1) / rows_per;
    const size_t nchunks = std::max<size_t>(nch, 1);
    const unsigned ny = (unsigned)((nd + 63) / 64);
    // one segment: the chunks follow from the offsets; several: a chunk table
    // [begin | end | first chunk of each segment | step], staged in pinned
    // memory and uploaded with one copy
    const bool table = nseg > 1;
    const size_t hdr = table ? 2 * nchunks * sizeof(uint64_t) +
                                   (nseg + 1 + nchunks) * sizeof(uint32_t)
                             : 0;
    const size_t hdr_al = (hdr + 255) / 256 * 256;
    const size_t part_bytes = 2 * nchunks * nd * sizeof(double);
    tmp.ensure(hdr_al + part_bytes + ny * nchunks * sizeof(uint32_t));
    char *base = tmp.as<char>();
    double *pmin = reinterpret_cast<double *>(base + hdr_al), *pmax = pmin + nchunks * nd;
    uint32_t *cbad = reinterpret_cast<uint32_t *>(base + hdr_al + part_bytes);
    MinmaxChunks ch{};
    ch.nchunks = (uint32_t)nchunks;
    const uint32_t *dc0 = nullptr;
    if (table) {
        stage.wait();  // the staging buffer may still feed the previous upload
        char *h = static_cast<char *>(stage.ensure(hdr));
        uint64_t *cb = reinterpret_cast<uint64_t *>(h), *ce = cb + nchunks;
        uint32_t *c0 = reinterpret_cast<uint32_t *>(ce + nchunks);
        uint32_t *cs = c0 + nseg + 1;
        size_t k = 0;
        for (uint32_t g = 0; g < nseg; g++) {
            c0[g] = (uint32_t)k;
            const uint32_t st = (uint32_t)seg_step(g);
            for (uint64_t r = seg_off[g]; r < seg_off[g + 1]; r += rows_per, k++) {
                cb[k] = r;
                ce[k] = std::min(r + rows_per, seg_off[g + 1]);
                cs[k] = st;
            }
        }
        c0[nseg] = (uint32_t)k;
        HJ_HIP(hipMemcpyAsync(base, h, hdr, hipMemcpyHostToDevice, s));
        stage.record(s);
        ch.begin = reinterpret_cast<const uint64_t *>(base);
        ch.end = ch.begin + nchunks;
        dc0 = reinterpret_cast<const uint32_t *>(base + 2 * nchunks * sizeof(uint64_t));
        ch.step = dc0 + nseg + 1;
    } else {
        ch.row0 = seg_off[0];
        ch.row_end = seg_off[1];
        ch.step1 = (uint32_t)seg_step(0);
        ch.rows_per = rows_per;
    }
    if (nch) {
        int lpr = 8;
        while (lpr < nd && lpr < 64) lpr *= 2;
        // chunks on x: a batch of many small segments can exceed 65535 chunks
        minmax_partial<<<dim3((unsigned)nch, ny), dim3(64, MM_ROWS), 0, s>>>(
            pts, d, k0, nd, lpr, ch, pmin, pmax, cbad);
        HJ_HIP(hipGetLastError());
    }
    minmax_final<<<dim3((unsigned)nd, nseg), 256, 0, s>>>(pmin, pmax, cbad, (uint32_t)nchunks, ny,
                                                          dc0, (uint32_t)nch, nd, mins, maxs,
                                                          segbad, clear);
    HJ_HIP(hipGetLastError());
}

void launch_keys(const double *pts, size_t n, int d, const SegDev *segs, uint32_t nseg,
                 const uint64_t *seg_off, const GridDev *grids, const int64_t *minc,
                 int kd_max, bool query_prefix, uint64_t *keys, uint32_t *vals, uint64_t *plo,
                 uint64_t *phi, uint32_t *widened, hipStream_t s) {
    if (n == 0) return;
    kd_max = std::max(1, kd_max);
    const size_t lds = (size_t)4 * 64 * (kd_max + 1) * sizeof(double);
    keys_kernel<<<blocks_for(n, 256), 256, lds, s>>>(pts, n, d, segs, nseg, seg_off, grids, minc,
                                                     kd_max, query_prefix ? 1 : 0, keys, vals,
                                                     plo, phi, widened);
    HJ_HIP(hipGetLastError());
}

void launch_gather(const double *pts, const uint32_t *perm, const SegDev *segs,
                   const uint32_t *tile_seg, size_t ntiles, bool target_side, int d, int dpad,
                   double *cols, double *pnorm, hipStream_t s) {
    if (ntiles == 0) return;
    if (dpad % 8 != 0) throw InvalidArg{"internal: tile rows not padded to 8 dimensions"};
    gather_kernel<<<(unsigned)ntiles, 256, 0, s>>>(pts, perm, segs, tile_seg, target_side ? 1 : 0,
                                                   d, dpad, cols, pnorm);
    HJ_HIP(hipGetLastError());
}

void launch_tile_index(const uint64_t *keys, const uint32_t *perm, const uint64_t *plo,
                       const uint64_t *phi, const SegDev *segs, const uint32_t *tile_seg,
                       size_t ntiles, bool target_side, const uint32_t *widened,
                       uint64_t *tfirst, uint64_t *tlo, uint64_t *thi, hipStream_t s) {
    if (ntiles == 0) return;
    tile_index_kernel<<<blocks_for(ntiles, 4), 256, 0, s>>>(keys, perm, plo, phi, segs, tile_seg,
                                                            ntiles, target_side ? 1 : 0, widened,
                                                            tfirst, tlo, thi);
    HJ_HIP(hipGetLastError());
}

void launch_prefix_classes(const uint64_t *keys, const SegDev *segs, uint32_t nseg,
                           const GridDev *grids, uint64_t *counts, uint64_t max_tn,
                           hipStream_t s) {
    launch_zero_u64(counts, (size_t)nseg * 8, s);
    if (nseg == 0 || max_tn == 0) return;
    dim3 grid((unsigned)((max_tn + PC_SLICE - 1) / PC_SLICE), nseg);
    prefix_classes_kernel<<<grid, 256, 0, s>>>(keys, segs, grids,
                                               reinterpret_cast<unsigned long long *>(counts));
    HJ_HIP(hipGetLastError());
}

void launch_choose_active(const uint64_t *tlo, const uint64_t *thi, const uint64_t *tkeys,
                          const uint64_t *tfirst, const SegDev *segs, uint32_t nseg,
                          GridDev *grids, uint32_t samples, uint64_t *est, float code_cost,
                          hipStream_t s) {
    if (nseg == 0 || samples == 0) return;
    launch_zero_u64(est, (size_t)nseg * 8, s);
    sample_ranges_kernel<<<dim3(samples, nseg, 7), SAMPLE_THREADS, 0, s>>>(
        tlo, thi, tkeys, tfirst, segs, grids, samples, code_cost, est);
    HJ_HIP(hipGetLastError());
    choose_active_kernel<<<blocks_for(nseg, 256), 256, 0, s>>>(est, samples, segs, nseg, grids,
                                                               code_cost);
    HJ_HIP(hipGetLastError());
}

void launch_tile_ranges(const uint64_t *tlo, const uint64_t *thi, const uint32_t *tile_seg,
                        size_t ntiles_q, const uint64_t *tkeys, const uint64_t *tfirst,
                        const SegDev *segs, const GridDev *grids, uint32_t *ivals,
                        uint32_t *icount, uint64_t *jcount, int max_iv, uint64_t *cand_total,
                        hipStream_t s) {
    if (ntiles_q == 0) return;
    if (max_iv > MAX_CODES) throw InvalidArg{"too many stripe codes"};
    tile_ranges_kernel<<<(unsigned)ntiles_q, 64, 0, s>>>(tlo, thi, tile_seg, ntiles_q, tkeys,
                                                         tfirst, segs, grids, ivals, icount,
                                                         jcount, max_iv, cand_total);
    HJ_HIP(hipGetLastError());
}

void launch_emit_pairs(const uint32_t *ivals, const uint32_t *icount, const uint64_t *off,
                       size_t ntiles_q, int max_iv, uint64_t *pairs, uint64_t cap,
                       hipStream_t s) {
    if (ntiles_q == 0) return;
    emit_pairs_kernel<<<(unsigned)ntiles_q, 64, 0, s>>>(ivals, icount, off, ntiles_q, max_iv,
                                                        pairs, cap);
    HJ_HIP(hipGetLastError());
}

void launch_hilbert_keys(const uint64_t *pairs, size_t npairs, int order, uint64_t *hkeys,
                         hipStream_t s) {
    if (npairs == 0) return;
    hilbert_pairs_kernel<<<blocks_for(npairs, 256), 256, 0, s>>>(pairs, npairs, order, hkeys);
    HJ_HIP(hipGetLastError());
}

void launch_hilbert_xy(const uint32_t *x, const uint32_t *y, size_t n, int order,
                       uint64_t *out, hipStream_t s) {
    if (n == 0) return;
    hilbert_xy_kernel<<<blocks_for(n, 256), 256, 0, s>>>(x, y, n, order, out);
    HJ_HIP(hipGetLastError());
}

void launch_copy16(const void *src, void *dst, size_t bytes, int num_cus, hipStream_t s) {
    size_t n16 = bytes / 16;
    if (n16 == 0) return;
    size_t blocks = std::min<size_t>((n16 + 255) / 256, (size_t)std::max(1, num_cus) * 16);
    copy16_kernel<<<(unsigned)blocks, 256, 0, s>>>(static_cast<const uint4 *>(src),
                                                   static_cast<uint4 *>(dst), n16);
    HJ_HIP(hipGetLastError());
}

void launch_dim_hist(const double *pts, size_t n, int d, double r, const int64_t *cmin,
                     const uint32_t *hoff, uint32_t nbins_total, uint32_t *hist, int num_cus,
                     hipStream_t s) {
    if (n == 0) return;
    const size_t total = n * (size_t)d;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 4095) / 4096,
                                                       (size_t)std::max(1, num_cus) * 8);
    if ((size_t)nbins_total * 4 <= DIM_HIST_LDS_BYTES) {
        dim_hist_kernel<true><<<blocks, 256, (size_t)nbins_total * 4, s>>>(
            pts, n, d, r, cmin, hoff, nbins_total, hist);
    } else {
        dim_hist_kernel<false><<<blocks, 256, 0, s>>>(pts, n, d, r, cmin, hoff, nbins_total,
                                                      hist);
    }
    HJ_HIP(hipGetLastError());
}

void launch_permute_cols(const double *in, double *out, size_t n, int d, const int32_t *perm,
                         int num_cus, hipStream_t s) {
    if (n == 0) return;
    const size_t total = n * (size_t)d;
    const unsigned blocks =
        (unsigned)std::min<size_t>((total + 255) / 256, (size_t)std::max(1, num_cus) * 16);
    permute_cols_kernel<<<blocks, 256, 0, s>>>(in, out, n, d, perm);
    HJ_HIP(hipGetLastError());
}

void launch_zero_u64(uint64_t *p, size_t n, hipStream_t s) {
    if (n == 0) return;
    zero_u64_kernel<<<blocks_for(n, 256), 256, 0, s>>>(p, n);
    HJ_HIP(hipGetLastError());
}

void launch_store_u64(const uint64_t *src, uint64_t *dst, size_t n, hipStream_t s) {
    if (n == 0) return;
    store_u64_kernel<<<blocks_for(n, 256), 256, 0, s>>>(src, dst, n);
    HJ_HIP(hipGetLastError());
}

void readback_ranges(PinnedBuf &buf, const U64Ranges &rs, hipStream_t s) {
    if (rs.nr < 0 || rs.nr > READBACK_RANGES) throw InvalidArg{"internal: read-back ranges"};
    size_t most = 0;
    for (int i = 0; i < rs.nr; i++) {
        // dst: byte offsets into buf
        const size_t off = reinterpret_cast<size_t>(rs.r[i].dst);
        if (off + rs.r[i].n * sizeof(uint64_t) > buf.bytes)
            throw InvalidArg{"internal: read-back past its buffer"};
        most = std::max<size_t>(most, rs.r[i].n);
    }
    if (most == 0) return;
    char *dp = static_cast<char *>(buf.dev_ptr());
    if (!dp) {
        for (int i = 0; i < rs.nr; i++)
            if (rs.r[i].n)
                HJ_HIP(hipMemcpyAsync(buf.as<char>() + reinterpret_cast<size_t>(rs.r[i].dst),
                                      rs.r[i].src, rs.r[i].n * sizeof(uint64_t),
                                      hipMemcpyDeviceToHost, s));
        return;
    }
    U64Ranges d = rs;
    for (int i = 0; i < d.nr; i++)
        d.r[i].dst = reinterpret_cast<uint64_t *>(dp + reinterpret_cast<size_t>(rs.r[i].dst));
    const unsigned blocks = (unsigned)std::min<size_t>(blocks_for(most, 256), 64);
    store_ranges_kernel<<<blocks, 256, 0, s>>>(d);
    HJ_HIP(hipGetLastError());
}

void readback_u64(PinnedBuf &buf, size_t off, const uint64_t *src, size_t n, hipStream_t s) {
    if (off + n * sizeof(uint64_t) > buf.bytes) throw InvalidArg{"internal: read-back past its buffer"};
    char *dp = static_cast<char *>(buf.dev_ptr());
    if (dp)
        launch_store_u64(src, reinterpret_cast<uint64_t *>(dp + off), n, s);
    else
        HJ_HIP(hipMemcpyAsync(buf.as<char>() + off, src, n * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, s));
}

}  // namespace hj

// ---------------------------------------------------------------------------
// CSR post-processing
// ---------------------------------------------------------------------------
namespace hj {
namespace {

// layout (row, col) -> sort key (query input row) << 32 | (target input row -
// toff of the pair's segment); the segment id goes to out_seg.  Mirrored self
// pairs (target row as the row) take the second half of the output.
__global__ void map_pairs_kernel(const uint64_t *__restrict__ in, uint64_t np,
                                 const SegDev *__restrict__ segs,
                                 const uint32_t *__restrict__ tile_seg,
                                 const uint32_t *__restrict__ qperm,
                                 const uint32_t *__restrict__ tperm, int mirror,
                                 uint64_t *__restrict__ keys, uint32_t *__restrict__ out_seg) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint64_t p = in[k];
    const uint32_t row = (uint32_t)p, col = (uint32_t)(p >> 32);
    const uint32_t sg = tile_seg[row >> 6];
    const SegDev &S = segs[sg];
    const uint64_t qs = S.qoff + 64ull * ((row >> 6) - S.qtile0) + (row & 63);
    const uint64_t ts = S.toff + 64ull * ((col >> 6) - S.ttile0) + (col & 63);
    const uint64_t a = qperm[qs], b = tperm[ts];
    keys[k] = (a << 32) | (b - S.toff);
    out_seg[k] = sg;
    if (mirror) {
        keys[np + k] = (b << 32) | (a - S.qoff);
        out_seg[np + k] = sg;
    }
}

__global__ void csr_rowptr_kernel(const uint64_t *__restrict__ keys, uint64_t nk,
                                  uint64_t nrows, int64_t *__restrict__ rowptr) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrows) return;
    uint64_t x = r << 32;
    uint64_t lo = 0, hi = nk;
    while (lo < hi) {
        uint64_t m = lo + ((hi - lo) >> 1);
        if (keys[m] < x)
            lo = m + 1;
        else
            hi = m;
    }
    rowptr[r] = (int64_t)lo;
}

__global__ void csr_cols_kernel(const uint64_t *__restrict__ keys, uint64_t nk,
                                int64_t *__restrict__ colidx) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nk) colidx[k] = (int64_t)(keys[k] & 0xffffffffull);
}

}  // namespace

void launch_map_pairs(const uint64_t *in, uint64_t np, const SegDev *segs,
                      const uint32_t *tile_seg, const uint32_t *qperm, const uint32_t *tperm,
                      bool mirror, uint64_t *out_keys, uint32_t *out_seg, hipStream_t s) {
    if (np == 0) return;
    map_pairs_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(
        in, np, segs, tile_seg, qperm, tperm, mirror ? 1 : 0, out_keys, out_seg);
    HJ_HIP(hipGetLastError());
}

void launch_csr_rowptr(const uint64_t *keys, uint64_t nk, uint64_t nrows, int64_t *rowptr,
                       hipStream_t s) {
    csr_rowptr_kernel<<<(unsigned)((nrows + 1 + 255) / 256), 256, 0, s>>>(keys, nk, nrows,
                                                                          rowptr);
    HJ_HIP(hipGetLastError());
}

void launch_csr_cols(const uint64_t *keys, uint64_t nk, int64_t *colidx, hipStream_t s) {
    if (nk == 0) return;
    csr_cols_kernel<<<(unsigned)((nk + 255) / 256), 256, 0, s>>>(keys, nk, colidx);
    HJ_HIP(hipGetLastError());
}

}  // namespace hj
