// hj_mfma.hip -- the reduced-precision MFMA filter with an FP64 band recheck.
//
// Same work list and count semantics as the FP64 kernel (the reference's tile
// loop EGO_LOOP_TRAN egojoin.h:311-392); only the distance arithmetic differs.
// Each wavefront computes the 64 x 64 block of a tile pair with four
// v_mfma_f32_32x32x16_{f16,bf16} accumulators (FP32 accumulate).  Operands are
// the centred, power-of-two-scaled coordinates y = (x - c) s rounded to
// FP16/BF16, extended by a few K slots that fold the squared norms in, so
//     acc = a.b - |a|^2/2 - |b|^2/2 = -|a - b|^2 / 2      (rounded operands)
// The host derives thresholds T_out <= T_in from a worst-case bound of
// operand rounding, FP32 accumulation and norm splitting (filter_thresholds
// in hj_api.cpp):
//     acc <  T_out        the pair is certainly farther than eps
//     acc >= T_in         the pair is certainly within eps
//     otherwise           the pair is queued and rechecked with the FP64
//                         fixed-order predicate
// so the count equals the FP64 oracle's (no false negatives or positives).
#include "hj_device.h"
#include "hj_internal.h"
#include "hj_work.h"

namespace hj {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// value of y after rounding to the operand type, as a double
__device__ __forceinline__ double round_op(double y, bool bf16) {
    float f = (float)y;
    if (bf16) return (double)(float)(__bf16)f;
    return (double)(float)(_Float16)f;
}

__device__ __forceinline__ uint16_t bits_op(double y, bool bf16) {
    float f = (float)y;
    if (bf16) return __builtin_bit_cast(uint16_t, (__bf16)f);
    return __builtin_bit_cast(uint16_t, (_Float16)f);
}

// Squared norms (FP64) of each point's rounded, scaled, centred coordinates,
// split at dimension `split`: norms[2p] over [0, split), norms[2p+1] over
// [split, d).  Squares of FP16/BF16 values are exact in FP64.
__global__ void op_norms_kernel(const double *__restrict__ rows, size_t n, int d, int dpad,
                                int split, const double *__restrict__ centre, double scale,
                                int bf16, double *__restrict__ norms) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < d; k++) {
        double v = round_op((rows[p * dpad + k] - centre[k]) * scale, bf16 != 0);
        if (k < split)
            s0 = __builtin_fma(v, v, s0);
        else
            s1 = __builtin_fma(v, v, s1);
    }
    norms[2 * p] = s0;
    norms[2 * p + 1] = s1;
}

struct PackParams {
    int d, dpad, split, steps1, steps, parts, bf16;
    double scale;
};

constexpr double NORM_UNIT = 32768.0;  // 2^15: the partner of each norm part
constexpr double PAD_PART = 32768.0;   // norm part of pad points: acc <= -2^30

// Operand fragments.  The K axis is cut in two chunks: k-steps [0, steps1)
// hold dimensions [0, split) followed by 2*parts norm slots of those
// dimensions; k-steps [steps1, steps) hold dimensions [split, d) and the norm
// slots of the rest.  Row-role arrays carry the point's norm parts in the first
// `parts` slots and -2^15 in the next `parts`; column-role arrays the reverse,
// so a row x column product adds -|a_c|^2/2 - |b_c|^2/2 for chunk c and the
// accumulator after chunk 1 is -(partial distance)^2/2, after both chunks
// -(distance)^2/2.  One wavefront per (tile, 32-point half, k-step): lane l
// holds point tile*64 + half*32 + l%32, K slots 16 t + 8 (l/32) + j, j = 0..7
// -- the A/B operand layout of v_mfma_f32_32x32x16.
__global__ __launch_bounds__(64) void frag_pack_kernel(const double *__restrict__ rows,
                                                       size_t n, const double *__restrict__ centre,
                                                       const double *__restrict__ norms,
                                                       PackParams pp, u32x4 *__restrict__ frow,
                                                       u32x4 *__restrict__ fcol) {
    const size_t blk = blockIdx.x;
    const int t = (int)(blk % pp.steps);
    const size_t th = blk / pp.steps;  // tile*2 + half
    const int lane = threadIdx.x;
    const size_t p = th * 32 + (lane & 31);
    const bool bf = pp.bf16 != 0;
    const bool first = t < pp.steps1;
    const int dim0 = first ? 0 : pp.split;
    const int ndata = first ? pp.split : pp.d - pp.split;
    const int slot0 = (first ? t : t - pp.steps1) * 16 + 8 * (lane >> 5);
    double part[3] = {0, 0, 0};
    if (p < n) {
        double v = norms[2 * p + (first ? 0 : 1)] * 0.5 / NORM_UNIT;
        for (int q = 0; q < pp.parts; q++) {
            part[q] = round_op(v, bf);
            v -= part[q];
        }
    } else {
        part[0] = PAD_PART;
    }
    uint16_t hr[8], hc[8];
    for (int j = 0; j < 8; j++) {
        const int q = slot0 + j;
        double vr = 0.0, vc = 0.0;
        if (q < ndata) {
            const int k = dim0 + q;
            if (p < n) vr = vc = (rows[p * pp.dpad + k] - centre[k]) * pp.scale;
        } else if (q < ndata + pp.parts) {
            vr = part[q - ndata];
            vc = -NORM_UNIT;
        } else if (q < ndata + 2 * pp.parts) {
            vr = -NORM_UNIT;
            vc = part[q - ndata - pp.parts];
        }
        hr[j] = bits_op(vr, bf);
        hc[j] = bits_op(vc, bf);
    }
    u32x4 r, c;
    for (int q = 0; q < 4; q++) {
        r[q] = (uint32_t)hr[2 * q] | ((uint32_t)hr[2 * q + 1] << 16);
        c[q] = (uint32_t)hc[2 * q] | ((uint32_t)hc[2 * q + 1] << 16);
    }
    const size_t o = (th * pp.steps + t) * 64 + lane;
    frow[o] = r;
    fcol[o] = c;
}

// Fused preparation of one 64-point tile for the filter (d <= FILTER_PREP_MAX_D):
// the tile's rows are gathered through the EGO permutation into LDS with
// row-contiguous loads, then the block computes the tile's widened key bounds
// (the FP64 path's tile_bounds_kernel), the chunk norms and all operand
// fragments (the layout of frag_pack_kernel) without an intermediate copy.
constexpr int FILTER_PREP_MAX_D = 128;

__global__ __launch_bounds__(256) void filter_prep_kernel(
    const double *__restrict__ pts, const uint32_t *__restrict__ perm, size_t n,
    const double *__restrict__ centre, PackParams pp, KeyParams kp,
    const int64_t *__restrict__ minc, uint64_t *__restrict__ tlo, uint64_t *__restrict__ thi,
    u32x4 *__restrict__ frow, u32x4 *__restrict__ fcol) {
    extern __shared__ double sx[];  // [64][d + 1] raw coordinates of the tile
    __shared__ double snorm[4][64][2];
    const int t = threadIdx.x;
    const size_t I = blockIdx.x;
    const int d = pp.d, stride = d + 1;
    const bool bf = pp.bf16 != 0;
    for (int e = t; e < 64 * d; e += 256) {
        const int p = e / d, k = e - p * d;
        const size_t i = I * 64 + p;
        sx[p * stride + k] = i < n ? pts[(size_t)perm[i] * d + k] : 0.0;
    }
    __syncthreads();
    const int q = t & 63, g = t >> 6;
    {
        double s0 = 0.0, s1 = 0.0;
        for (int k = g; k < d; k += 4) {
            const double v = round_op((sx[q * stride + k] - centre[k]) * pp.scale, bf);
            if (k < pp.split)
                s0 = __builtin_fma(v, v, s0);
            else
                s1 = __builtin_fma(v, v, s1);
        }
        snorm[g][q][0] = s0;
        snorm[g][q][1] = s1;
    }
    if (g == 0) {
        uint64_t lo = ~0ull, hi = 0ull;
        if (I * 64 + q < n) point_key_bounds(sx + q * stride, 1, kp, minc, &lo, &hi);
        wave_min_max_u64(lo, hi);
        if (q == 0) {
            tlo[I] = lo;
            thi[I] = hi;
        }
    }
    __syncthreads();
    for (int e = t; e < 2 * pp.steps * 64; e += 256) {
        const int lane = e & 63, hs = e >> 6;
        const int h = hs / pp.steps, st = hs - h * pp.steps;
        const int pt = h * 32 + (lane & 31);
        const size_t p = I * 64 + pt;
        const bool first = st < pp.steps1;
        const int dim0 = first ? 0 : pp.split;
        const int ndata = first ? pp.split : d - pp.split;
        const int slot0 = (first ? st : st - pp.steps1) * 16 + 8 * (lane >> 5);
        double part[3] = {0, 0, 0};
        if (p < n) {
            const int c = first ? 0 : 1;
            double v = (snorm[0][pt][c] + snorm[1][pt][c] + snorm[2][pt][c] + snorm[3][pt][c]) *
                       0.5 / NORM_UNIT;
            for (int r = 0; r < pp.parts; r++) {
                part[r] = round_op(v, bf);
                v -= part[r];
            }
        } else {
            part[0] = PAD_PART;
        }
        uint16_t hr[8], hc[8];
        for (int j = 0; j < 8; j++) {
            const int qq = slot0 + j;
            double vr = 0.0, vc = 0.0;
            if (qq < ndata) {
                const int k = dim0 + qq;
                if (p < n) vr = vc = (sx[pt * stride + k] - centre[k]) * pp.scale;
            } else if (qq < ndata + pp.parts) {
                vr = part[qq - ndata];
                vc = -NORM_UNIT;
            } else if (qq < ndata + 2 * pp.parts) {
                vr = -NORM_UNIT;
                vc = part[qq - ndata - pp.parts];
            }
            hr[j] = bits_op(vr, bf);
            hc[j] = bits_op(vc, bf);
        }
        u32x4 r, c;
        for (int w = 0; w < 4; w++) {
            r[w] = (uint32_t)hr[2 * w] | ((uint32_t)hr[2 * w + 1] << 16);
            c[w] = (uint32_t)hc[2 * w] | ((uint32_t)hc[2 * w + 1] << 16);
        }
        const size_t o = ((I * 2 + h) * pp.steps + st) * 64 + lane;
        frow[o] = r;
        fcol[o] = c;
    }
}

template <bool BF16>
__device__ __forceinline__ f32x16 mfma(u32x4 a, u32x4 b, f32x16 c) {
    if constexpr (BF16)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8, a),
                                                      __builtin_bit_cast(half8, b), c, 0, 0, 0);
}

__device__ __forceinline__ float max16(const f32x16 &v) {
    float m0 = fmaxf(fmaxf(v[0], v[1]), v[2]);
    float m1 = fmaxf(fmaxf(v[3], v[4]), v[5]);
    float m2 = fmaxf(fmaxf(v[6], v[7]), v[8]);
    float m3 = fmaxf(fmaxf(v[9], v[10]), v[11]);
    float m4 = fmaxf(fmaxf(v[12], v[13]), v[14]);
    return fmaxf(fmaxf(fmaxf(m0, m1), fmaxf(m2, m3)), fmaxf(m4, v[15]));
}

// Classifies the 32 x 32 block at (row0, col) of a tile pair: counts the
// certain pairs in a per-lane counter and appends band pairs (sorted indices
// i | j << 32) with per-lane atomics -- band pairs are rare, so no
// wavefront-wide ballots are spent on them.
template <bool SELF>
__device__ __forceinline__ void classify(const f32x16 &acc, uint32_t row0, uint32_t col,
                                         uint32_t nA, uint32_t nB, bool diag, float t_in,
                                         float t_out, uint32_t &cnt, uint64_t *__restrict__ ctr,
                                         uint64_t *__restrict__ band, uint64_t band_cap) {
    const uint32_t lane = threadIdx.x & 63;
    const bool vcol = col < nB;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t row = row0 + 4 * (lane >> 5) + (r & 3) + 8 * (r >> 2);
        const bool valid = vcol && row < nA && (!SELF || !diag || row < col);
        const float v = acc[r];
        cnt += (valid && v >= t_in) ? 1u : 0u;
        if (valid && v >= t_out && v < t_in) {
            const uint64_t o = atomicAdd((unsigned long long *)&ctr[4], 1ull);
            if (o < band_cap) band[o] = (uint64_t)row | ((uint64_t)col << 32);
        }
    }
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_u64(v, m);
    return v;
}

// Tile pairs per chunk of the filter's schedule: the column fragments of all
// pairs of a chunk are loaded before the first MFMA of the chunk, so their
// memory latency overlaps.
constexpr uint32_t MFMA_CHUNK = 4;

// Full K of one tile pair followed by the classification of its 64 x 64
// pairs, one 32 x 32 block at a time with a single accumulator (the slow path:
// diagonal tiles and blocks where some pair may be within eps).
template <bool SELF, bool BF16>
__device__ __forceinline__ uint32_t full_tile_pair(
    const u32x4 *__restrict__ fa, const u32x4 *__restrict__ fb, uint32_t I, uint32_t J, int steps,
    uint32_t nA, uint32_t nB, float t_in, float t_out, uint64_t *__restrict__ ctr,
    uint64_t *__restrict__ band, uint64_t band_cap) {
    const uint32_t lane = threadIdx.x & 63;
    const size_t half = (size_t)steps * 64;
    const bool diag = SELF && I == J;
    uint32_t c = 0;
#pragma unroll 1
    for (int blk = 0; blk < 4; blk++) {
        const uint32_t bi = blk >> 1, bj = blk & 1;
        const u32x4 *a = fa + (size_t)(2 * I + bi) * half + lane;
        const u32x4 *b = fb + (size_t)(2 * J + bj) * half + lane;
        f32x16 acc = {};
        for (int t = 0; t < steps; t++) acc = mfma<BF16>(a[t * 64], b[t * 64], acc);
        classify<SELF>(acc, I * 64 + bi * 32, J * 64 + bj * 32 + (lane & 31), nA, nB, diag, t_in,
                       t_out, c, ctr, band, band_cap);
    }
    return c;
}

// One wavefront per 64 x 64 tile pair.  Fast path: the first K chunk
// (k-steps [0, steps1); all of K when steps1 == steps) gives, in four 32 x 32
// accumulators, -(partial distance)^2 / 2 for all 4096 pairs; when every pair
// is certainly farther than eps (a partial distance never exceeds the full
// one) the tile pair is done.  Diagonal tiles and blocks where some pair may
// be within eps go to full_tile_pair.  The operand fragments of a whole chunk
// of tile pairs are loaded before its first MFMA, and the schedule keeps each
// XCD on one segment of the list.
template <bool SELF, bool BF16>
__global__ __launch_bounds__(256) void mfma_count_kernel(
    const u32x4 *__restrict__ fa, const u32x4 *__restrict__ fb, const uint64_t *__restrict__ pairs,
    uint64_t npairs, uint32_t nA, uint32_t nB, int steps1, int steps, float t_in, float t_out,
    uint64_t *__restrict__ ctr, uint64_t *__restrict__ band, uint64_t band_cap) {
    const uint32_t lane = threadIdx.x & 63;
    uint64_t cand = 0, exits = 0, cnt = 0;
    WorkCursor<MFMA_CHUNK, true> cur(npairs);
    const size_t half = (size_t)steps * 64;
    uint64_t w0;
    uint32_t len;
    uint32_t curI = 0xFFFFFFFFu;
    u32x4 A0c = {}, A1c = {};  // k-step-0 row fragments of row tile curI
    while (cur.next_chunk(w0, len)) {
        uint32_t slow = 0;
        uint64_t pr[MFMA_CHUNK];
        u32x4 B0[MFMA_CHUNK], B1[MFMA_CHUNK];
#pragma unroll
        for (uint32_t q = 0; q < MFMA_CHUNK; q++) pr[q] = pairs[w0 + (q < len ? q : 0)];
#pragma unroll
        for (uint32_t q = 0; q < MFMA_CHUNK; q++) {
            const u32x4 *b = fb + (size_t)(2 * (uint32_t)(pr[q] >> 32)) * half + lane;
            B0[q] = b[0];
            B1[q] = b[half];
        }
#pragma unroll
        for (uint32_t q = 0; q < MFMA_CHUNK; q++) {
            if (q >= len) break;
            const uint32_t I = (uint32_t)pr[q], J = (uint32_t)(pr[q] >> 32);
            const bool diag = SELF && I == J;
            const uint32_t nvi = min(64u, nA - I * 64), nvj = min(64u, nB - J * 64);
            cand += diag ? (uint64_t)nvj * (nvj - 1) / 2 : (uint64_t)nvi * nvj;
            if (diag) {
                slow |= 1u << q;
                continue;
            }
            const u32x4 *a0 = fa + (size_t)(2 * I) * half + lane;
            if (I != curI) {  // consecutive pairs in row order share the row tile
                A0c = a0[0];
                A1c = a0[half];
                curI = I;
            }
            const u32x4 *b0 = fb + (size_t)(2 * J) * half + lane;
            float m = -INFINITY;
#pragma unroll
            for (int bi = 0; bi < 2; bi++) {
                f32x16 c0 = {}, c1 = {};
                const u32x4 Ab = bi ? A1c : A0c;
                c0 = mfma<BF16>(Ab, B0[q], c0);
                c1 = mfma<BF16>(Ab, B1[q], c1);
                for (int t = 1; t < steps1; t++) {
                    const u32x4 At = a0[bi * half + t * 64], Bt0 = b0[t * 64],
                                Bt1 = b0[half + t * 64];
                    c0 = mfma<BF16>(At, Bt0, c0);
                    c1 = mfma<BF16>(At, Bt1, c1);
                }
                m = fmaxf(m, fmaxf(max16(c0), max16(c1)));
            }
            if (__any(m >= t_out))
                slow |= 1u << q;
            else
                exits++;
        }
        while (slow) {
            const uint32_t q = __builtin_ctz(slow);
            slow &= slow - 1;
            const uint64_t p = pairs[w0 + q];
            cnt += full_tile_pair<SELF, BF16>(fa, fb, (uint32_t)p, (uint32_t)(p >> 32), steps, nA,
                                              nB, t_in, t_out, ctr, band, band_cap);
        }
    }
    cnt = wave_sum_u64(cnt);
    if (lane == 0) {
        atomicAdd((unsigned long long *)&ctr[1], (unsigned long long)cnt);
        atomicAdd((unsigned long long *)&ctr[3], (unsigned long long)cand);
        atomicAdd((unsigned long long *)&ctr[7], (unsigned long long)exits);
    }
}

// FP64 fixed-order recheck of the queued band pairs (sorted indices, read
// through the EGO permutations from the original row-major points).
__global__ __launch_bounds__(256) void band_recheck_kernel(
    const double *__restrict__ pa, const uint32_t *__restrict__ perma,
    const double *__restrict__ pb, const uint32_t *__restrict__ permb, int d, double eps2,
    const uint64_t *__restrict__ band, uint64_t nband, uint64_t *__restrict__ ctr) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    if (k < nband) {
        const uint64_t e = band[k];
        const double *a = pa + (size_t)perma[(uint32_t)e] * d;
        const double *b = pb + (size_t)permb[(uint32_t)(e >> 32)] * d;
        double s = 0.0;
        int c = 0;
        for (; c < d; c++) {
            const double t = a[c] - b[c];
            s = __builtin_fma(t, t, s);
            if (s > eps2) break;
        }
        in = c == d && s <= eps2;
    }
    const uint64_t bal = __ballot(in);
    if ((threadIdx.x & 63) == 0 && bal)
        atomicAdd((unsigned long long *)&ctr[1], (unsigned long long)__popcll(bal));
}

}  // namespace

void launch_op_norms(const double *rows, size_t n, int d, int dpad, int split,
                     const double *centre, double scale, bool bf16, double *norms,
                     hipStream_t s) {
    if (n == 0) return;
    op_norms_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(rows, n, d, dpad, split, centre,
                                                                scale, bf16 ? 1 : 0, norms);
    HJ_HIP(hipGetLastError());
}

void launch_frag_pack(const double *rows, size_t n, int d, int dpad, const double *centre,
                      const double *norms, double scale, const FilterParams &fp, void *frow,
                      void *fcol, hipStream_t s) {
    size_t ntiles = (n + 63) / 64;
    if (ntiles == 0) return;
    PackParams pp{d, dpad, fp.split, fp.steps1, fp.steps, fp.parts, fp.bf16, scale};
    frag_pack_kernel<<<(unsigned)(ntiles * 2 * fp.steps), 64, 0, s>>>(
        rows, n, centre, norms, pp, static_cast<u32x4 *>(frow), static_cast<u32x4 *>(fcol));
    HJ_HIP(hipGetLastError());
}

template <bool SELF, bool BF16>
void mfma_launch(const u32x4 *a, const u32x4 *b, const uint64_t *pairs, uint64_t npairs,
                 uint32_t nA, uint32_t nB, const FilterParams &fp, uint64_t *counters,
                 uint64_t *band, uint64_t band_cap, int grid, int num_cus, hipStream_t s) {
    auto k = mfma_count_kernel<SELF, BF16>;
    k<<<persistent_grid(k, 256, num_cus, grid), 256, 0, s>>>(a, b, pairs, npairs, nA, nB,
                                                             fp.steps1, fp.steps, fp.t_in,
                                                             fp.t_out, counters, band, band_cap);
}

void launch_mfma_count(const void *fragA, const void *fragB, const uint64_t *pairs,
                       uint64_t npairs, uint32_t nA, uint32_t nB, const FilterParams &fp,
                       bool self, uint64_t *counters, uint64_t *band, uint64_t band_cap,
                       int grid, int num_cus, hipStream_t s) {
    if (npairs == 0) return;
    const u32x4 *a = static_cast<const u32x4 *>(fragA);
    const u32x4 *b = static_cast<const u32x4 *>(fragB);
    if (self) {
        if (fp.bf16)
            mfma_launch<true, true>(a, b, pairs, npairs, nA, nB, fp, counters, band, band_cap,
                                    grid, num_cus, s);
        else
            mfma_launch<true, false>(a, b, pairs, npairs, nA, nB, fp, counters, band, band_cap,
                                     grid, num_cus, s);
    } else {
        if (fp.bf16)
            mfma_launch<false, true>(a, b, pairs, npairs, nA, nB, fp, counters, band, band_cap,
                                     grid, num_cus, s);
        else
            mfma_launch<false, false>(a, b, pairs, npairs, nA, nB, fp, counters, band, band_cap,
                                      grid, num_cus, s);
    }
    HJ_HIP(hipGetLastError());
}

void launch_band_recheck(const double *ptsA, const uint32_t *permA, const double *ptsB,
                         const uint32_t *permB, int d, double eps2, const uint64_t *band,
                         uint64_t nband, uint64_t *counters, hipStream_t s) {
    if (nband == 0) return;
    band_recheck_kernel<<<(unsigned)((nband + 255) / 256), 256, 0, s>>>(
        ptsA, permA, ptsB, permB, d, eps2, band, nband, counters);
    HJ_HIP(hipGetLastError());
}

bool filter_prep_fused(int d) { return d <= FILTER_PREP_MAX_D; }

void launch_filter_prep(const double *pts, const uint32_t *perm, size_t n, int d,
                        const double *centre, double scale, const FilterParams &fp,
                        const Grid &g, const int64_t *minc, uint64_t *tlo, uint64_t *thi,
                        void *frow, void *fcol, hipStream_t s) {
    size_t ntiles = (n + 63) / 64;
    if (ntiles == 0) return;
    PackParams pp{d, d, fp.split, fp.steps1, fp.steps, fp.parts, fp.bf16, scale};
    KeyParams kp{g.r, g.key_dims, g.bits};
    const size_t lds = (size_t)64 * (d + 1) * sizeof(double);
    filter_prep_kernel<<<(unsigned)ntiles, 256, lds, s>>>(pts, perm, n, centre, pp, kp, minc, tlo,
                                                          thi, static_cast<u32x4 *>(frow),
                                                          static_cast<u32x4 *>(fcol));
    HJ_HIP(hipGetLastError());
}

}  // namespace hj
