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

// Classifies the 32 x 32 block (bi, bj) of a tile pair: counts certain
// pairs, queues band pairs (sorted indices i | j << 32).
template <bool SELF>
__device__ __forceinline__ void classify(const f32x16 &acc, uint32_t row0, uint32_t col,
                                         uint32_t nA, uint32_t nB, bool diag, float t_in,
                                         float t_out, uint64_t &cnt, uint64_t *__restrict__ ctr,
                                         uint64_t *__restrict__ band, uint64_t band_cap) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t row = row0 + 4 * (lane >> 5) + (r & 3) + 8 * (r >> 2);
        const bool valid = row < nA && col < nB && (!SELF || !diag || row < col);
        const float v = acc[r];
        const bool in = valid && v >= t_in;
        const bool bnd = valid && v >= t_out && v < t_in;
        cnt += __popcll(__ballot(in));
        const uint64_t bb = __ballot(bnd);
        if (bb) {
            uint64_t base = 0;
            if (lane == 0)
                base = atomicAdd((unsigned long long *)&ctr[4], (unsigned long long)__popcll(bb));
            base = __shfl(base, 0);
            if (bnd) {
                const uint64_t o = base + __popcll(bb & lt);
                if (o < band_cap) band[o] = (uint64_t)row | ((uint64_t)col << 32);
            }
        }
    }
}

// One wavefront per 64 x 64 tile pair: four 32 x 32 accumulators.  With
// EXIT, the first chunk (k-steps [0, steps1)) yields -(partial distance)^2/2;
// a block whose every pair is already certainly farther than eps skips the
// rest of K (a partial distance never exceeds the full one).  The row tile's
// fragments of that first chunk stay in registers while consecutive work items
// share the row tile.
template <bool SELF, bool BF16, bool EXIT>
__global__ __launch_bounds__(256) void mfma_count_kernel(
    const u32x4 *__restrict__ fa, const u32x4 *__restrict__ fb, const uint64_t *__restrict__ pairs,
    uint64_t npairs, uint32_t nA, uint32_t nB, int steps1, int steps, float t_in, float t_out,
    uint64_t *__restrict__ ctr, uint64_t *__restrict__ band, uint64_t band_cap) {
    const uint32_t lane = threadIdx.x & 63;
    uint64_t cnt = 0, cand = 0, exits = 0;
    WorkCursor cur(npairs);
    uint64_t w;
    uint32_t prevI = 0xFFFFFFFFu;
    u32x4 A0f = {}, A1f = {};  // row fragments of k-step 0 (EXIT: the whole first chunk)
    while (cur.next(w)) {
        const uint64_t pr = pairs[w];
        const uint32_t I = (uint32_t)pr, J = (uint32_t)(pr >> 32);
        const bool diag = SELF && I == J;
        const uint32_t nvi = min(64u, nA - I * 64), nvj = min(64u, nB - J * 64);
        cand += diag ? (uint64_t)nvj * (nvj - 1) / 2 : (uint64_t)nvi * nvj;
        const u32x4 *a0 = fa + (size_t)(2 * I) * steps * 64 + lane;
        const u32x4 *a1 = a0 + (size_t)steps * 64;
        const u32x4 *b0 = fb + (size_t)(2 * J) * steps * 64 + lane;
        const u32x4 *b1 = b0 + (size_t)steps * 64;
        if (I != prevI) {
            A0f = a0[0];
            A1f = a1[0];
            prevI = I;
        }
        f32x16 c00 = {}, c01 = {}, c10 = {}, c11 = {};
        {
            const u32x4 B0 = b0[0], B1 = b1[0];
            c00 = mfma<BF16>(A0f, B0, c00);
            c01 = mfma<BF16>(A0f, B1, c01);
            c10 = mfma<BF16>(A1f, B0, c10);
            c11 = mfma<BF16>(A1f, B1, c11);
        }
        for (int t = 1; t < steps1; t++) {
            const u32x4 A0 = a0[t * 64], A1 = a1[t * 64], B0 = b0[t * 64], B1 = b1[t * 64];
            c00 = mfma<BF16>(A0, B0, c00);
            c01 = mfma<BF16>(A0, B1, c01);
            c10 = mfma<BF16>(A1, B0, c10);
            c11 = mfma<BF16>(A1, B1, c11);
        }
        if constexpr (EXIT) {
            const float m1 = fmaxf(fmaxf(max16(c00), max16(c01)), fmaxf(max16(c10), max16(c11)));
            if (!__any(m1 >= t_out) && !diag) {
                exits++;
                continue;
            }
            for (int t = steps1; t < steps; t++) {
                const u32x4 A0 = a0[t * 64], A1 = a1[t * 64], B0 = b0[t * 64], B1 = b1[t * 64];
                c00 = mfma<BF16>(A0, B0, c00);
                c01 = mfma<BF16>(A0, B1, c01);
                c10 = mfma<BF16>(A1, B0, c10);
                c11 = mfma<BF16>(A1, B1, c11);
            }
        }
        const float m = fmaxf(fmaxf(max16(c00), max16(c01)), fmaxf(max16(c10), max16(c11)));
        if (__any(m >= t_out)) {
            const uint32_t col0 = J * 64 + (lane & 31);
            classify<SELF>(c00, I * 64, col0, nA, nB, diag, t_in, t_out, cnt, ctr, band, band_cap);
            classify<SELF>(c01, I * 64, col0 + 32, nA, nB, diag, t_in, t_out, cnt, ctr, band,
                           band_cap);
            classify<SELF>(c10, I * 64 + 32, col0, nA, nB, diag, t_in, t_out, cnt, ctr, band,
                           band_cap);
            classify<SELF>(c11, I * 64 + 32, col0 + 32, nA, nB, diag, t_in, t_out, cnt, ctr, band,
                           band_cap);
        }
    }
    if (lane == 0) {
        atomicAdd((unsigned long long *)&ctr[1], (unsigned long long)cnt);
        atomicAdd((unsigned long long *)&ctr[3], (unsigned long long)cand);
        atomicAdd((unsigned long long *)&ctr[7], (unsigned long long)exits);
    }
}

// FP64 fixed-order recheck of the queued band pairs (sorted indices).
__global__ __launch_bounds__(256) void band_recheck_kernel(const double *__restrict__ ra,
                                                           const double *__restrict__ rb,
                                                           int dpad, double eps2,
                                                           const uint64_t *__restrict__ band,
                                                           uint64_t nband,
                                                           uint64_t *__restrict__ ctr) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    if (k < nband) {
        const uint64_t e = band[k];
        const double *a = ra + (size_t)(uint32_t)e * dpad;
        const double *b = rb + (size_t)(uint32_t)(e >> 32) * dpad;
        double s = 0.0;
        int c = 0;
        for (; c < dpad; c++) {
            const double t = a[c] - b[c];
            s = __builtin_fma(t, t, s);
            if (s > eps2) break;
        }
        in = c == dpad && s <= eps2;
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

template <bool SELF, bool BF16, bool EXIT>
void mfma_launch(const u32x4 *a, const u32x4 *b, const uint64_t *pairs, uint64_t npairs,
                 uint32_t nA, uint32_t nB, const FilterParams &fp, uint64_t *counters,
                 uint64_t *band, uint64_t band_cap, int grid, int num_cus, hipStream_t s) {
    auto k = mfma_count_kernel<SELF, BF16, EXIT>;
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
    const bool ex = fp.steps1 < fp.steps;
#define HJ_MFMA_LAUNCH(S, B, E)                                                                \
    mfma_launch<S, B, E>(a, b, pairs, npairs, nA, nB, fp, counters, band, band_cap, grid,      \
                         num_cus, s)
    if (self) {
        if (fp.bf16) {
            if (ex) HJ_MFMA_LAUNCH(true, true, true); else HJ_MFMA_LAUNCH(true, true, false);
        } else {
            if (ex) HJ_MFMA_LAUNCH(true, false, true); else HJ_MFMA_LAUNCH(true, false, false);
        }
    } else {
        if (fp.bf16) {
            if (ex) HJ_MFMA_LAUNCH(false, true, true); else HJ_MFMA_LAUNCH(false, true, false);
        } else {
            if (ex) HJ_MFMA_LAUNCH(false, false, true); else HJ_MFMA_LAUNCH(false, false, false);
        }
    }
#undef HJ_MFMA_LAUNCH
    HJ_HIP(hipGetLastError());
}

void launch_band_recheck(const double *rowsA, const double *rowsB, int dpad, double eps2,
                         const uint64_t *band, uint64_t nband, uint64_t *counters,
                         hipStream_t s) {
    if (nband == 0) return;
    band_recheck_kernel<<<(unsigned)((nband + 255) / 256), 256, 0, s>>>(rowsA, rowsB, dpad, eps2,
                                                                        band, nband, counters);
    HJ_HIP(hipGetLastError());
}

}  // namespace hj
