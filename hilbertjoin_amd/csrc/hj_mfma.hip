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

// squared norm (FP64) of each point's rounded, scaled, centred coordinates
__global__ void op_norms_kernel(const double *__restrict__ rows, size_t n, int d, int dpad,
                                const double *__restrict__ centre, double scale, int bf16,
                                double *__restrict__ norms) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double s = 0.0;
    for (int k = 0; k < d; k++) {
        double v = round_op((rows[p * dpad + k] - centre[k]) * scale, bf16 != 0);
        s = __builtin_fma(v, v, s);
    }
    norms[p] = s;
}

struct PackParams {
    int d, dpad, ksteps, parts, bf16;
    double scale;
};

constexpr double NORM_UNIT = 32768.0;  // 2^15: the partner of each norm part
constexpr double PAD_PART = 32768.0;   // norm part of pad points: acc <= -2^30

// One wavefront per (tile, 32-point half, k-step): lane l holds point
// tile*64 + half*32 + l%32, dimensions 16 t + 8 (l/32) + j, j = 0..7 -- the
// A/B operand layout of v_mfma_f32_32x32x16.  The row-role array carries the
// point's norm parts in slots [d, d+parts) and -2^15 in [d+parts, d+2 parts);
// the column-role array the reverse, so row x column adds -|a|^2/2 - |b|^2/2.
__global__ __launch_bounds__(64) void frag_pack_kernel(const double *__restrict__ rows,
                                                       size_t n, const double *__restrict__ centre,
                                                       const double *__restrict__ norms,
                                                       PackParams pp, u32x4 *__restrict__ frow,
                                                       u32x4 *__restrict__ fcol) {
    const size_t blk = blockIdx.x;
    const int t = (int)(blk % pp.ksteps);
    const size_t th = blk / pp.ksteps;  // tile*2 + half
    const int lane = threadIdx.x;
    const size_t p = th * 32 + (lane & 31);
    const int k0 = t * 16 + 8 * (lane >> 5);
    const bool bf = pp.bf16 != 0;
    double part[3] = {0, 0, 0};
    if (p < n) {
        double v = norms[p] * 0.5 / NORM_UNIT;
        for (int q = 0; q < pp.parts; q++) {
            part[q] = round_op(v, bf);
            v -= part[q];
        }
    } else {
        part[0] = PAD_PART;
    }
    uint16_t hr[8], hc[8];
    for (int j = 0; j < 8; j++) {
        const int k = k0 + j;
        double vr = 0.0, vc = 0.0;
        if (k < pp.d) {
            if (p < n) vr = vc = (rows[p * pp.dpad + k] - centre[k]) * pp.scale;
        } else if (k < pp.d + pp.parts) {
            vr = part[k - pp.d];
            vc = -NORM_UNIT;
        } else if (k < pp.d + 2 * pp.parts) {
            vr = -NORM_UNIT;
            vc = part[k - pp.d - pp.parts];
        }
        hr[j] = bits_op(vr, bf);
        hc[j] = bits_op(vc, bf);
    }
    u32x4 r, c;
    for (int q = 0; q < 4; q++) {
        r[q] = (uint32_t)hr[2 * q] | ((uint32_t)hr[2 * q + 1] << 16);
        c[q] = (uint32_t)hc[2 * q] | ((uint32_t)hc[2 * q + 1] << 16);
    }
    const size_t o = (th * pp.ksteps + t) * 64 + lane;
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

template <bool SELF, bool BF16>
__global__ __launch_bounds__(256) void mfma_count_kernel(
    const u32x4 *__restrict__ fa, const u32x4 *__restrict__ fb, const uint64_t *__restrict__ pairs,
    uint64_t npairs, uint32_t nA, uint32_t nB, int ksteps, float t_in, float t_out,
    uint64_t *__restrict__ ctr, uint64_t *__restrict__ band, uint64_t band_cap) {
    const uint32_t lane = threadIdx.x & 63;
    uint64_t cnt = 0, cand = 0;
    WorkCursor cur(npairs);
    uint64_t w;
    while (cur.next(w)) {
        const uint64_t pr = pairs[w];
        const uint32_t I = (uint32_t)pr, J = (uint32_t)(pr >> 32);
        const bool diag = SELF && I == J;
        const uint32_t nvi = min(64u, nA - I * 64), nvj = min(64u, nB - J * 64);
        cand += diag ? (uint64_t)nvj * (nvj - 1) / 2 : (uint64_t)nvi * nvj;
        const u32x4 *a0 = fa + (size_t)(2 * I) * ksteps * 64 + lane;
        const u32x4 *a1 = a0 + (size_t)ksteps * 64;
        const u32x4 *b0 = fb + (size_t)(2 * J) * ksteps * 64 + lane;
        const u32x4 *b1 = b0 + (size_t)ksteps * 64;
        f32x16 c00 = {}, c01 = {}, c10 = {}, c11 = {};
        for (int t = 0; t < ksteps; t++) {
            const u32x4 A0 = a0[t * 64], A1 = a1[t * 64], B0 = b0[t * 64], B1 = b1[t * 64];
            c00 = mfma<BF16>(A0, B0, c00);
            c01 = mfma<BF16>(A0, B1, c01);
            c10 = mfma<BF16>(A1, B0, c10);
            c11 = mfma<BF16>(A1, B1, c11);
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

void launch_op_norms(const double *rows, size_t n, int d, int dpad, const double *centre,
                     double scale, bool bf16, double *norms, hipStream_t s) {
    if (n == 0) return;
    op_norms_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(rows, n, d, dpad, centre, scale,
                                                                bf16 ? 1 : 0, norms);
    HJ_HIP(hipGetLastError());
}

void launch_frag_pack(const double *rows, size_t n, int d, int dpad, const double *centre,
                      const double *norms, double scale, int ksteps, int parts, bool bf16,
                      void *frow, void *fcol, hipStream_t s) {
    size_t ntiles = (n + 63) / 64;
    if (ntiles == 0) return;
    PackParams pp{d, dpad, ksteps, parts, bf16 ? 1 : 0, scale};
    frag_pack_kernel<<<(unsigned)(ntiles * 2 * ksteps), 64, 0, s>>>(
        rows, n, centre, norms, pp, static_cast<u32x4 *>(frow), static_cast<u32x4 *>(fcol));
    HJ_HIP(hipGetLastError());
}

void launch_mfma_count(const void *fragA, const void *fragB, const uint64_t *pairs,
                       uint64_t npairs, uint32_t nA, uint32_t nB, const FilterParams &fp,
                       bool self, uint64_t *counters, uint64_t *band, uint64_t band_cap,
                       int grid, int num_cus, hipStream_t s) {
    if (npairs == 0) return;
    const u32x4 *a = static_cast<const u32x4 *>(fragA);
    const u32x4 *b = static_cast<const u32x4 *>(fragB);
#define HJ_MFMA_LAUNCH(S, B)                                                                   \
    do {                                                                                       \
        auto k = mfma_count_kernel<S, B>;                                                      \
        k<<<persistent_grid(k, 256, num_cus, grid), 256, 0, s>>>(                              \
            a, b, pairs, npairs, nA, nB, fp.ksteps, fp.t_in, fp.t_out, counters, band, band_cap); \
    } while (0)
    if (self) {
        if (fp.bf16)
            HJ_MFMA_LAUNCH(true, true);
        else
            HJ_MFMA_LAUNCH(true, false);
    } else {
        if (fp.bf16)
            HJ_MFMA_LAUNCH(false, true);
        else
            HJ_MFMA_LAUNCH(false, false);
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
