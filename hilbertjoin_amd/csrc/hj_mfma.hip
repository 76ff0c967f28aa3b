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
// in hj_join.cpp):
//     acc <  T_out        the pair is certainly farther than eps
//     acc >= T_in         the pair is certainly within eps
//     otherwise           the pair is queued and rechecked with the FP64
//                         fixed-order predicate
// so the count equals the FP64 oracle's (no false negatives or positives).
// Band pairs, and for pair lists the certain pairs, leave through
// per-wavefront LDS queues (hj_queue.h).
#include <algorithm>

#include "hj_device.h"
#include "hj_frames.h"
#include "hj_internal.h"
#include "hj_queue.h"
#include "hj_work.h"

namespace hj {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

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

// value of an FP16 / BF16 bit pattern, as a double
__device__ __forceinline__ double op_value(uint16_t h, bool bf16) {
    if (bf16) return (double)(float)__builtin_bit_cast(__bf16, h);
    return (double)(float)__builtin_bit_cast(_Float16, h);
}

// Row stride of a staged tile ([64][stride] 16-bit entries): 4 times an odd
// number, so a row is 8-byte aligned (two 8-byte loads read 8 consecutive
// dimensions) and the rows of 32 points fall on distinct LDS bank pairs.
__host__ __device__ constexpr int stage_stride(int d) { return 4 * (((d + 3) / 4) | 1); }
// 16-bit entries of the staged tile (a multiple of 4, so the FP64 data that
// follows is aligned)
__host__ __device__ constexpr size_t stage_halfs(int d) { return (size_t)64 * stage_stride(d); }

constexpr double NORM_UNIT = 32768.0;  // 2^15: the partner of each norm part
constexpr double PAD_PART = 32768.0;   // norm part of pad points: acc <= -2^30
constexpr int FILTER_PREP_MAX_D = 128;  // LDS-staged tiles up to this d

struct PackParams {
    int d, split, steps1, steps, parts, bf16;
    int target;  // the set is the target side of its segments
};

// Norm partial sums of a tile's points over the K chunks (grp-th share of
// the dimensions), from its rounded coordinates ybits(p, k) (16-bit patterns).
template <int WAVES, bool BF16, class YBits>
__device__ __forceinline__ void tile_norms(const PackParams &pp, uint32_t nv, YBits ybits,
                                           double (*snorm)[64][2]) {
    const int t = threadIdx.x, q = t & 63, grp = t >> 6;
    constexpr bool bf = BF16;
    double s0 = 0.0, s1 = 0.0;
    if ((uint32_t)q < nv) {
        for (int k = grp; k < pp.d; k += WAVES) {
            const double v = op_value(ybits(q, k), bf);
            if (k < pp.split)
                s0 = __builtin_fma(v, v, s0);
            else
                s1 = __builtin_fma(v, v, s1);
        }
    }
    snorm[grp][q][0] = s0;
    snorm[grp][q][1] = s1;
}

// Norm parts of the tile's points per K chunk as operand bit patterns
// (spart[chunk][point][part]; a pad point's first part is PAD_PART).
template <int WAVES, bool BF16>
__device__ __forceinline__ void tile_parts(const PackParams &pp, uint32_t nv,
                                           const double (*snorm)[64][2],
                                           uint16_t (*spart)[64][3]) {
    for (int i = threadIdx.x; i < 128; i += 64 * WAVES) {
        const int pt = i & 63, c = i >> 6;
        double part[3] = {0, 0, 0};
        if ((uint32_t)pt < nv) {
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < WAVES; w++) v += snorm[w][pt][c];
            v *= 0.5 / NORM_UNIT;
            for (int r = 0; r < pp.parts; r++) {
                part[r] = round_op(v, BF16);
                v -= part[r];
            }
        } else {
            part[0] = PAD_PART;
        }
#pragma unroll
        for (int r = 0; r < 3; r++) spart[c][pt][r] = bits_op(part[r], BF16);
    }
}

__device__ __forceinline__ u32x4 pack8(const uint16_t (&h)[8]) {
    u32x4 r;
#pragma unroll
    for (int w = 0; w < 4; w++) r[w] = (uint32_t)h[2 * w] | ((uint32_t)h[2 * w + 1] << 16);
    return r;
}

// The fragments of tile T (both roles; a null array skips its role), written
// in memory order: entry e of the tile is k-step e / 128, half (e / 64) % 2,
// lane e % 64.  yrow8(p, k) gives the rounded coordinates k .. k + 7 of point
// p packed as an entry (only called with k + 8 within the chunk's data), y(p,
// k) one of them; the norm parts come from spart (tile_parts).  Data slots
// hold the same value in both roles; of a chunk's 2 * parts norm slots the
// row role holds the parts and then -2^15, the column role the reverse.
template <int WAVES, bool BF16, class YRow8, class YBits>
__device__ __forceinline__ void tile_frags_write(const PackParams &pp, size_t T, YRow8 yrow8,
                                                 YBits ybits, const uint16_t (*spart)[64][3],
                                                 u32x4 *__restrict__ frow,
                                                 u32x4 *__restrict__ fcol) {
    const uint16_t neg = bits_op(-NORM_UNIT, BF16);
    const int n = 2 * pp.steps * 64;
    const size_t base = T * (size_t)n;
    for (int e = threadIdx.x; e < n; e += 64 * WAVES) {
        const int lane = e & 63, h = (e >> 6) & 1, st = e >> 7;
        const int pt = h * 32 + (lane & 31);
        const bool first = st < pp.steps1;
        const int dim0 = first ? 0 : pp.split;
        const int ndata = first ? pp.split : pp.d - pp.split;
        const int slot0 = (first ? st : st - pp.steps1) * 16 + 8 * (lane >> 5);
        u32x4 r, c;
        if (slot0 + 8 <= ndata) {
            r = c = yrow8(pt, dim0 + slot0);  // coordinate slots only
        } else {
            // first norm slot of the chunk: the last 2 * parts slots of its k-steps
            const int norm0 = 16 * (first ? pp.steps1 : pp.steps - pp.steps1) - 2 * pp.parts;
            const int ch = first ? 0 : 1;
            uint16_t hr[8], hc[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int qq = slot0 + j;
                uint16_t vr = 0, vc = 0;
                if (qq < ndata) {
                    vr = vc = ybits(pt, dim0 + qq);
                } else if (qq >= norm0 + pp.parts) {
                    vr = neg;
                    vc = spart[ch][pt][qq - norm0 - pp.parts];
                } else if (qq >= norm0) {
                    vr = spart[ch][pt][qq - norm0];
                    vc = neg;
                }
                hr[j] = vr;
                hc[j] = vc;
            }
            r = pack8(hr);
            c = pack8(hc);
        }
        if (frow) frow[base + e] = r;
        if (fcol) fcol[base + e] = c;
    }
}

// 8 consecutive staged coordinates (k, k + 1, ..., k + 7 of point p) as one
// fragment entry: two 8-byte, four 4-byte or eight 2-byte LDS loads by the
// alignment of k (k is wave-uniform in tile_frags_write: the first dimension
// of a K chunk plus a multiple of 8).
__device__ __forceinline__ u32x4 staged_row8(const uint16_t *sy, int stride, int p, int k) {
    const uint16_t *q = sy + p * stride + k;
    u32x4 r;
    if ((k & 3) == 0) {
        const uint2 a = *reinterpret_cast<const uint2 *>(q);
        const uint2 b = *reinterpret_cast<const uint2 *>(q + 4);
        r[0] = a.x;
        r[1] = a.y;
        r[2] = b.x;
        r[3] = b.y;
    } else if ((k & 1) == 0) {
#pragma unroll
        for (int w = 0; w < 4; w++) r[w] = *reinterpret_cast<const uint32_t *>(q + 2 * w);
    } else {
#pragma unroll
        for (int w = 0; w < 4; w++) r[w] = (uint32_t)q[2 * w] | ((uint32_t)q[2 * w + 1] << 16);
    }
    return r;
}

// Stages the rows of a tile (wavefront grp takes rows grp, grp + WAVES, ...:
// coordinate-contiguous loads, 16 rows in flight at once) and hands
// (point, dimension, coordinate) of its real points to emit.
template <int WAVES, class Emit>
__device__ __forceinline__ void stage_rows(const double *__restrict__ pts, const uint32_t *srow,
                                           int d, uint32_t nv, Emit emit) {
    const int q = threadIdx.x & 63, grp = threadIdx.x >> 6;
    for (int i0 = 0; i0 < 64 / WAVES; i0 += 16) {
        for (int k = q; k < d; k += 64) {
            double v[16];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int p = grp + WAVES * (i0 + i);
                v[i] = (uint32_t)p < nv ? pts[(size_t)srow[p] * d + k] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < 16; i++) emit(grp + WAVES * (i0 + i), k, v[i]);
        }
    }
}

// Operand fragments of one layout tile (one block of 256 threads per tile).
// The K axis is cut in two chunks: k-steps [0, steps1) hold dimensions
// [0, split), k-steps [steps1, steps) dimensions [split, d); each chunk ends
// with 2*parts norm slots of its dimensions, the last 2*parts slots of its
// last k-step (zeros between the data and the norm slots).  Row-role arrays
// carry the point's norm parts in the first `parts` of them and -2^15 in the
// next `parts`; column-role arrays the reverse, so a row x column product
// adds -|a_c|^2/2 - |b_c|^2/2 for chunk c, and the accumulator after chunk 1
// is -(partial distance)^2/2, after both chunks -(distance)^2/2.  (The norm
// slots of a chunk are therefore always in the upper 32 lanes of its last
// k-step, elements 8 - 2*parts .. 7.)
// Fragment entry (tile, k-step t, half h, lane l) holds point 32 h + l % 32 of
// the tile, K slots 16 t + 8 (l / 32) + j, j = 0..7 -- the A/B operand layout
// of v_mfma_f32_32x32x16.  The two halves of a k-step are adjacent (1 KiB
// apart), so the second one is an immediate offset from the first; a tile
// spans steps * 2 KiB.  The points come straight from the input rows
// through the EGO permutation; with STAGED the tile's rounded scaled
// coordinates are first computed into LDS (FP16 / BF16 bit patterns) from
// row-contiguous loads (d <= FILTER_PREP_MAX_D).  A null frow or fcol skips
// that role (a query side needs only rows, a target side only columns).
// STAGED with opflag: every real point's scaled coordinates are checked
// against the grid's bounds |z_k| <= zmax_k, and *opflag set when one is
// outside (or not a number).
template <bool STAGED, int WAVES, bool BF16>
__global__ __launch_bounds__(64 * WAVES) void frag_tile_kernel(
    const double *__restrict__ pts, const uint32_t *__restrict__ perm,
    const SegDev *__restrict__ segs, const uint32_t *__restrict__ tile_seg,
    const GridDev *__restrict__ grids, const double *__restrict__ centre_all, PackParams pp,
    u32x4 *__restrict__ frow, u32x4 *__restrict__ fcol, const double *__restrict__ zmax_all,
    uint32_t *__restrict__ opflag) {
    // STAGED: [64][stride] rounded scaled coordinates y of the tile as FP16 /
    // BF16 bit patterns, then the centre [d] and the bounds [d] in FP64
    extern __shared__ uint16_t sy[];
    __shared__ double snorm[WAVES][64][2];
    __shared__ uint16_t spart[2][64][3];
    __shared__ uint32_t srow[64];  // input row of each tile point
    const int t = threadIdx.x;
    const size_t T = blockIdx.x;
    const int d = pp.d, stride = stage_stride(d);
    constexpr bool bf = BF16;
    const SegDev &S = segs[tile_seg[T]];
    const GridDev g = grids[S.grid];
    const double *centre = centre_all + g.centre_off;
    const double scale = g.scale;
    uint64_t start;
    uint32_t nv;
    tile_span(S, (uint32_t)T, pp.target != 0, &start, &nv);
    if (t < 64) srow[t] = (uint32_t)t < nv ? perm[start + t] : 0u;
    if (STAGED) {
        double *sc = reinterpret_cast<double *>(sy + stage_halfs(d));
        double *szm = sc + d;
        const double *zmax = zmax_all + g.centre_off;
        for (int k = t; k < d; k += 64 * WAVES) {
            sc[k] = centre[k];
            if (opflag) szm[k] = zmax[k];
        }
        __syncthreads();
        bool miss = false;
        stage_rows<WAVES>(pts, srow, d, nv, [&](int p, int k, double v) {
            const bool real = (uint32_t)p < nv;
            const double z = (v - sc[k]) * scale;
            sy[p * stride + k] = real ? bits_op(z, bf) : (uint16_t)0;
            if (opflag && real && !(fabs(z) <= szm[k])) miss = true;
        });
        if (opflag) {
            if (__syncthreads_or(miss) && t == 0) atomicOr(opflag, 1u);
        }
    }
    __syncthreads();
    // rounded scaled coordinate k of tile point p (p < nv)
    auto ybits = [&](int p, int k) -> uint16_t {
        if (STAGED) return sy[p * stride + k];
        return bits_op((pts[(size_t)srow[p] * d + k] - centre[k]) * scale, bf);
    };
    tile_norms<WAVES, BF16>(pp, nv, ybits, snorm);
    __syncthreads();
    tile_parts<WAVES, BF16>(pp, nv, snorm, spart);
    __syncthreads();
    auto yrow8 = [&](int p, int k) -> u32x4 {
        if (STAGED) return staged_row8(sy, stride, p, k);
        uint16_t h[8];
#pragma unroll
        for (int j = 0; j < 8; j++) h[j] = ybits(p, k + j);
        return pack8(h);
    };
    tile_frags_write<WAVES, BF16>(pp, T, yrow8, ybits, spart, frow, fcol);
}

// Tile frames (hj_frames.h), tiles staged in LDS (d <= FRAME_MAX_D): the
// grid-frame fragments of every tile as frag_tile_kernel (frow, fcol), and,
// for a tile that takes its own frame, its fragments in that frame (frow2,
// fcol2; the other tiles' entries there are left unwritten: the filter reads
// their grid-frame fragments).  The anchor is the centre of the tile's bounding
// box in scaled coordinates z = (x - c) s, rounded to the grid's lattice; the
// tile takes its own frame when the bound sqrt(sum_k (w_k / 2 + |m_k -
// anchor_k|)^2) (w: box widths, m: box centres) on its radius around the
// anchor is at most FRAME_GAIN times its radius around the grid centre.  Per
// tile the kernel writes {rho, framed} (finfo; rho bounds the operand vectors'
// norm in the tile's frame), for a framed tile its anchor per K slot (fanch:
// zero off the data slots; not written for grid-frame tiles, whose anchor is
// 0), and sets *fflag for a framed tile.
// With opflag the points are checked as in frag_tile_kernel.
template <int WAVES, bool BF16>
__global__ __launch_bounds__(64 * WAVES) void frag_frames_kernel(
    const double *__restrict__ pts, const uint32_t *__restrict__ perm,
    const SegDev *__restrict__ segs, const uint32_t *__restrict__ tile_seg,
    const GridDev *__restrict__ grids, const double *__restrict__ centre_all, PackParams pp,
    u32x4 *__restrict__ frow, u32x4 *__restrict__ fcol, u32x4 *__restrict__ frow2,
    u32x4 *__restrict__ fcol2, float *__restrict__ fanch, float2 *__restrict__ finfo,
    uint32_t *__restrict__ fflag, const double *__restrict__ zmax_all,
    uint32_t *__restrict__ opflag) {
    // [64][stride] rounded coordinates (16-bit), then FP64: the centre [d],
    // the anchor [d], box bounds [WAVES][2][d], the bounds on |z| [d]
    extern __shared__ uint16_t sy[];
    __shared__ double snorm[WAVES][64][2];
    __shared__ uint16_t spart[2][64][3];
    __shared__ uint32_t srow[64];
    __shared__ int sframed;
    __shared__ float srho;
    const int t = threadIdx.x, q = t & 63, grp = t >> 6;
    const size_t T = blockIdx.x;
    const int d = pp.d, stride = stage_stride(d);
    constexpr bool bf = BF16;
    const SegDev &S = segs[tile_seg[T]];
    const GridDev g = grids[S.grid];
    const double *centre = centre_all + g.centre_off;
    const double scale = g.scale;
    double *sc = reinterpret_cast<double *>(sy + stage_halfs(d));
    double *sa = sc + d;
    double *sbox = sa + d;  // [WAVES][2][d]
    double *szm = sbox + 2 * WAVES * d;
    const double *zmax = zmax_all + g.centre_off;
    uint64_t start;
    uint32_t nv;
    tile_span(S, (uint32_t)T, pp.target != 0, &start, &nv);
    if (t < 64) srow[t] = (uint32_t)t < nv ? perm[start + t] : 0u;
    for (int k = t; k < d; k += 64 * WAVES) {
        sc[k] = centre[k];
        if (opflag) szm[k] = zmax[k];
    }
    __syncthreads();
    // grid frame, and the box of the thread's rows in its dimension (one
    // dimension per thread: d <= 64)
    double lo = INFINITY, hi = -INFINITY;
    bool miss = false;
    stage_rows<WAVES>(pts, srow, d, nv, [&](int p, int k, double v) {
        const bool real = (uint32_t)p < nv;
        const double z = (v - sc[k]) * scale;
        sy[p * stride + k] = real ? bits_op(z, bf) : (uint16_t)0;
        if (real) {
            lo = fmin(lo, z);
            hi = fmax(hi, z);
            if (opflag && !(fabs(z) <= szm[k])) miss = true;
        }
    });
    if (q < d) {
        sbox[(grp * 2) * d + q] = lo;
        sbox[(grp * 2 + 1) * d + q] = hi;
    }
    if (opflag) {
        if (__syncthreads_or(miss) && t == 0) atomicOr(opflag, 1u);
    } else {
        __syncthreads();
    }
    auto ybits = [&](int p, int k) -> uint16_t { return sy[p * stride + k]; };
    auto yrow8 = [&](int p, int k) -> u32x4 { return staged_row8(sy, stride, p, k); };
    tile_norms<WAVES, BF16>(pp, nv, ybits, snorm);
    __syncthreads();
    tile_parts<WAVES, BF16>(pp, nv, snorm, spart);
    __syncthreads();
    tile_frags_write<WAVES, BF16>(pp, T, yrow8, ybits, spart, frow, fcol);
    if (t < 64) {  // wave 0: the anchor (lane k) and the frame decision
        double rg = 0.0;  // lane t: |y|^2 of point t in the grid frame
        for (int w = 0; w < WAVES; w++) rg += snorm[w][t][0] + snorm[w][t][1];
        double bnd = 0.0;  // lane k: its term of the radius bound
        if (t < d) {
            double blo = INFINITY, bhi = -INFINITY;
            for (int w = 0; w < WAVES; w++) {
                blo = fmin(blo, sbox[(w * 2) * d + t]);
                bhi = fmax(bhi, sbox[(w * 2 + 1) * d + t]);
            }
            const double fq = (double)g.fr_q;  // a power of two
            const double amax = bf ? FRAME_ANCHOR_MAX_BF16 : FRAME_ANCHOR_MAX_F16;
            const double m = 0.5 * blo + 0.5 * bhi;
            const double a = fmin(fmax(rint(m / fq), -amax), amax) * fq;
            sa[t] = a;
            const double r = (0.5 * bhi - 0.5 * blo) + fabs(m - a);
            bnd = r * r;
        }
        for (int m = 32; m >= 1; m >>= 1) {
            rg = fmax(rg, __shfl_xor(rg, m));
            bnd += __shfl_xor(bnd, m);
        }
        if (t == 0) {
            // with a margin for the rounding of both sides of the comparison
            const bool framed = bnd * (1.0 + 0x1p-30) <= FRAME_GAIN * FRAME_GAIN * rg;
            sframed = framed ? 1 : 0;
            srho = (float)(sqrt(rg) * (1.0 + 0x1p-20));  // grid-frame rho
        }
    }
    __syncthreads();
    if (sframed == 0) {
        if (t == 0) finfo[T] = make_float2(srho, 0.0f);
        return;
    }
    for (int e = t; e < pp.steps * 16; e += 64 * WAVES) {
        const int st = e >> 4, slot = e & 15;
        const bool first = st < pp.steps1;
        const int local = (first ? st : st - pp.steps1) * 16 + slot;
        const int ndata = first ? pp.split : d - pp.split;
        const int dim = (first ? 0 : pp.split) + local;
        fanch[(T * pp.steps + st) * 16 + slot] = local < ndata ? (float)sa[dim] : 0.0f;
    }
    // own frame: the rows again (from L2), rounded around the anchor
    stage_rows<WAVES>(pts, srow, d, nv, [&](int p, int k, double v) {
        sy[p * stride + k] =
            (uint32_t)p < nv ? bits_op((v - sc[k]) * scale - sa[k], bf) : (uint16_t)0;
    });
    __syncthreads();
    tile_norms<WAVES, BF16>(pp, nv, ybits, snorm);
    __syncthreads();
    tile_parts<WAVES, BF16>(pp, nv, snorm, spart);
    __syncthreads();
    tile_frags_write<WAVES, BF16>(pp, T, yrow8, ybits, spart, frow2, fcol2);
    if (t < 64) {
        double ro = 0.0;
        for (int w = 0; w < WAVES; w++) ro += snorm[w][t][0] + snorm[w][t][1];
        for (int m = 32; m >= 1; m >>= 1) ro = fmax(ro, __shfl_xor(ro, m));
        if (t == 0) {
            finfo[T] = make_float2((float)(sqrt(ro) * (1.0 + 0x1p-20)), 1.0f);
            *fflag = 1u;  // every writer stores the same value
        }
    }
}

// True in every lane when p holds in some lane (the vote stays in scalar
// registers).
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// Fragments of tile T: tbytes = steps * 2 KiB per tile; a 32-bit tile index
// times a 32-bit stride (two scalar multiplies, no 64 x 64-bit product).
__device__ __forceinline__ const u32x4 *tile_frags(const u32x4 *f, uint32_t T, uint32_t tbytes) {
    return reinterpret_cast<const u32x4 *>(reinterpret_cast<const char *>(f) +
                                           (uint64_t)T * tbytes);
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

__device__ __forceinline__ uint64_t min64(uint64_t a, uint64_t b) { return a < b ? a : b; }

__device__ __forceinline__ float max3f(float a, float b, float c) {
    return fmaxf(fmaxf(a, b), c);  // one v_max3_f32
}

// Maximum of m and the 32 values of two accumulators: a tree of 16
// three-input maxima (the accumulators are read straight from VGPRs).
__device__ __forceinline__ float max32(const f32x16 &a, const f32x16 &b, float m) {
    const float l0 = max3f(a[0], a[1], a[2]), l1 = max3f(a[3], a[4], a[5]);
    const float l2 = max3f(a[6], a[7], a[8]), l3 = max3f(a[9], a[10], a[11]);
    const float l4 = max3f(a[12], a[13], a[14]), l5 = max3f(a[15], b[0], b[1]);
    const float l6 = max3f(b[2], b[3], b[4]), l7 = max3f(b[5], b[6], b[7]);
    const float l8 = max3f(b[8], b[9], b[10]), l9 = max3f(b[11], b[12], b[13]);
    const float k0 = max3f(l0, l1, l2), k1 = max3f(l3, l4, l5), k2 = max3f(l6, l7, l8);
    const float k3 = max3f(l9, b[14], b[15]);
    return max3f(max3f(k0, k1, k2), k3, m);
}

// Queues the pairs of a 32 x 32 block selected by each lane's accumulator
// mask (bit r: accumulator r, row rbase + (r & 3) + 8 (r >> 2)), one
// accumulator index at a time: a loop, so the queue code appears once.
template <uint32_t NQ, uint32_t Q>
__device__ __forceinline__ void push_rows(QueueLds<NQ> &qs, uint32_t mask, uint32_t row0,
                                          uint32_t col) {
    const uint32_t rbase = row0 + 4 * ((threadIdx.x & 63) >> 5);
#pragma unroll 1
    for (uint32_t r = 0; r < 16; r++) {
        const uint32_t row = rbase + (r & 3) + 8 * (r >> 2);
        queue_push<NQ, Q>(qs, (mask >> r) & 1u, (uint64_t)row | ((uint64_t)col << 32));
    }
}

template <bool SELF, bool EMIT, uint32_t NQ>
__device__ __forceinline__ void classify(const f32x16 &acc, uint32_t row0, uint32_t col,
                                         uint32_t nA, uint32_t nB, bool diag, float t_in,
                                         float t_out, uint32_t &cnt, QueueLds<NQ> &qs) {
    const uint32_t rb = row0 + 4 * ((threadIdx.x & 63) >> 5);
    const bool vcol = col < nB;
    uint32_t inm = 0, bandm = 0;  // bit r: accumulator r (shifted in from r = 15 down)
#pragma unroll
    for (int r = 15; r >= 0; r--) {
        const uint32_t row = rb + (r & 3) + 8 * (r >> 2);
        const bool valid = vcol && row < nA && (!SELF || !diag || row < col);
        const float v = acc[r];
        inm = (inm << 1) | (valid && v >= t_in ? 1u : 0u);
        bandm = (bandm << 1) | (valid && v >= t_out && v < t_in ? 1u : 0u);
    }
    cnt += (uint32_t)__popc(inm);
    if (wave_any(bandm != 0)) push_rows<NQ, 0>(qs, bandm, row0, col);
    if constexpr (EMIT)
        if (wave_any(inm != 0)) push_rows<NQ, NQ - 1>(qs, inm, row0, col);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_u64(v, m);
    return v;
}


// Maximum of m and the 64 x 64 accumulators of one tile pair after one
// k-step (row fragments A0, A1, column fragments B0, B1).
template <bool BF16>
__device__ __forceinline__ float pair_max(const u32x4 &A0, const u32x4 &A1, const u32x4 &B0,
                                          const u32x4 &B1, float m) {
    const f32x16 c00 = mfma<BF16>(A0, B0, f32x16{});
    const f32x16 c01 = mfma<BF16>(A0, B1, f32x16{});
    const f32x16 c10 = mfma<BF16>(A1, B0, f32x16{});
    const f32x16 c11 = mfma<BF16>(A1, B1, f32x16{});
    return max32(c10, c11, max32(c00, c01, m));
}

__device__ __forceinline__ float max16(const f32x16 &acc) {
    float m = max3f(acc[0], acc[1], acc[2]);
#pragma unroll
    for (int r = 3; r < 15; r += 2) m = max3f(m, acc[r], acc[r + 1]);
    return fmaxf(m, acc[15]);
}

// The slow path of one tile pair, one 32 x 32 block at a time with a single
// accumulator: the first K chunk (k-steps [0, steps1)); a block whose partial
// distances are all certainly beyond eps is done (a partial distance never
// exceeds the full one); otherwise the rest of K, and a block with some
// accumulator >= t_out is classified (diagonal tiles and blocks where some
// pair may be within eps).
template <bool SELF, bool BF16, bool EMIT, uint32_t NQ>
__device__ __forceinline__ uint32_t full_tile_pair(
    const u32x4 *__restrict__ fa, const u32x4 *__restrict__ fb, uint32_t I, uint32_t J,
    int steps1, int steps, uint32_t nA, uint32_t nB, float t_in, float t_out, QueueLds<NQ> &qs,
    uint32_t &mfmas) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t tbytes = (uint32_t)steps * 2048u;
    const bool diag = SELF && I == J;
    uint32_t c = 0;
#pragma unroll 1
    for (int blk = 0; blk < 4; blk++) {
        const uint32_t bi = blk >> 1, bj = blk & 1;
        const u32x4 *a = tile_frags(fa, I, tbytes) + bi * 64;  // wave-uniform bases
        const u32x4 *b = tile_frags(fb, J, tbytes) + bj * 64;
        f32x16 acc = {};
        // the next k-step's operands are loaded while the current one multiplies
        u32x4 at = a[lane], bt = b[lane];
        for (int t = 0; t < steps; t++) {
            u32x4 an = at, bn = bt;
            if (t + 1 < steps) {
                an = a[(t + 1) * 128 + lane];
                bn = b[(t + 1) * 128 + lane];
            }
            acc = mfma<BF16>(at, bt, acc);
            mfmas++;
            at = an;
            bt = bn;
            if (t + 1 == steps1 && steps1 < steps && !wave_any(max16(acc) >= t_out)) break;
        }
        if (wave_any(max16(acc) >= t_out))
            classify<SELF, EMIT, NQ>(acc, I * 64 + bi * 32, J * 64 + bj * 32 + (lane & 31),
                                     nA, nB, diag, t_in, t_out, c, qs);
    }
    return c;
}

// ---- tile frames (hj_frames.h) ----------------------------------------------
// This lane's 8 anchor slots of tile T, k-step t (slots 8 (lane / 32) + j).
__device__ __forceinline__ void anchor_slots(const float *__restrict__ fanch, uint32_t T,
                                             int steps, int t, uint32_t lane, float (&a)[8]) {
    const float4 *p = reinterpret_cast<const float4 *>(
        fanch + ((size_t)T * (uint32_t)steps + (uint32_t)t) * 16 + 8 * (lane >> 5));
    const float4 x = p[0], y = p[1];
    a[0] = x.x, a[1] = x.y, a[2] = x.z, a[3] = x.w;
    a[4] = y.x, a[5] = y.y, a[6] = y.z, a[7] = y.w;
}

// Anchor differences in the operand type: exact (lattice multiples of at most
// 11 resp. 8 significant bits), so truncating conversions are exact too.
template <bool BF16> __device__ __forceinline__ u32x4 pack_delta(const float (&dl)[8]) {
    u32x4 r;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        if constexpr (BF16) {
            const float d0 = dl[2 * w], d1 = dl[2 * w + 1];
            r[w] = (__builtin_bit_cast(uint32_t, d0) >> 16) |
                   (__builtin_bit_cast(uint32_t, d1) & 0xFFFF0000u);
        } else {
            r[w] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(dl[2 * w], dl[2 * w + 1]));
        }
    }
    return r;
}

// acc + the dot product of this lane's 8 operand slots with 8 delta slots
// (products exact in FP32, FP32 accumulation).  Vector elements are copied
// to scalars before __builtin_bit_cast: hipcc 7.2 reads element 0 for every
// element of a bit-cast vector subscript.
template <bool BF16> __device__ __forceinline__ float dot8(u32x4 a, u32x4 dl, float acc) {
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint32_t aw = a[w], dw = dl[w];
        if constexpr (BF16)
            acc = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, aw),
                                                  __builtin_bit_cast(bf16x2, dw), acc, false);
        else
            acc = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2v, aw),
                                         __builtin_bit_cast(half2v, dw), acc, false);
    }
    return acc;
}

// Sum of the values of lanes l and l ^ 32 (the two slot halves of one point),
// in every lane.
__device__ __forceinline__ float half_sum(float x) {
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    const auto pr = __builtin_amdgcn_permlane32_swap(b, b, false, false);
    const uint32_t p0 = pr[0], p1 = pr[1];
    return __builtin_bit_cast(float, p0) + __builtin_bit_cast(float, p1);
}

__device__ __forceinline__ float lo_f(uint32_t w, bool bf) {
    const uint16_t h = (uint16_t)(w & 0xFFFFu);
    return bf ? (float)__builtin_bit_cast(__bf16, h) : (float)__builtin_bit_cast(_Float16, h);
}
__device__ __forceinline__ float hi_f(uint32_t w, bool bf) { return lo_f(w >> 16, bf); }
__device__ __forceinline__ uint32_t op_bits(float v, bool bf) {
    return bf ? (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)v)
              : (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
}

// Replaces the norm parts of one role in a chunk's last k-step (upper lanes:
// elements 8 - 2p .. 8 - p - 1 for the row role, 8 - p .. 7 for the column
// role) by the parts of (their FP32 sum + corr / 2^15), split again in FP32.
// (The sum is exact unless a low part is subnormal; the band allows 2^-21 of
// the norms for this sum, the correction and the split.)
template <bool BF16, bool ROW>
__device__ __forceinline__ u32x4 adjust_norm(u32x4 f, float corr) {
    if ((threadIdx.x & 63) < 32) return f;
    constexpr float inv = 1.0f / 32768.0f;
    if constexpr (!BF16) {
        const int w = ROW ? 2 : 3;  // elements 4, 5 or 6, 7
        const float v = __builtin_fmaf(corr, inv, lo_f(f[w], false) + hi_f(f[w], false));
        const float q0 = (float)(_Float16)v;
        f[w] = op_bits(v, false) | (op_bits(v - q0, false) << 16);
    } else if constexpr (ROW) {  // elements 2, 3, 4
        const float v = __builtin_fmaf(
            corr, inv, (lo_f(f[1], true) + hi_f(f[1], true)) + lo_f(f[2], true));
        const float q0 = (float)(__bf16)v, r1 = v - q0, q1 = (float)(__bf16)r1;
        f[1] = op_bits(v, true) | (op_bits(r1, true) << 16);
        f[2] = (f[2] & 0xFFFF0000u) | op_bits(r1 - q1, true);
    } else {  // elements 5, 6, 7
        const float v = __builtin_fmaf(
            corr, inv, (hi_f(f[2], true) + lo_f(f[3], true)) + hi_f(f[3], true));
        const float q0 = (float)(__bf16)v, r1 = v - q0, q1 = (float)(__bf16)r1;
        f[2] = (f[2] & 0xFFFFu) | (op_bits(v, true) << 16);
        f[3] = op_bits(r1, true) | (op_bits(r1 - q1, true) << 16);
    }
    return f;
}

// Frame data of one filter launch.  Framed tiles have their fragments in the
// own-frame arrays fa2 / fb2 (the other tiles' entries there are not
// written: their grid-frame fragments in fa / fb are the same) and their
// anchor slots in qanch / tanch (not written either for the others: their
// anchor is 0).
struct FrameArgsDev {
    const u32x4 *fa2, *fb2;       // own-frame fragments (row role, column role)
    const float2 *qinfo, *tinfo;  // per tile {rho, framed}
    const float *qanch, *tanch;   // per tile anchor slots
    const uint32_t *qflag, *tflag;  // per side: some tile is framed
    int fast;                     // the fast path in the tiles' own frames too
};

// This lane's anchor slots of tile T at k-step t: zero for a grid-frame tile.
__device__ __forceinline__ void tile_anchor(const float *__restrict__ fanch, bool framed,
                                            uint32_t T, int steps, int t, uint32_t lane,
                                            float (&a)[8]) {
    if (framed) {
        anchor_slots(fanch, T, steps, t, lane, a);
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = 0.0f;
    }
}

// The slow path of a tile pair with tile frames: as full_tile_pair, each tile
// in its own frame where it has one, with the corrections r, c of both K
// chunks computed first (one pass over the pair's operands) and applied in
// each chunk's last k-step, and the pair's own thresholds: tout1 for the
// early exit after the first chunk (shift |delta_1|^2 / 2), t_in, t_out over
// all of K.
template <bool SELF, bool BF16, bool EMIT, uint32_t NQ>
__device__ __forceinline__ uint32_t full_tile_pair_fr(
    const u32x4 *__restrict__ fa, const u32x4 *__restrict__ fb, const FrameArgsDev &fr,
    uint32_t I, uint32_t J, int steps1, int steps, uint32_t nA, uint32_t nB,
    const FrameConst &fc, QueueLds<NQ> &qs, uint32_t &mfmas) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t tbytes = (uint32_t)steps * 2048u;
    const bool diag = SELF && I == J;
    const float2 fi = fr.qinfo[I], fj = fr.tinfo[J];
    const bool fri = fi.y != 0.0f, frj = fj.y != 0.0f;
    // [chunk][half]
    float rc[2][2] = {{0.f, 0.f}, {0.f, 0.f}}, cc[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
    float dd[2] = {0.f, 0.f};
    const bool corr = (fri || frj) && !diag;
    const u32x4 *a = tile_frags(fri ? fr.fa2 : fa, I, tbytes);
    const u32x4 *b = tile_frags(frj ? fr.fb2 : fb, J, tbytes);
    if (corr) {
#pragma unroll 1
        for (int t = 0; t < steps; t++) {
            const int ch = t < steps1 ? 0 : 1;
            float ai[8], aj[8], dl[8];
            tile_anchor(fr.qanch, fri, I, steps, t, lane, ai);
            tile_anchor(fr.tanch, frj, J, steps, t, lane, aj);
            float s = dd[ch];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                dl[k] = aj[k] - ai[k];
                s = __builtin_fmaf(dl[k], dl[k], s);
            }
            dd[ch] = s;
            const u32x4 dp = pack_delta<BF16>(dl);
            const u32x4 a0 = a[t * 128 + lane], a1 = a[t * 128 + 64 + lane];
            const u32x4 b0 = b[t * 128 + lane], b1 = b[t * 128 + 64 + lane];
            rc[ch][0] = dot8<BF16>(a0, dp, rc[ch][0]);
            rc[ch][1] = dot8<BF16>(a1, dp, rc[ch][1]);
            cc[ch][0] = dot8<BF16>(b0, dp, cc[ch][0]);
            cc[ch][1] = dot8<BF16>(b1, dp, cc[ch][1]);
        }
#pragma unroll
        for (int ch = 0; ch < 2; ch++) {
            dd[ch] = half_sum(dd[ch]);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                rc[ch][h] = half_sum(rc[ch][h]);
                cc[ch][h] = half_sum(cc[ch][h]);
            }
        }
    }
    float tin1, tout1, t_in, t_out;
    frame_thresholds(fc, fi.x, fj.x, dd[0], dd[0], &tin1, &tout1);
    frame_thresholds(fc, fi.x, fj.x, dd[0] + dd[1], dd[0] + dd[1], &t_in, &t_out);
    if (steps1 == steps) tout1 = t_out;
    uint32_t c = 0;
#pragma unroll 1
    for (int blk = 0; blk < 4; blk++) {
        const uint32_t bi = blk >> 1, bj = blk & 1;
        const u32x4 *ab = a + bi * 64;  // wave-uniform bases
        const u32x4 *bb = b + bj * 64;
        const float r1 = bi ? rc[0][1] : rc[0][0], c1 = bj ? cc[0][1] : cc[0][0];
        const float r2 = bi ? rc[1][1] : rc[1][0], c2 = bj ? cc[1][1] : cc[1][0];
        f32x16 acc = {};
        u32x4 at = ab[lane], bt = bb[lane];
        for (int t = 0; t < steps; t++) {
            u32x4 an = at, bn = bt;
            if (t + 1 < steps) {
                an = ab[(t + 1) * 128 + lane];
                bn = bb[(t + 1) * 128 + lane];
            }
            if (corr && t == steps1 - 1) {
                at = adjust_norm<BF16, true>(at, -r1);
                bt = adjust_norm<BF16, false>(bt, c1);
            } else if (corr && t == steps - 1) {
                at = adjust_norm<BF16, true>(at, -r2);
                bt = adjust_norm<BF16, false>(bt, c2);
            }
            acc = mfma<BF16>(at, bt, acc);
            mfmas++;
            at = an;
            bt = bn;
            if (t + 1 == steps1 && steps1 < steps && !wave_any(max16(acc) >= tout1)) break;
        }
        if (wave_any(max16(acc) >= t_out))
            classify<SELF, EMIT, NQ>(acc, I * 64 + bi * 32, J * 64 + bj * 32 + (lane & 31),
                                     nA, nB, diag, t_in, t_out, c, qs);
    }
    return c;
}

// One wavefront per 64 x 64 tile pair.  Fast path: the first K chunk
// (k-steps [0, steps1); all of K when steps1 == steps) gives, in four 32 x 32
// accumulators, -(partial distance)^2 / 2 for all 4096 pairs; when every pair
// is certainly farther than eps (a partial distance never exceeds the full
// one) the tile pair is done.  Diagonal tiles and blocks where some pair may
// be within eps go to full_tile_pair.  The operand fragments of a whole chunk
// of tile pairs are loaded before its first MFMA, and the schedule
// (XcdCursor) keeps each XCD on one narrow window of its segment of the list.
// CHUNK: tile pairs per chunk of the schedule; the column fragments of all
// pairs of a chunk are loaded before the first MFMA of the chunk, so their
// memory latency overlaps.
// FRM: tile frames (hj_frames.h; steps1 == 1).  Where they are possible the
// host launches the FRM == 0 kernel and then the FRM == 1 or 2 one, and of the
// two the one that does not apply returns at once (some tile of either side
// took its own frame -- the sides' flags, read at the start -- or none), so
// the choice needs no read-back and no kernel carries another's registers.
// FRM >= 1: the slow path tests each tile pair in the tiles' own frames
// against the pair's own thresholds (from the tiles' radii and the anchors'
// difference), after correcting the norm slots of its operands when a tile
// is framed (full_tile_pair_fr); FRM == 2 does so in the fast path too (for
// wide grid-frame bands), FRM == 1 keeps the grid frame there.
template <bool SELF, bool BF16, uint32_t CHUNK, bool EMIT, int FRM, bool UNI>
__global__ __launch_bounds__(256) void mfma_count_kernel(
    const u32x4 *__restrict__ fa, const u32x4 *__restrict__ fb,
    const uint64_t *__restrict__ pairs, uint64_t npairs, const SegDev *__restrict__ segs,
    const uint32_t *__restrict__ tile_seg, int steps1, int steps, uint64_t *__restrict__ ctr,
    uint64_t *__restrict__ seg_counts, uint32_t *__restrict__ sched, FilterOut out,
    FrameArgsDev fr, const uint64_t *__restrict__ npairs_dev, const uint32_t *__restrict__ opflag) {
    const uint32_t lane = threadIdx.x & 63;
    // tile frames in use (wave-uniform): in the slow path, and in the fast one
    const bool framed = fr.qflag && (*fr.qflag | *fr.tflag) != 0;
    if (FRM == 0 ? framed : !framed) return;  // the other launch joins
    constexpr bool FAST = FRM == 2;
    static_assert(!(UNI && FAST), "one-row chunks are tested in the grid frame");
    uint64_t cnt = 0, segcnt = 0;
    uint32_t exits = 0, fast = 0, slow_mfmas = 0;
    uint32_t cur_seg = 0xFFFFFFFFu;
    // the list's length from the device where it was not read back (never
    // past npairs, the entries written); none when a point lay outside the
    // ranges the operands were scaled for (the host runs the join again)
    const uint64_t nlist =
        opflag && *opflag ? 0 : (npairs_dev ? min64(*npairs_dev, npairs) : npairs);
    XcdCursor<CHUNK, UNI ? 8 : XCD_FETCH> cur(nlist, sched);
    constexpr uint32_t NQ = EMIT ? 2 : 1;
    __shared__ QueueLds<NQ> qs;  // band pairs at ctr[4]; emitted pairs at ctr[6]
    queue_clear_fills(qs);
    if (threadIdx.x == 0) {
        queue_setup(qs, 0, out.band, out.band_cap, &ctr[4]);
        if constexpr (EMIT) queue_setup(qs, NQ - 1, out.pairs, out.pairs_cap, &ctr[6]);
        // which frames the launch used (every writer stores the same value)
        if (blockIdx.x == 0) ctr[5] = (uint64_t)FRM;
    }
    __syncthreads();
    const uint32_t tbytes = (uint32_t)steps * 2048u;
    uint64_t w0;
    uint32_t len;
    uint32_t curI = 0xFFFFFFFFu;
    u32x4 A0c = {}, A1c = {};  // k-step-0 row fragments of row tile curI
    float t_out = 0.f;  // threshold of row tile curI's segment
    FrameConst fcI{};   // frames in the fast path: band constants of curI's segment
    float2 fiI{};       //   {rho, framed} of row tile curI
    float ancI[8];      //   its anchor slots (k-step 0)
    uint64_t pr[CHUNK];
    u32x4 B0[CHUNK], B1[CHUNK];
    // one scalar load of a whole chunk: the list is padded by WORK_LIST_PAD
    // entries, and entries past n repeat the first pair
    auto load_pairs = [&](uint64_t w, uint32_t n, uint64_t(&p)[CHUNK]) {
#pragma unroll
        for (uint32_t q = 0; q < CHUNK; q++) p[q] = pairs[w + q];
#pragma unroll
        for (uint32_t q = 1; q < CHUNK; q++) p[q] = q < n ? p[q] : p[0];
    };
    // the chunk pr[0, len): its column fragments, the fast path, the slow path
    auto run_chunk = [&](uint32_t len) {
#pragma unroll
        for (uint32_t q = 0; q < CHUNK; q++) {
            // wave-uniform tile base + per-lane offset (SGPR-based loads); a
            // framed column tile's own-frame fragments in the fast path
            const uint32_t J = (uint32_t)(pr[q] >> 32);
            const u32x4 *fbq = FAST && fr.tinfo[J].y != 0.0f ? fr.fb2 : fb;
            const u32x4 *b = tile_frags(fbq, J, tbytes);
            B0[q] = b[lane];
            B1[q] = b[64 + lane];
        }
        uint32_t slow = 0;
        bool uniform = false;
        if constexpr (UNI) {
            // a chunk on one row tile (the list groups a row tile's pairs, so
            // its first and last pair share it) and one k-step: one vote for
            // the whole chunk, per pair only when some pair may be near (a
            // diagonal pair always is; its slow path masks the diagonal)
            const uint32_t I0 = (uint32_t)pr[0];
            uniform = steps1 == 1 && (uint32_t)pr[CHUNK - 1] == I0;
            if (uniform) {
                if (I0 != curI) {
                    t_out = segs[tile_seg[I0]].t_out;
                    const u32x4 *aI = tile_frags(fa, I0, tbytes);  // wave-uniform
                    A0c = aI[lane];
                    A1c = aI[64 + lane];
                    curI = I0;
                }
                float mq[CHUNK], m = -INFINITY;
#pragma unroll
                for (uint32_t q = 0; q < CHUNK; q++) {
                    mq[q] = pair_max<BF16>(A0c, A1c, B0[q], B1[q], -INFINITY);
                    m = fmaxf(m, mq[q]);
                }
                if (wave_any(m >= t_out)) {
#pragma unroll
                    for (uint32_t q = 0; q < CHUNK; q++)
                        if (q < len) slow |= (uint32_t)wave_any(mq[q] >= t_out) << q;
                }
                exits += len - (uint32_t)__popc(slow);
                fast += len;
            }
        }
        if (!uniform)
#pragma unroll
        for (uint32_t q = 0; q < CHUNK; q++) {
            if (q >= len) break;
            const uint32_t I = (uint32_t)pr[q], J = (uint32_t)(pr[q] >> 32);
            if (I != curI) {  // consecutive pairs in row order share the row tile
                const SegDev &S = segs[tile_seg[I]];
                t_out = S.t_out;
                const u32x4 *faI = fa;
                if constexpr (FAST) {
                    fcI = S.fc;
                    fiI = fr.qinfo[I];
                    tile_anchor(fr.qanch, fiI.y != 0.0f, I, steps, 0, lane, ancI);
                    if (fiI.y != 0.0f) faI = fr.fa2;
                }
                const u32x4 *aI = tile_frags(faI, I, tbytes);  // wave-uniform
                A0c = aI[lane];
                A1c = aI[64 + lane];
                curI = I;
            }
            const bool diag = SELF && I == J;
            if (diag) {
                slow |= 1u << q;
                continue;
            }
            float m = -INFINITY;
            float tq = t_out;
            if constexpr (FAST) {
                // one k-step: correct the norm slots of the pair's operands
                // when a tile has its own frame, then the pair's threshold
                const float2 fj = fr.tinfo[J];
                u32x4 A0 = A0c, A1 = A1c, Bq0 = B0[q], Bq1 = B1[q];
                float dd = 0.f;
                if (fiI.y != 0.f || fj.y != 0.f) {
                    float aj[8], dl[8];
                    tile_anchor(fr.tanch, fj.y != 0.f, J, steps, 0, lane, aj);
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        dl[k] = aj[k] - ancI[k];
                        dd = __builtin_fmaf(dl[k], dl[k], dd);
                    }
                    dd = half_sum(dd);
                    const u32x4 dp = pack_delta<BF16>(dl);
                    A0 = adjust_norm<BF16, true>(A0c, -half_sum(dot8<BF16>(A0c, dp, 0.f)));
                    A1 = adjust_norm<BF16, true>(A1c, -half_sum(dot8<BF16>(A1c, dp, 0.f)));
                    Bq0 = adjust_norm<BF16, false>(Bq0, half_sum(dot8<BF16>(Bq0, dp, 0.f)));
                    Bq1 = adjust_norm<BF16, false>(Bq1, half_sum(dot8<BF16>(Bq1, dp, 0.f)));
                }
                float tin_unused;
                frame_thresholds(fcI, fiI.x, fj.x, dd, dd, &tin_unused, &tq);
#pragma unroll
                for (int bi = 0; bi < 2; bi++) {
                    f32x16 c0 = {}, c1 = {};
                    const u32x4 Ab = bi ? A1 : A0;
                    c0 = mfma<BF16>(Ab, Bq0, c0);
                    c1 = mfma<BF16>(Ab, Bq1, c1);
                    m = max32(c0, c1, m);
                }
            } else {
                const u32x4 *aI = tile_frags(fa, I, tbytes);  // wave-uniform
                const u32x4 *bJ = tile_frags(fb, J, tbytes);
#pragma unroll
                for (int bi = 0; bi < 2; bi++) {
                    f32x16 c0 = {}, c1 = {};
                    const u32x4 Ab = bi ? A1c : A0c;
                    c0 = mfma<BF16>(Ab, B0[q], c0);
                    c1 = mfma<BF16>(Ab, B1[q], c1);
                    for (int t = 1; t < steps1; t++) {
                        const u32x4 At = aI[t * 128 + bi * 64 + lane], Bt0 = bJ[t * 128 + lane],
                                    Bt1 = bJ[t * 128 + 64 + lane];
                        c0 = mfma<BF16>(At, Bt0, c0);
                        c1 = mfma<BF16>(At, Bt1, c1);
                    }
                    m = max32(c0, c1, m);
                }
            }
            const bool near = wave_any(m >= tq);
            slow |= (uint32_t)near << q;
            exits += near ? 0u : 1u;
            fast++;
        }
        while (slow) {
            const uint32_t q = __builtin_ctz(slow);
            slow &= slow - 1;
            const uint32_t I = (uint32_t)pr[q], J = (uint32_t)(pr[q] >> 32);
            const uint32_t sg = tile_seg[I];
            const SegDev &S = segs[sg];
            const uint32_t nvi = tile_points(S.qn - 64ull * (I - S.qtile0));
            const uint32_t nvj = tile_points(S.tn - 64ull * (J - S.ttile0));
            uint32_t c;
            if constexpr (FRM != 0)
                c = full_tile_pair_fr<SELF, BF16, EMIT, NQ>(fa, fb, fr, I, J, steps1, steps,
                                                            I * 64 + nvi, J * 64 + nvj, S.fc, qs,
                                                            slow_mfmas);
            else
                c = full_tile_pair<SELF, BF16, EMIT, NQ>(fa, fb, I, J, steps1, steps,
                                                         I * 64 + nvi, J * 64 + nvj, S.t_in,
                                                         S.t_out, qs, slow_mfmas);
            const uint64_t cw = wave_sum_u64(c);
            cnt += cw;
            if (sg != cur_seg) {
                if (cur_seg != 0xFFFFFFFFu && segcnt && lane == 0)
                    atomicAdd((unsigned long long *)&seg_counts[cur_seg],
                              (unsigned long long)segcnt);
                cur_seg = sg;
                segcnt = 0;
            }
            segcnt += cw;
        }
    };
    if constexpr (UNI) {
        // the next chunk's pairs are loaded while this one runs
        uint64_t prn[CHUNK];
        bool have = cur.next_chunk(w0, len);
        if (have) load_pairs(w0, len, pr);
        while (have) {
            uint64_t w0n = 0;
            uint32_t lenn = 0;
            const bool haven = cur.next_chunk(w0n, lenn);
            load_pairs(haven ? w0n : w0, haven ? lenn : len, prn);  // the last: itself
            run_chunk(len);
#pragma unroll
            for (uint32_t q = 0; q < CHUNK; q++) pr[q] = prn[q];
            w0 = haven ? w0n : w0;
            len = lenn;
            have = haven;
        }
    } else {
        while (cur.next_chunk(w0, len)) {
            load_pairs(w0, len, pr);
            run_chunk(len);
        }
    }
    queue_drain<NQ, 0>(qs);
    if constexpr (EMIT) queue_drain<NQ, NQ - 1>(qs);
    if (cur_seg != 0xFFFFFFFFu && segcnt && lane == 0)
        atomicAdd((unsigned long long *)&seg_counts[cur_seg], (unsigned long long)segcnt);
    if (lane == 0) {
        atomicAdd((unsigned long long *)&ctr[1], (unsigned long long)cnt);
        atomicAdd((unsigned long long *)&ctr[7], (unsigned long long)exits);
        // MFMAs issued: four per k-step of the first chunk for every tile pair
        // of the fast path, one per k-step and block in the slow path
        atomicAdd((unsigned long long *)&ctr[0],
                  (unsigned long long)fast * 4ull * (unsigned)steps1 + slow_mfmas);
    }
}

// FP64 fixed-order recheck of the queued band pairs (layout positions, read
// through the segments and the EGO permutations from the original rows).
// Each lane evaluates its pair in the oracle's order (one FMA per dimension,
// ascending); the rows come in eight dimensions at a time, all loads of a
// group in flight together (16-byte loads when V2: even d, aligned rows),
// and the early exit is taken between groups -- the running sum never
// decreases, so the result is that of an exit after every dimension.
template <bool V2>
__global__ __launch_bounds__(256) void band_recheck_kernel(
    const double *__restrict__ pa, const uint32_t *__restrict__ perma,
    const double *__restrict__ pb, const uint32_t *__restrict__ permb,
    const SegDev *__restrict__ segs, const uint32_t *__restrict__ tile_seg, int d,
    const uint64_t *__restrict__ nband_dev, FilterOut out, uint64_t *__restrict__ ctr,
    uint64_t *__restrict__ seg_counts) {
    // the queue length is read on the device (no host round trip); entries
    // past the capacity were not stored and are redone by the caller
    const uint64_t *band = out.band;
    const uint64_t nband = min64(*nband_dev, out.band_cap);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t lane = threadIdx.x & 63;
    uint64_t total = 0;
    for (uint64_t k0 = (uint64_t)blockIdx.x * blockDim.x; k0 < nband; k0 += stride) {
        const uint64_t k = k0 + threadIdx.x;
        bool in = false;
        uint64_t e = 0;
        uint32_t sg = 0;
        if (k < nband) {
            e = band[k];
            const uint32_t row = (uint32_t)e, col = (uint32_t)(e >> 32);
            sg = tile_seg[row >> 6];
            const SegDev &S = segs[sg];
            const uint64_t qs = S.qoff + 64ull * ((row >> 6) - S.qtile0) + (row & 63);
            const uint64_t ts = S.toff + 64ull * ((col >> 6) - S.ttile0) + (col & 63);
            const double *a = pa + (size_t)perma[qs] * d;
            const double *b = pb + (size_t)permb[ts] * d;
            const double eps2 = S.eps2;
            double s = 0.0;
            for (int c0 = 0; c0 < d; c0 += 8) {
                double av[8], bv[8];
                if constexpr (V2) {
#pragma unroll
                    for (int u = 0; u < 8; u += 2) {
                        const int c = c0 + u;  // d even: c < d covers c + 1
                        double2 x = make_double2(0.0, 0.0), y = x;
                        if (c < d) {
                            x = *reinterpret_cast<const double2 *>(a + c);
                            y = *reinterpret_cast<const double2 *>(b + c);
                        }
                        av[u] = x.x;
                        av[u + 1] = x.y;
                        bv[u] = y.x;
                        bv[u + 1] = y.y;
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int c = c0 + u;
                        av[u] = c < d ? a[c] : 0.0;
                        bv[u] = c < d ? b[c] : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {  // dimensions past d add exactly +0
                    const double t = av[u] - bv[u];
                    s = __builtin_fma(t, t, s);
                }
                if (s > eps2) break;
            }
            in = s <= eps2;
        }
        // one atomic per wavefront for each counter (the band of dense data
        // holds many pairs); per-lane segment atomics only where a
        // wavefront's accepted pairs span several segments
        const uint64_t hits = __ballot(in);
        if (hits) {
            const uint32_t first = (uint32_t)__builtin_ctzll(hits);
            if (seg_counts) {
                const uint32_t sg0 = __builtin_amdgcn_readlane(sg, first);
                if (__all(!in || sg == sg0)) {
                    if (lane == first)
                        atomicAdd((unsigned long long *)&seg_counts[sg0],
                                  (unsigned long long)__popcll(hits));
                } else if (in) {
                    atomicAdd((unsigned long long *)&seg_counts[sg], 1ull);
                }
            }
            if (out.pairs) {
                uint64_t base = 0;
                if (lane == first)
                    base = atomicAdd((unsigned long long *)&ctr[6],
                                     (unsigned long long)__popcll(hits));
                const uint64_t o = read_lane_u64(base, first) + lanes_below(hits);
                if (in && o < out.pairs_cap) out.pairs[o] = e;
            }
        }
        total += (uint64_t)__popcll(hits);
    }
    if (lane == 0 && total) atomicAdd((unsigned long long *)&ctr[1], (unsigned long long)total);
}

}  // namespace

bool filter_prep_fused(int d) { return d <= FILTER_PREP_MAX_D; }

template <int WAVES>
void frag_frames(size_t ntiles, const double *pts, const uint32_t *perm, const SegDev *segs,
                 const uint32_t *tile_seg, const GridDev *grids, const double *centre,
                 const PackParams &pp, void *frow, void *fcol, const FrameOut &fo,
                 const double *zmax, uint32_t *opflag, hipStream_t s) {
    const int d = pp.d;
    const size_t lds = stage_halfs(d) * sizeof(uint16_t) +
                       ((size_t)3 * d + (size_t)2 * WAVES * d) * sizeof(double);
    auto k = pp.bf16 ? frag_frames_kernel<WAVES, true> : frag_frames_kernel<WAVES, false>;
    k<<<(unsigned)ntiles, 64 * WAVES, lds, s>>>(
        pts, perm, segs, tile_seg, grids, centre, pp, static_cast<u32x4 *>(frow),
        static_cast<u32x4 *>(fcol), static_cast<u32x4 *>(fo.frow2), static_cast<u32x4 *>(fo.fcol2),
        fo.anch, fo.info, fo.flag, zmax, opflag);
}

void launch_filter_prep(const double *pts, const uint32_t *perm, const SegDev *segs,
                        const uint32_t *tile_seg, size_t ntiles, bool target_side, int d,
                        const GridDev *grids, const double *centre, const FilterParams &fp,
                        void *frow, void *fcol, const FrameOut &fo, const double *zmax,
                        uint32_t *opflag, hipStream_t s) {
    if (ntiles == 0) return;
    if (opflag && !(filter_prep_fused(d) && zmax))
        throw InvalidArg{"internal: the range check needs a staged tile and its bounds"};
    PackParams pp{d, fp.split, fp.steps1, fp.steps, fp.parts, fp.bf16, target_side ? 1 : 0};
    if (fo.anch) {
        if (!(d <= FRAME_MAX_D && filter_prep_fused(d)) || (frow && !fo.frow2) ||
            (fcol && !fo.fcol2) || !fo.info || !fo.flag)
            throw InvalidArg{"internal: tile frames need a staged tile and their arrays"};
        // short rows: one wavefront per tile keeps four times more tiles in
        // flight (their latency dominates); longer rows: four per tile
        if (d <= 32)
            frag_frames<1>(ntiles, pts, perm, segs, tile_seg, grids, centre, pp, frow, fcol, fo,
                           zmax, opflag, s);
        else
            frag_frames<4>(ntiles, pts, perm, segs, tile_seg, grids, centre, pp, frow, fcol, fo,
                           zmax, opflag, s);
    } else if (filter_prep_fused(d)) {
        const size_t lds = stage_halfs(d) * sizeof(uint16_t) + (size_t)2 * d * sizeof(double);
        if (d <= 32) {
            auto k = pp.bf16 ? frag_tile_kernel<true, 1, true> : frag_tile_kernel<true, 1, false>;
            k<<<(unsigned)ntiles, 64, lds, s>>>(pts, perm, segs, tile_seg, grids, centre, pp,
                                               static_cast<u32x4 *>(frow),
                                               static_cast<u32x4 *>(fcol), zmax, opflag);
        } else {
            auto k = pp.bf16 ? frag_tile_kernel<true, 4, true> : frag_tile_kernel<true, 4, false>;
            k<<<(unsigned)ntiles, 256, lds, s>>>(pts, perm, segs, tile_seg, grids, centre, pp,
                                                static_cast<u32x4 *>(frow),
                                                static_cast<u32x4 *>(fcol), zmax, opflag);
        }
    } else {
        auto k = pp.bf16 ? frag_tile_kernel<false, 4, true> : frag_tile_kernel<false, 4, false>;
        k<<<(unsigned)ntiles, 256, 0, s>>>(pts, perm, segs, tile_seg, grids, centre, pp,
                                          static_cast<u32x4 *>(frow), static_cast<u32x4 *>(fcol),
                                          nullptr, nullptr);
    }
    HJ_HIP(hipGetLastError());
}

template <bool SELF, bool BF16, uint32_t CHUNK, bool EMIT, int FRM, bool UNI>
void mfma_launch_k(const JoinArgs &ja, const u32x4 *a, const u32x4 *b, const FilterParams &fp,
                   const FilterOut &out, const FrameArgsDev &fr, hipStream_t s) {
    auto k = mfma_count_kernel<SELF, BF16, CHUNK, EMIT, FRM, UNI>;
    k<<<persistent_grid(k, 256, ja.num_cus, ja.grid), 256, 0, s>>>(
        a, b, ja.pairs, ja.npairs, ja.segs, ja.tile_seg, fp.steps1, fp.steps, ja.ctr,
        ja.seg_counts, ja.sched, out, fr, ja.npairs_dev, ja.opflag);
}

// One-row chunks (UNI) where the list groups a row tile's pairs, for FP16
// (measured: BF16's more frequent slow path loses with them).
template <bool SELF, bool BF16, uint32_t CHUNK, bool EMIT, int FRM>
void mfma_launch_fr(const JoinArgs &ja, const u32x4 *a, const u32x4 *b, const FilterParams &fp,
                    const FilterOut &out, const FrameArgsDev &fr, hipStream_t s) {
    if constexpr (!BF16 && FRM != 2)
        if (ja.rows_grouped) return mfma_launch_k<SELF, BF16, CHUNK, EMIT, FRM, true>(ja, a, b, fp, out, fr, s);
    mfma_launch_k<SELF, BF16, CHUNK, EMIT, FRM, false>(ja, a, b, fp, out, fr, s);
}

template <bool SELF, bool BF16, uint32_t CHUNK, bool EMIT = false>
void mfma_launch(const JoinArgs &ja, const u32x4 *a, const u32x4 *b, const FilterParams &fp,
                 const FilterOut &out, hipStream_t s) {
    if (ja.frames) {
        const auto &f = ja.fr;
        if (fp.steps1 != 1 || !f.qinfo || !f.tinfo || !f.qanch || !f.tanch || !f.qflag ||
            !f.tflag || (!f.fa2 && !f.fb2))
            throw InvalidArg{"internal: tile frames without their tables"};
        const FrameArgsDev fr{static_cast<const u32x4 *>(f.fa2), static_cast<const u32x4 *>(f.fb2),
                              f.qinfo, f.tinfo, f.qanch, f.tanch, f.qflag, f.tflag,
                              fp.frames_fast};
        mfma_launch_fr<SELF, BF16, CHUNK, EMIT, 0>(ja, a, b, fp, out, fr, s);
        if (fr.fast)
            mfma_launch_fr<SELF, BF16, CHUNK, EMIT, 2>(ja, a, b, fp, out, fr, s);
        else
            mfma_launch_fr<SELF, BF16, CHUNK, EMIT, 1>(ja, a, b, fp, out, fr, s);
    } else {
        mfma_launch_fr<SELF, BF16, CHUNK, EMIT, 0>(ja, a, b, fp, out, FrameArgsDev{}, s);
    }
}

template <bool SELF, bool BF16>
void mfma_chunked(const JoinArgs &ja, const u32x4 *a, const u32x4 *b, const FilterParams &fp,
                  const FilterOut &out, hipStream_t s) {
    // pair emission: one schedule only (its pairs go through a sort anyway)
    if (out.pairs)
        mfma_launch<SELF, BF16, 4, true>(ja, a, b, fp, out, s);
    else if (ja.chunk == 2)
        mfma_launch<SELF, BF16, 2>(ja, a, b, fp, out, s);
    else if (ja.chunk == 8)
        mfma_launch<SELF, BF16, 8>(ja, a, b, fp, out, s);
    else
        mfma_launch<SELF, BF16, 4>(ja, a, b, fp, out, s);
}

void launch_mfma_count(const JoinArgs &ja, const void *fragA, const void *fragB,
                       const FilterParams &fp, bool self, const FilterOut &out, hipStream_t s) {
    if (ja.npairs == 0) return;
    if (!work_list_fits(ja.npairs, 2)) throw InvalidArg{"work list too long"};
    if (!ja.sched) throw InvalidArg{"internal: filter launch without scheduling counters"};
    const u32x4 *a = static_cast<const u32x4 *>(fragA);
    const u32x4 *b = static_cast<const u32x4 *>(fragB);
    if (self) {
        if (fp.bf16)
            mfma_chunked<true, true>(ja, a, b, fp, out, s);
        else
            mfma_chunked<true, false>(ja, a, b, fp, out, s);
    } else {
        if (fp.bf16)
            mfma_chunked<false, true>(ja, a, b, fp, out, s);
        else
            mfma_chunked<false, false>(ja, a, b, fp, out, s);
    }
    HJ_HIP(hipGetLastError());
}

void launch_band_recheck(const double *ptsA, const uint32_t *permA, const double *ptsB,
                         const uint32_t *permB, const SegDev *segs, const uint32_t *tile_seg,
                         int d, const uint64_t *nband_dev, const FilterOut &out, uint64_t *ctr,
                         uint64_t *seg_counts, int num_cus, hipStream_t s) {
    const unsigned blocks = (unsigned)std::max(1, num_cus) * 8;
    // 16-byte row loads when every row starts on a 16-byte boundary
    const bool v2 = d % 2 == 0 && reinterpret_cast<uintptr_t>(ptsA) % 16 == 0 &&
                    reinterpret_cast<uintptr_t>(ptsB) % 16 == 0;
    if (v2)
        band_recheck_kernel<true><<<blocks, 256, 0, s>>>(ptsA, permA, ptsB, permB, segs,
                                                         tile_seg, d, nband_dev, out, ctr,
                                                         seg_counts);
    else
        band_recheck_kernel<false><<<blocks, 256, 0, s>>>(ptsA, permA, ptsB, permB, segs,
                                                          tile_seg, d, nband_dev, out, ctr,
                                                          seg_counts);
    HJ_HIP(hipGetLastError());
}

}  // namespace hj
