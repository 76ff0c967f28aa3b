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
//   launch_tile_bounds /          stripe bounds: prepareStripes
//   launch_tile_ranges            (egojoin.cpp:270-464) and prepareStripesNoSelf
//                                 (:1126-1240) find, per point and stripe, the
//                                 sorted range of candidate partners with
//                                 galloping searches, then take block min/max
//                                 (:437-447).  Here a whole 64-point tile is
//                                 bounded at once (see launch_tile_ranges).
//   launch_emit_pairs /           the Hilbert traversal FGF_HILBERT_FOR
//   launch_hilbert_keys           (hilloop.h:9-82) with its cond/cond2 pruning
//                                 (egojoin.h:325-328) becomes an explicit list
//                                 of candidate tile pairs, ordered by the
//                                 textbook Hilbert index of (I, J).
#include <algorithm>

#include "hj_internal.h"
#include "hj_device.h"

namespace hj {

namespace {

constexpr int MM_ROWS = 4;
constexpr int MM_UNROLL = 8;

// Per-dimension min / max: block (64 dimensions) x (MM_ROWS row lanes); each
// thread keeps MM_UNROLL independent loads in flight.
__global__ __launch_bounds__(256) void minmax_partial(const double *__restrict__ pts, size_t n,
                                                      int d, double *__restrict__ pmin,
                                                      double *__restrict__ pmax,
                                                      uint32_t *__restrict__ bad) {
    __shared__ double smin[MM_ROWS][64], smax[MM_ROWS][64];
    const int dim = blockIdx.x * 64 + threadIdx.x;
    double lo = INFINITY, hi = -INFINITY;
    uint32_t nonfinite = 0;
    if (dim < d) {
        const size_t step = (size_t)gridDim.y * MM_ROWS;
        for (size_t r0 = (size_t)blockIdx.y * MM_ROWS + threadIdx.y; r0 < n;
             r0 += step * MM_UNROLL) {
            double v[MM_UNROLL];
#pragma unroll
            for (int u = 0; u < MM_UNROLL; u++) {
                const size_t r = r0 + u * step;
                v[u] = r < n ? pts[r * d + dim] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < MM_UNROLL; u++) {
                if (r0 + u * step < n) {
                    if (!isfinite(v[u])) nonfinite = 1;
                    lo = fmin(lo, v[u]);
                    hi = fmax(hi, v[u]);
                }
            }
        }
    }
    if (nonfinite) atomicOr(bad, 1u);
    smin[threadIdx.y][threadIdx.x] = lo;
    smax[threadIdx.y][threadIdx.x] = hi;
    __syncthreads();
    if (threadIdx.y == 0 && dim < d) {
        for (int y = 1; y < MM_ROWS; y++) {
            lo = fmin(lo, smin[y][threadIdx.x]);
            hi = fmax(hi, smax[y][threadIdx.x]);
        }
        pmin[(size_t)blockIdx.y * d + dim] = lo;
        pmax[(size_t)blockIdx.y * d + dim] = hi;
    }
}

// one block per dimension reduces the partial results
__global__ __launch_bounds__(256) void minmax_final(const double *__restrict__ pmin,
                                                    const double *__restrict__ pmax, int parts,
                                                    int d, double *__restrict__ mins,
                                                    double *__restrict__ maxs) {
    __shared__ double slo[256], shi[256];
    const int dim = blockIdx.x;
    double lo = INFINITY, hi = -INFINITY;
    for (int p = threadIdx.x; p < parts; p += 256) {
        lo = fmin(lo, pmin[(size_t)p * d + dim]);
        hi = fmax(hi, pmax[(size_t)p * d + dim]);
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
        mins[dim] = slo[0];
        maxs[dim] = shi[0];
    }
}

// One wavefront per 64 consecutive points: the leading key_dims coordinates of
// the 64 rows are staged in LDS with row-contiguous (coalesced) loads, then each
// lane packs its point's key.
__global__ __launch_bounds__(256) void keys_kernel(const double *__restrict__ pts, size_t n, int d,
                                                   KeyParams kp, const int64_t *__restrict__ minc,
                                                   uint64_t *__restrict__ keys,
                                                   uint32_t *__restrict__ vals) {
    extern __shared__ double kx[];  // [4 waves][64][key_dims + 1]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kd = kp.key_dims, stride = kd + 1;
    double *x = kx + (size_t)wave * 64 * stride;
    const size_t base = ((size_t)blockIdx.x * 4 + wave) * 64;
    for (int e = lane; e < 64 * kd; e += 64) {
        const int p = e / kd, k = e - p * kd;
        const size_t i = base + p;
        x[p * stride + k] = i < n ? pts[i * d + k] : 0.0;
    }
    __syncthreads();
    const size_t i = base + lane;
    if (i >= n) return;
    keys[i] = point_key(x + lane * stride, 1, kp, minc);
    vals[i] = (uint32_t)i;
}

// One block per 64-point tile: rows[i][k] (row-major, padded to dpad
// dimensions) and cols[I][k][p] (dimension-major inside the tile).
__global__ __launch_bounds__(256) void gather_kernel(const double *__restrict__ pts,
                                                     const uint32_t *__restrict__ perm, size_t n,
                                                     int d, int dpad, double *__restrict__ rows,
                                                     double *__restrict__ cols) {
    __shared__ double lds[64][65];
    const size_t I = blockIdx.x;
    const int t = threadIdx.x;
    for (int c0 = 0; c0 < dpad; c0 += 64) {
        for (int e = t; e < 64 * 64; e += 256) {
            int p = e >> 6, kk = e & 63, k = c0 + kk;
            size_t i = I * 64 + p;
            double v;
            if (i < n)
                v = k < d ? pts[(size_t)perm[i] * d + k] : 0.0;
            else
                v = k < d ? PAD : 0.0;
            if (k < dpad) rows[i * dpad + k] = v;
            lds[p][kk] = v;
        }
        __syncthreads();
        for (int e = t; e < 64 * 64; e += 256) {
            int kk = e >> 6, p = e & 63, k = c0 + kk;
            size_t i = I * 64 + p;
            if (k < dpad) {
                double v = lds[p][kk];
                if (i >= n && k < d) v = -PAD;
                cols[(I * dpad + k) * 64 + p] = v;
            }
        }
        __syncthreads();
    }
}

// Per tile: tlo = min over its points of (key - Wlo), thi = max of (key + Whi)
// (see tile_key_bounds), read from the transposed tile layout (coalesced).
__global__ __launch_bounds__(256) void tile_bounds_kernel(const double *__restrict__ cols,
                                                          size_t n, int dpad, KeyParams kp,
                                                          const int64_t *__restrict__ minc,
                                                          size_t ntiles, uint64_t *__restrict__ tlo,
                                                          uint64_t *__restrict__ thi) {
    const size_t I = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (I >= ntiles) return;
    const size_t i = I * 64 + lane;
    uint64_t lo = ~0ull, hi = 0ull;
    if (i < n) point_key_bounds(cols + I * dpad * 64 + lane, 64, kp, minc, &lo, &hi);
    wave_min_max_u64(lo, hi);
    if (lane == 0) {
        tlo[I] = lo;
        thi[I] = hi;
    }
}

__device__ __forceinline__ size_t lower_bound_u64(const uint64_t *__restrict__ a, size_t n,
                                                  uint64_t x) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t m = lo + ((hi - lo) >> 1);
        if (a[m] < x)
            lo = m + 1;
        else
            hi = m;
    }
    return lo;
}

__device__ __forceinline__ size_t upper_bound_u64(const uint64_t *__restrict__ a, size_t n,
                                                  uint64_t x) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t m = lo + ((hi - lo) >> 1);
        if (a[m] <= x)
            lo = m + 1;
        else
            hi = m;
    }
    return lo;
}

struct RangeParams {
    int key_dims, bits, active;
    int code_begin, code_end;  // base-3 codes enumerated
    int self;
};

constexpr int MAX_CODES = 729;  // 3^6

// One wavefront per query tile.  Code c (base 3, most significant digit =
// dimension 0) encodes the offsets delta_k = digit_k - 1 of the first
// `active` key fields.  For code c the candidates are the targets with keys in
//   [tlo + D_lo(c), thi + D_hi(c)],
//   D_lo(c) = sum_k<A delta_k 2^s_k - sum_A<=k<D' 2^s_k,
//   D_hi(c) = sum_k<A delta_k 2^s_k + sum_A<=k<D' 2^s_k   (s_k = bits (D'-1-k)),
// i.e. the stripes of the reference with exact cell arithmetic instead of
// displaced coordinates.  Self-joins need only codes >=lex 0 (partners later
// in the EGO order) and tiles J >= I.  Bounds grow with c, so the tile
// intervals arrive sorted and are merged in one pass.
__global__ __launch_bounds__(64) void tile_ranges_kernel(
    const uint64_t *__restrict__ tlo, const uint64_t *__restrict__ thi, size_t ntq,
    const uint64_t *__restrict__ tkeys, size_t nt, RangeParams rp, uint32_t *__restrict__ ivals,
    uint32_t *__restrict__ icount, uint64_t *__restrict__ jcount, int max_iv) {
    __shared__ uint32_t s_start[MAX_CODES], s_end[MAX_CODES];
    const size_t I = blockIdx.x;
    const int lane = threadIdx.x;
    if (I >= ntq) return;
    const uint64_t lo0 = tlo[I], hi0 = thi[I];
    int64_t rest = 0;
    for (int k = rp.active; k < rp.key_dims; k++)
        rest += (int64_t)1 << (rp.bits * (rp.key_dims - 1 - k));
    const uint32_t ntt = (uint32_t)((nt + 63) / 64);
    const int ncodes = rp.code_end - rp.code_begin;
    for (int ci = lane; ci < ncodes; ci += 64) {
        int code = rp.code_begin + ci;
        int64_t dsum = 0;
        int cc = code;
        for (int k = rp.active - 1; k >= 0; k--) {
            int digit = cc % 3;
            cc /= 3;
            dsum += (int64_t)(digit - 1) * ((int64_t)1 << (rp.bits * (rp.key_dims - 1 - k)));
        }
        uint64_t qlo = lo0 + (uint64_t)(dsum - rest);
        uint64_t qhi = hi0 + (uint64_t)(dsum + rest);
        size_t s = lower_bound_u64(tkeys, nt, qlo);
        size_t e = upper_bound_u64(tkeys, nt, qhi);
        uint32_t ts = (uint32_t)(s / 64), te = (uint32_t)((e + 63) / 64);
        if (s >= e) ts = te = 0;
        if (rp.self && ts < (uint32_t)I) ts = (uint32_t)I;
        if (te > ntt) te = ntt;
        s_start[ci] = ts;
        s_end[ci] = te;
    }
    __syncthreads();
    if (lane == 0) {
        uint32_t *out = ivals + I * (size_t)max_iv * 2;
        int m = 0;
        uint64_t tot = 0;
        uint32_t cs = 0, ce = 0;
        for (int ci = 0; ci < ncodes; ci++) {
            uint32_t s = s_start[ci], e = s_end[ci];
            if (s >= e) continue;
            if (m > 0 && s <= ce) {
                if (e > ce) ce = e;
            } else {
                if (m > 0) {
                    out[2 * (m - 1)] = cs;
                    out[2 * (m - 1) + 1] = ce;
                    tot += ce - cs;
                }
                cs = s;
                ce = e;
                m++;
            }
        }
        if (m > 0) {
            out[2 * (m - 1)] = cs;
            out[2 * (m - 1) + 1] = ce;
            tot += ce - cs;
        }
        icount[I] = (uint32_t)m;
        jcount[I] = tot;
    }
}

__global__ __launch_bounds__(64) void emit_pairs_kernel(const uint32_t *__restrict__ ivals,
                                                        const uint32_t *__restrict__ icount,
                                                        const uint64_t *__restrict__ off,
                                                        size_t ntq, int max_iv,
                                                        uint64_t *__restrict__ pairs) {
    const size_t I = blockIdx.x;
    if (I >= ntq) return;
    const uint32_t *iv = ivals + I * (size_t)max_iv * 2;
    uint64_t o = off[I];
    const uint32_t m = icount[I];
    for (uint32_t q = 0; q < m; q++) {
        uint32_t s = iv[2 * q], e = iv[2 * q + 1];
        for (uint32_t J = s + threadIdx.x; J < e; J += 64)
            pairs[o + (J - s)] = (uint64_t)I | ((uint64_t)J << 32);
        o += e - s;
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

// Number of distinct key prefixes over the first a fields, a = 1..6: counts[a]
// += number of i with prefix_a(key[i]) != prefix_a(key[i-1]) (plus the first
// point).  Sorted keys make equal prefixes contiguous.
__global__ __launch_bounds__(256) void prefix_classes_kernel(const uint64_t *__restrict__ keys,
                                                             size_t n, int key_dims, int bits,
                                                             unsigned long long *__restrict__ counts) {
    __shared__ unsigned int sc[7];
    if (threadIdx.x < 7) sc[threadIdx.x] = 0;
    __syncthreads();
    unsigned int c[7] = {0, 0, 0, 0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        const uint64_t p = i ? keys[i - 1] : ~k;
        for (int a = 1; a <= 6 && a <= key_dims; a++) {
            const int sh = bits * (key_dims - a);
            c[a] += (k >> sh) != (p >> sh) ? 1u : 0u;
        }
    }
    for (int a = 1; a <= 6; a++)
        if (c[a]) atomicAdd(&sc[a], c[a]);
    __syncthreads();
    if (threadIdx.x >= 1 && threadIdx.x < 7 && sc[threadIdx.x])
        atomicAdd(&counts[threadIdx.x], (unsigned long long)sc[threadIdx.x]);
}

__global__ void zero_u64_kernel(uint64_t *p, size_t n) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

inline unsigned blocks_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

void launch_minmax(const double *pts, size_t n, int d, double *mins, double *maxs,
                   uint32_t *bad, DevBuf &tmp, hipStream_t s) {
    const int dblocks = (d + 63) / 64;
    int parts = (int)((n + 8191) / 8192);
    parts = std::max(1, std::min(parts, std::max(1, 2048 / dblocks)));
    tmp.ensure((size_t)2 * parts * d * sizeof(double));
    double *pmin = tmp.as<double>(), *pmax = pmin + (size_t)parts * d;
    dim3 grid((unsigned)dblocks, (unsigned)parts);
    minmax_partial<<<grid, dim3(64, MM_ROWS), 0, s>>>(pts, n, d, pmin, pmax, bad);
    HJ_HIP(hipGetLastError());
    minmax_final<<<(unsigned)d, 256, 0, s>>>(pmin, pmax, parts, d, mins, maxs);
    HJ_HIP(hipGetLastError());
}

void launch_keys(const double *pts, size_t n, int d, const Grid &g, const int64_t *minc_dev,
                 uint64_t *keys, uint32_t *vals, hipStream_t s) {
    if (n == 0) return;
    KeyParams kp{g.r, g.key_dims, g.bits};
    const size_t lds = (size_t)4 * 64 * (g.key_dims + 1) * sizeof(double);
    keys_kernel<<<blocks_for(n, 256), 256, lds, s>>>(pts, n, d, kp, minc_dev, keys, vals);
    HJ_HIP(hipGetLastError());
}

void launch_gather(const double *pts, const uint32_t *perm, size_t n, int d, int dpad,
                   double *rows, double *cols, hipStream_t s) {
    size_t ntiles = (n + 63) / 64;
    if (ntiles == 0) return;
    gather_kernel<<<(unsigned)ntiles, 256, 0, s>>>(pts, perm, n, d, dpad, rows, cols);
    HJ_HIP(hipGetLastError());
}

void launch_tile_bounds(const double *cols, size_t n, int dpad, const Grid &g,
                        const int64_t *minc_dev, uint64_t *tlo, uint64_t *thi, hipStream_t s) {
    size_t ntiles = (n + 63) / 64;
    if (ntiles == 0) return;
    KeyParams kp{g.r, g.key_dims, g.bits};
    tile_bounds_kernel<<<blocks_for(ntiles, 4), 256, 0, s>>>(cols, n, dpad, kp, minc_dev,
                                                              ntiles, tlo, thi);
    HJ_HIP(hipGetLastError());
}

void launch_tile_ranges(const uint64_t *tlo, const uint64_t *thi, size_t ntiles_q,
                        const uint64_t *tkeys, size_t nt, const Grid &g, bool self,
                        uint32_t *ivals, uint32_t *icount, uint64_t *jcount, int max_iv,
                        hipStream_t s) {
    if (ntiles_q == 0) return;
    int total = 1;
    for (int k = 0; k < g.active; k++) total *= 3;
    RangeParams rp;
    rp.key_dims = g.key_dims;
    rp.bits = g.bits;
    rp.active = g.active;
    rp.code_begin = self ? (total - 1) / 2 : 0;
    rp.code_end = total;
    rp.self = self ? 1 : 0;
    if (rp.code_end - rp.code_begin > max_iv || total > MAX_CODES)
        throw InvalidArg{"too many stripe codes"};
    tile_ranges_kernel<<<(unsigned)ntiles_q, 64, 0, s>>>(tlo, thi, ntiles_q, tkeys, nt, rp,
                                                         ivals, icount, jcount, max_iv);
    HJ_HIP(hipGetLastError());
}

void launch_emit_pairs(const uint32_t *ivals, const uint32_t *icount, const uint64_t *off,
                       size_t ntiles_q, int max_iv, uint64_t *pairs, hipStream_t s) {
    if (ntiles_q == 0) return;
    emit_pairs_kernel<<<(unsigned)ntiles_q, 64, 0, s>>>(ivals, icount, off, ntiles_q, max_iv,
                                                        pairs);
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

void launch_prefix_classes(const uint64_t *keys, size_t n, int key_dims, int bits,
                           uint64_t *counts, int num_cus, hipStream_t s) {
    launch_zero_u64(counts, 8, s);
    if (n == 0) return;
    size_t blocks = std::min<size_t>((n + 255) / 256, (size_t)std::max(1, num_cus) * 4);
    prefix_classes_kernel<<<(unsigned)blocks, 256, 0, s>>>(
        keys, n, key_dims, bits, reinterpret_cast<unsigned long long *>(counts));
    HJ_HIP(hipGetLastError());
}

void launch_zero_u64(uint64_t *p, size_t n, hipStream_t s) {
    if (n == 0) return;
    zero_u64_kernel<<<blocks_for(n, 256), 256, 0, s>>>(p, n);
    HJ_HIP(hipGetLastError());
}

}  // namespace hj

// ---------------------------------------------------------------------------
// CSR post-processing
// ---------------------------------------------------------------------------
namespace hj {
namespace {

__global__ void map_pairs_kernel(const uint64_t *__restrict__ in, uint64_t np,
                                 const uint32_t *__restrict__ permA,
                                 const uint32_t *__restrict__ permB, int mirror,
                                 uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    uint64_t p = in[k];
    uint64_t a = permA[(uint32_t)p], b = permB[(uint32_t)(p >> 32)];
    keys[k] = (a << 32) | b;
    vals[k] = 0;
    if (mirror) {
        keys[np + k] = (b << 32) | a;
        vals[np + k] = 0;
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

void launch_map_pairs(const uint64_t *in, uint64_t np, const uint32_t *permA,
                      const uint32_t *permB, bool mirror, uint64_t *out_keys, uint32_t *out_vals,
                      hipStream_t s) {
    if (np == 0) return;
    map_pairs_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(in, np, permA, permB,
                                                                  mirror ? 1 : 0, out_keys,
                                                                  out_vals);
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
