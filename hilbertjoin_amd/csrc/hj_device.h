// hj_device.h -- device helpers shared by the preparation kernels: grid
// keys of a point and their widened tile bounds.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hj_internal.h"

namespace hj {

// Local key of one point: field k (k < key_dims, dimension 0 most
// significant) holds v_k = floor(x_k * r) - minc_k, clamped to
// [2, 2^bits - 3].  Grids are built so that every point of the set the grid
// was made for has v_k in [2, 2^bits - 3] (minc_k = min cell - 2 and 2^bits >
// range + 4), so adding or removing up to 2 in any field needs no carry or
// borrow; a query point outside a resident index's grid is clamped, which only
// widens its candidate set.
__device__ __forceinline__ uint64_t point_key(const double *x, size_t stride, const GridDev &g,
                                              const int64_t *__restrict__ minc) {
    uint64_t key = 0;
    const int64_t vmax = ((int64_t)1 << g.bits) - 3;
    for (uint32_t k = 0; k < g.key_dims; k++) {
        double X = x[(size_t)k * stride] * g.r;
        int64_t v = (int64_t)floor(X) - minc[k];
        v = v < 2 ? 2 : (v > vmax ? vmax : v);
        key = (key << g.bits) | (uint64_t)v;
    }
    return key;
}

// Key field of one coordinate x_k (cell minus minc_k, clamped to [2, 2^bits - 3])
// and its widening flags: l = 1 if the cell could be one lower, h = 1 if one
// higher, within FP64 rounding of the cell boundary (point_key_bounds).
__device__ __forceinline__ void key_field(double xk, const GridDev &g, int64_t minck, uint64_t *v,
                                          uint64_t *l, uint64_t *h) {
    const int64_t vmax = ((int64_t)1 << g.bits) - 3;
    const double X = xk * g.r;
    const double c = floor(X);
    const double f = X - c;
    const double eta = ldexp(1.0 + fabs(X), -48);
    int64_t w = (int64_t)c - minck;
    w = w < 2 ? 2 : (w > vmax ? vmax : w);
    *v = (uint64_t)w;
    *l = f < eta ? 1ull : 0ull;
    *h = f >= 1.0 - eta ? 1ull : 0ull;
}

// Widened local key bounds of one point: lo = key - Wlo, hi = key + Whi, where
// Wlo / Whi carry a 1 in every field whose coordinate lies within eta of the
// lower / upper cell boundary.  A partner q of p accepted by the FP64
// predicate has |fl(x_q r) - fl(x_p r)| <= 1 + eta with eta = 2^-48 (1+|x_p r|)
// (|x_q - x_p| <= eps (1 + 3u), r = fl(1/eps)), so its cell differs by at most
// one, or by two across such a boundary.
__device__ __forceinline__ void point_key_bounds(const double *x, size_t stride, const GridDev &g,
                                                 const int64_t *__restrict__ minc, uint64_t *lo,
                                                 uint64_t *hi) {
    uint64_t key = 0, wlo = 0, whi = 0;
    for (uint32_t k = 0; k < g.key_dims; k++) {
        uint64_t v, l, h;
        key_field(x[(size_t)k * stride], g, minc[k], &v, &l, &h);
        key = (key << g.bits) | v;
        wlo = (wlo << g.bits) | l;
        whi = (whi << g.bits) | h;
    }
    *lo = key - wlo;
    *hi = key + whi;
}

// Segment of input row i: the largest s with off[s] <= i (off has nseg + 1
// entries, ascending).
__device__ __forceinline__ uint32_t find_seg(const uint64_t *__restrict__ off, uint32_t nseg,
                                             uint64_t i) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (off[m] <= i)
            lo = m;
        else
            hi = m;
    }
    return lo;
}

// First sorted position and number of real points of layout tile t of a
// segment, on the query or the target side.
__device__ __forceinline__ void tile_span(const SegDev &s, uint32_t t, bool target,
                                          uint64_t *start, uint32_t *nv) {
    const uint32_t lt = t - (target ? s.ttile0 : s.qtile0);
    const uint64_t n = target ? s.tn : s.qn;
    *start = (target ? s.toff : s.qoff) + 64ull * lt;
    const uint64_t rem = n - 64ull * lt;
    *nv = rem < 64 ? (uint32_t)rem : 64u;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, m);
    hi = __shfl_xor(hi, m);
    return ((uint64_t)hi << 32) | lo;
}

// wavefront-wide min of lo and max of hi (all 64 lanes must call)
__device__ __forceinline__ void wave_min_max_u64(uint64_t &lo, uint64_t &hi) {
    for (int m = 32; m >= 1; m >>= 1) {
        uint64_t olo = shfl_xor_u64(lo, m), ohi = shfl_xor_u64(hi, m);
        lo = olo < lo ? olo : lo;
        hi = ohi > hi ? ohi : hi;
    }
}

}  // namespace hj
