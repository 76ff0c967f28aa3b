// hj_device.h -- device helpers shared by the preparation kernels: grid
// keys of a point and their widened tile bounds.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hj {

struct KeyParams {
    double r;      // 1 / eps (FP64), the cell rule floor(x * r) of egojoin.cpp:9-17
    int key_dims;  // leading dimensions packed in the key
    int bits;      // bits per packed dimension
};

// Packed key of one point: field k (k < key_dims, dimension 0 most
// significant) holds v_k = c_k - minc_k + 2 with c_k = floor(x_k * r).  Fields
// never leave [0, 2^bits) because 2^bits > (max c - min c) + 4, so adding or
// removing up to 2 in any field needs no carry or borrow.
__device__ __forceinline__ uint64_t point_key(const double *x, size_t stride, KeyParams kp,
                                              const int64_t *__restrict__ minc) {
    uint64_t key = 0;
    for (int k = 0; k < kp.key_dims; k++) {
        double X = x[(size_t)k * stride] * kp.r;
        int64_t c = (int64_t)floor(X);
        key = (key << kp.bits) | (uint64_t)(c - minc[k] + 2);
    }
    return key;
}

// Widened key bounds of one point: lo = key - Wlo, hi = key + Whi, where
// Wlo / Whi carry a 1 in every field whose coordinate lies within eta of the
// lower / upper cell boundary.  A partner q of p accepted by the FP64
// predicate has |fl(x_q r) - fl(x_p r)| <= 1 + eta with eta = 2^-48 (1+|x_p r|)
// (|x_q - x_p| <= eps (1 + 3u), r = fl(1/eps)), so its cell differs by at most
// one, or by two across such a boundary.
__device__ __forceinline__ void point_key_bounds(const double *x, size_t stride, KeyParams kp,
                                                 const int64_t *__restrict__ minc, uint64_t *lo,
                                                 uint64_t *hi) {
    uint64_t key = 0, wlo = 0, whi = 0;
    for (int k = 0; k < kp.key_dims; k++) {
        double X = x[(size_t)k * stride] * kp.r;
        double c = floor(X);
        double f = X - c;
        double eta = ldexp(1.0 + fabs(X), -48);
        key = (key << kp.bits) | (uint64_t)((int64_t)c - minc[k] + 2);
        wlo = (wlo << kp.bits) | (f < eta ? 1ull : 0ull);
        whi = (whi << kp.bits) | (f >= 1.0 - eta ? 1ull : 0ull);
    }
    *lo = key - wlo;
    *hi = key + whi;
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
