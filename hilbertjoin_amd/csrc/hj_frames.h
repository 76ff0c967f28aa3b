// hj_frames.h -- tile-relative operand frames of the MFMA filter and the error
// band of one tile pair (shared by the host and the device).
//
// With one frame per join (operands y = (x - c) s around the grid centre c),
// the filter's error band grows with the data's extent from c relative to
// eps.  A tile whose points lie close together takes its own frame instead:
// an anchor a_T (a lattice point near the tile's bounding-box centre, in the
// scaled units of the grid) and operands y~ = round(z - a_T), z = (x - c) s.
// For a tile pair (I, J) with delta = a_J - a_I the filter then computes
//     y~_a . y~_b - (|y~_a|^2 / 2 - r) - (|y~_b|^2 / 2 + c)
//       = -|y~_a - y~_b - delta|^2 / 2 + |delta|^2 / 2,
// r = y~_a . delta and c = y~_b . delta, per K chunk: the corrected norms go
// into the norm slots of the pair's operands, and the thresholds move up by
// |delta|^2 / 2.  Since y~_a - y~_b - delta = s (x_a - x_b) up to the operand
// rounding, the band now scales with the tiles' own radii (rho) and with
// |delta|, not with the extent of the whole join.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace hj {

// Anchors are multiples of the grid's lattice step q (a power of two) with
// |a / q| <= FRAME_ANCHOR_MAX: a difference of two anchors is an integer
// times q of at most 11 (FP16) or 8 (BF16) significant bits, exact in the
// operand type, and every anchor is exact in FP32.
constexpr int FRAME_ANCHOR_MAX_F16 = 1024;
constexpr int FRAME_ANCHOR_MAX_BF16 = 128;
// tiles are staged in FP64 in LDS to find their anchor: up to this d
constexpr int FRAME_MAX_D = 64;
// a tile takes its own frame when its radius around its anchor is at most
// FRAME_GAIN times its radius around the grid centre
constexpr double FRAME_GAIN = 0.5;
// relative width of the grid-frame band (outer edge over eps, minus 1) above
// which the fast path also tests in the tiles' own frames: measured on
// Gaussian mixtures, FP16 bands (0.09-0.2) leave few tile pairs a tighter band
// would decide, BF16 bands (0.27-0.75) many
constexpr double FRAME_FAST_BAND = 0.25;

// Per-segment constants of the tile-pair band (frame_consts in hj_join.cpp).
struct FrameConst {
    float sek_in, sek_out;  // s eps sqrt(1 -+ mu)(1 -+ 2^-30), rounded inward / outward
    float u;                // operand unit roundoff: error per unit of radius
    float e0;               // absolute operand error (flushed subnormals, FP64 rounding)
    float cP;               // accumulation, norm-split and correction error per unit of P
    float c0;               // absolute accumulated error
    float cd;               // error of the computed |delta|^2 per unit of it
    float pad_;
};

// Classification thresholds of one tile pair: rho_a, rho_b bound |y~| over the
// two tiles (each in its own frame), dd_bound bounds |delta|^2 over the K
// slots the test has accumulated (it enters the error bound) and dd_shift is
// the computed |delta|^2 of those slots (the threshold shift).
//   acc <  t_out   certainly farther than eps
//   acc >= t_in    certainly within eps
// Bound (S = rho_a + rho_b, D = |delta|): the operands' rounding moves the
// computed distance by at most e = u S + e0; the FP32 accumulation of the K
// products, the norm split, the corrections r and c and the corrected norms'
// re-split add at most cP P + c0 with P = S^2 / 2 + S D >= the sum of the
// absolute values of all products; the computed |delta|^2 is within cd
// |delta|^2 of the exact one.  The thresholds are then moved outward by
// 2^-19 of their terms' magnitude, more than the FP32 rounding of this
// evaluation (host and device round alike: each operation once, to nearest).
__host__ __device__ inline void frame_thresholds(const FrameConst &f, float rho_a, float rho_b,
                                                 float dd_bound, float dd_shift, float *t_in,
                                                 float *t_out) {
    const float S = rho_a + rho_b;
    const float D = sqrtf(dd_bound) * (1.0f + 0x1p-20f);
    const float P = 0.5f * S * S + S * D;
    const float e = f.u * S + f.e0;
    const float Er = f.cP * P + f.c0 + f.cd * dd_bound;
    const float half_shift = 0.5f * dd_shift;
    const float tho = f.sek_out + e;
    const float qo = 0.5f * tho * tho;
    *t_out = (half_shift - qo - Er) - (half_shift + qo + Er) * 0x1p-19f;
    const float thi = f.sek_in - e;
    if (thi > 0.0f) {
        const float qi = 0.5f * thi * thi;
        *t_in = (half_shift - qi + Er) + (half_shift + qi + Er) * 0x1p-19f;
    } else {
        *t_in = INFINITY;
    }
}

}  // namespace hj
