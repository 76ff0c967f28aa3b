// hj_api.cpp -- C ABI of the MI355X epsilon join: contexts, the join
// pipeline and its phase timings.
//
// Pipeline of one join (reference: EGO_PARALLEL_TRAN egojoin.h:164-222 and
// EGO_PARALLEL_TRAN_NOSELF :395-479 set up, EGO_LOOP_TRAN :311-392 runs):
//   1. grid     per-dimension coordinate range -> minimum cell, bits per
//               field, number of key dimensions; 64-bit key per point
//   2. sort     LSD radix sort of (key, index)
//   3. index    gather sorted rows / transposed tiles, per-tile key bounds,
//               per-tile candidate intervals, tile-pair list, Hilbert order
//   4. join     persistent distance-count kernel over the tile-pair list
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>

#include "hj_internal.h"

struct hj_ctx {
    hj::Ctx c;
};

using namespace hj;

namespace {

// Active dimensions of the stripe enumeration when the caller does not fix
// them: the largest a <= 6 whose key-prefix classes (points sharing their
// first a cells) still hold AUTO_CLASS_POINTS points on average in the target
// set -- deeper prefixes cut no further once a class fits in one or two
// tiles, while the enumeration grows as 3^a.
constexpr int AUTO_CLASS_POINTS = 128;
constexpr int DEFAULT_KBLOCK = 8;

int dpad_for(int d) { return (d + 7) / 8 * 8; }

size_t ntiles_of(size_t n) { return (n + 63) / 64; }

int ipow3(int a) {
    int r = 1;
    for (int k = 0; k < a; k++) r *= 3;
    return r;
}

struct Opts {
    int kernel = HJ_KERNEL_FP64;
    int active = -1;
    int order = HJ_ORDER_ROW;
    int kblock = DEFAULT_KBLOCK;
    int grid = 0;
    int split = 0;  // filter: first-chunk dimensions (0 auto, -1 no early exit)
};

Opts read_opts(const hj_opts *o) {
    Opts r;
    if (!o) return r;
    r.kernel = o->kernel;
    r.active = o->active_dims;
    r.order = o->order;
    r.kblock = o->kblock == 0 ? DEFAULT_KBLOCK : o->kblock;
    r.grid = o->grid;
    r.split = o->filter_split;
    if (r.kernel < HJ_KERNEL_FP64 || r.kernel > HJ_KERNEL_MFMA_BF16)
        throw InvalidArg{"unknown kernel"};
    if (r.order != HJ_ORDER_HILBERT && r.order != HJ_ORDER_ROW) throw InvalidArg{"unknown order"};
    if (r.kblock != 4 && r.kblock != 8) throw InvalidArg{"kblock must be 4 or 8"};
    if (r.active > 6) throw InvalidArg{"active_dims must be in [-1, 6]"};
    if (r.split < -1) throw InvalidArg{"filter_split must be >= -1"};
    return r;
}

void check_sizes(uint64_t n, uint32_t d) {
    if (d == 0) throw InvalidArg{"d must be >= 1"};
    if (d > 65536) throw InvalidArg{"d must be <= 65536"};
    if (n > 0xFFFFFF00ull) throw InvalidArg{"n must be < 2^32 - 256"};
}

struct Phase {
    hipEvent_t *ev;
    hipStream_t s;
    void mark(int k) { HJ_HIP(hipEventRecord(ev[k], s)); }
    float ms(int a, int b) {
        float t = 0;
        HJ_HIP(hipEventElapsedTime(&t, ev[a], ev[b]));
        return t;
    }
};

// ---- 1. grid ---------------------------------------------------------------
Grid make_grid(Ctx &c, const double *A, size_t nA, const double *B, size_t nB, int d,
               double eps, int active_req) {
    if (!(eps > 0) || !std::isfinite(eps)) throw InvalidArg{"eps must be finite and > 0"};
    Grid g;
    g.eps = eps;
    g.r = 1.0 / eps;
    g.d = d;
    if (!std::isfinite(g.r)) throw InvalidArg{"1/eps is not finite"};
    hipStream_t s = c.stream;
    // min/max of both sets, then (upload_minc) the minimum cells
    c.minmax.ensure((size_t)4 * d * sizeof(double) + 64 * sizeof(int64_t));
    double *mm = c.minmax.as<double>();
    c.flags.ensure(64);
    uint32_t *bad = c.flags.as<uint32_t>();
    HJ_HIP(hipMemsetAsync(bad, 0, 4, s));
    if (nA) launch_minmax(A, nA, d, mm, mm + d, bad, c.scratch_scan, s);
    if (B && nB) launch_minmax(B, nB, d, mm + 2 * d, mm + 3 * d, bad, c.scratch_scan, s);
    std::vector<double> h((size_t)4 * d);
    uint32_t hbad = 0;
    HJ_HIP(hipMemcpyAsync(h.data(), mm, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HJ_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
    HJ_HIP(hipStreamSynchronize(s));
    if (hbad) throw InvalidArg{"input contains a non-finite coordinate"};
    std::vector<double> lo(d), hi(d);
    for (int k = 0; k < d; k++) {
        double l = INFINITY, u = -INFINITY;
        if (nA) {
            l = std::min(l, h[k]);
            u = std::max(u, h[d + k]);
        }
        if (B && nB) {
            l = std::min(l, h[2 * d + k]);
            u = std::max(u, h[3 * d + k]);
        }
        lo[k] = l;
        hi[k] = u;
    }
    int64_t range = 0;
    std::vector<int64_t> cmin(d);
    for (int k = 0; k < d; k++) {
        double cl = std::floor(lo[k] * g.r), ch = std::floor(hi[k] * g.r);
        if (!(std::fabs(cl) < 4e18) || !(std::fabs(ch) < 4e18) || ch - cl > 1e18)
            throw InvalidArg{"eps is too small for the coordinate range"};
        cmin[k] = (int64_t)cl;
        range = std::max(range, (int64_t)ch - (int64_t)cl);
    }
    if (range > ((int64_t)1 << 58)) throw InvalidArg{"eps is too small for the coordinate range"};
    int bits = 1;
    while (((uint64_t)1 << bits) <= (uint64_t)range + 4) bits++;
    g.bits = bits;
    g.key_dims = std::min(d, 64 / bits);
    g.minc.assign(cmin.begin(), cmin.begin() + g.key_dims);
    g.active = active_req < 0 ? -1 : std::max(0, std::min(active_req, std::min(6, g.key_dims)));
    // Filter operands: y = (x - c) * s with c the per-dimension centre and s a
    // power of two such that every |y| and every norm |y| stays <= 2^14.
    // mbound bounds |x - c| over all points (with a 2^-40 relative margin for
    // the rounding of this host computation).
    g.centre.resize(d);
    double m2 = 0;
    for (int k = 0; k < d; k++) {
        g.centre[k] = 0.5 * lo[k] + 0.5 * hi[k];
        double a = std::max(hi[k] - g.centre[k], g.centre[k] - lo[k]);
        m2 += a * a;
    }
    g.mbound = std::sqrt(m2) * (1.0 + std::ldexp(1.0, -40));
    g.scale = 1.0;
    if (g.mbound > 0) {
        int e;
        std::frexp(g.mbound, &e);  // mbound in [2^(e-1), 2^e)
        g.scale = std::ldexp(1.0, 14 - e);
    }
    return g;
}

int64_t *upload_minc(Ctx &c, const Grid &g) {
    // minc lives after the 4*d doubles of the min/max scratch
    size_t off = (size_t)4 * g.d * sizeof(double);
    c.minmax.ensure(off + 64 * sizeof(int64_t));
    // ensure() may reallocate only when growing; the min/max values are no
    // longer needed here
    int64_t *dst = reinterpret_cast<int64_t *>(c.minmax.as<char>() + off);
    HJ_HIP(hipMemcpyAsync(dst, g.minc.data(), g.minc.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice, c.stream));
    return dst;
}

// ---- 2. + 3a. keys, sort, gather, tile bounds ------------------------------------
struct Prepared {
    size_t n = 0, ntiles = 0;
    int dpad = 0;
    const uint64_t *keys = nullptr;
    const uint32_t *perm = nullptr;
    const double *rows = nullptr, *cols = nullptr;
    const uint64_t *tlo = nullptr, *thi = nullptr;
};

void make_keys(Ctx &c, SetBufs &S, const double *pts, size_t n, const Grid &g,
               const int64_t *minc, Prepared &P) {
    hipStream_t s = c.stream;
    P.n = n;
    P.ntiles = ntiles_of(n);
    P.dpad = dpad_for(g.d);
    size_t nk = std::max<size_t>(n, 1);
    for (int b = 0; b < 2; b++) {
        S.keys[b].ensure(nk * sizeof(uint64_t));
        S.vals[b].ensure(nk * sizeof(uint32_t));
    }
    launch_keys(pts, n, g.d, g, minc, S.keys[0].as<uint64_t>(), S.vals[0].as<uint32_t>(), s);
}

void sort_keys(Ctx &c, SetBufs &S, const Grid &g, Prepared &P) {
    hipStream_t s = c.stream;
    size_t n = P.n;
    uint64_t *kk[2] = {S.keys[0].as<uint64_t>(), S.keys[1].as<uint64_t>()};
    uint32_t *vv[2] = {S.vals[0].as<uint32_t>(), S.vals[1].as<uint32_t>()};
    int which = 0;
    radix_sort_u64_u32(kk, vv, n, g.bits * g.key_dims, &which, c.scratch_hist, c.scratch_scan, s);
    S.cur = which;
    P.keys = kk[which];
    P.perm = vv[which];
}

void gather_and_bound(Ctx &c, SetBufs &S, const double *pts, const Grid &g, const int64_t *minc,
                      Prepared &P) {
    hipStream_t s = c.stream;
    size_t nrows = std::max<size_t>(P.ntiles, 1) * 64;
    S.rows.ensure(nrows * P.dpad * sizeof(double));
    S.cols.ensure(nrows * P.dpad * sizeof(double));
    S.tlo.ensure(std::max<size_t>(P.ntiles, 1) * sizeof(uint64_t));
    S.thi.ensure(std::max<size_t>(P.ntiles, 1) * sizeof(uint64_t));
    launch_gather(pts, P.perm, P.n, g.d, P.dpad, S.rows.as<double>(), S.cols.as<double>(), s);
    launch_tile_bounds(S.cols.as<double>(), P.n, P.dpad, g, minc, S.tlo.as<uint64_t>(),
                       S.thi.as<uint64_t>(), s);
    P.rows = S.rows.as<double>();
    P.cols = S.cols.as<double>();
    P.tlo = S.tlo.as<uint64_t>();
    P.thi = S.thi.as<uint64_t>();
}

// ---- 3b. active dimensions (when not fixed by the caller) -----------------------
void choose_active(Ctx &c, Grid &g, const Prepared &T) {
    if (g.active >= 0) return;
    c.counters.ensure(8 * sizeof(uint64_t));
    uint64_t *cnt = c.counters.as<uint64_t>();
    launch_prefix_classes(T.keys, T.n, g.key_dims, g.bits, cnt, c.num_cus, c.stream);
    uint64_t h[8];
    HJ_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    HJ_HIP(hipStreamSynchronize(c.stream));
    int a = 0;
    for (int k = 1; k <= std::min(6, g.key_dims); k++)
        if (h[k] > 0 && (double)T.n / (double)h[k] >= AUTO_CLASS_POINTS) a = k;
    g.active = a;
}

// ---- 3c. work list ---------------------------------------------------------------
struct WorkList {
    const uint64_t *pairs = nullptr;
    uint64_t npairs = 0;
};

WorkList build_worklist(Ctx &c, const Prepared &Q, const Prepared &T, const Grid &g, bool self,
                        int order) {
    hipStream_t s = c.stream;
    WorkList wl;
    size_t ntq = Q.ntiles;
    if (ntq == 0 || T.n == 0) return wl;
    int total = ipow3(g.active);
    int ncodes = self ? total - (total - 1) / 2 : total;
    int max_iv = ncodes;
    c.ivals.ensure(ntq * (size_t)max_iv * 2 * sizeof(uint32_t));
    c.icount.ensure(ntq * sizeof(uint32_t));
    c.ioff.ensure((2 * ntq + 2) * sizeof(uint64_t));
    uint64_t *jcount = c.ioff.as<uint64_t>();
    uint64_t *off = jcount + ntq + 1;
    launch_zero_u64(jcount, ntq + 1, s);
    launch_tile_ranges(Q.tlo, Q.thi, ntq, T.keys, T.n, g, self, c.ivals.as<uint32_t>(),
                       c.icount.as<uint32_t>(), jcount, max_iv, s);
    exclusive_scan_u64(jcount, off, ntq + 1, c.scratch_scan, s);
    uint64_t np = 0;
    HJ_HIP(hipMemcpyAsync(&np, off + ntq, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HJ_HIP(hipStreamSynchronize(s));
    wl.npairs = np;
    if (np == 0) return wl;
    for (int b = 0; b < 2; b++) c.pairs[b].ensure(np * sizeof(uint64_t));
    launch_emit_pairs(c.ivals.as<uint32_t>(), c.icount.as<uint32_t>(), off, ntq, max_iv,
                      c.pairs[0].as<uint64_t>(), s);
    wl.pairs = c.pairs[0].as<uint64_t>();
    if (order == HJ_ORDER_HILBERT && np > 1) {
        size_t side = std::max(ntq, T.ntiles);
        int L = 0;
        while (((size_t)1 << L) < side) L++;
        for (int b = 0; b < 2; b++) c.hkeys[b].ensure(np * sizeof(uint64_t));
        launch_hilbert_keys(c.pairs[0].as<uint64_t>(), np, L, c.hkeys[0].as<uint64_t>(), s);
        uint64_t *kk[2] = {c.hkeys[0].as<uint64_t>(), c.hkeys[1].as<uint64_t>()};
        uint64_t *vv[2] = {c.pairs[0].as<uint64_t>(), c.pairs[1].as<uint64_t>()};
        int which = 0;
        radix_sort_u64_u64(kk, vv, np, 2 * L, &which, c.scratch_hist, c.scratch_scan, s);
        wl.pairs = vv[which];
    }
    return wl;
}

// requested persistent grid (0: sized from the kernel's occupancy)
int default_grid(Ctx &, const Opts &o) { return o.grid > 0 ? o.grid : 0; }

// ---- reduced-precision filter: thresholds and operands -------------------------
float round_up_f(double x) {
    if (std::isinf(x)) return (float)x;
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

float round_down_f(double x) {
    if (std::isinf(x)) return (float)x;
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}

// Worst-case classification thresholds of the MFMA filter.
//
// Scaled operands y = (x - c) s (s a power of two, |y| <= M = s * mbound),
// rounded to the operand type with unit roundoff u (double -> float -> half:
// u = 2^-11 (1 + 2^-12); bfloat16: 2^-8 (1 + 2^-15)); subnormal operands may
// be flushed, an absolute error <= 2^-14 (half) per coordinate.
//   E1: rho = |a^ - b^| differs from the exact scaled distance D by at most
//       e = 2 u M + 2 sqrt(d) 2^-14 + 2^-49 M (the last term: FP64 rounding
//       of x - c).
//   FP64 oracle: its sum differs from the exact |x_a - x_b|^2 by a relative
//       (d+2) 2^-52 and eps*eps by 2^-53, so pairs with exact distance below
//       s eps sqrt(1-mu) are "in" for it and above s eps sqrt(1+mu) "out".
//   E2: the MFMA accumulates K exact products in FP32; bounded by
//       (K+1) 2^-24 sum|products| <= (K+1) 2^-24 2 M^2, doubled for safety.
//   R:  the squared norms folded in as 2 (half) or 3 (bfloat16) parts leave a
//       residual <= 2^-22 (resp. 2^-24) |y|^2 per side, plus the FP64 sum of
//       squares (d 2^-53 M^2) and flushed tiny parts.
// acc = -rho^2/2 + r + E with |r| + |E| <= Er, so
//   acc >= -theta_in^2/2 + Er   =>  rho + e <= theta_in + e, pair in
//   acc <  -theta_out^2/2 - Er  =>  D > s eps sqrt(1+mu),  pair out.
FilterParams filter_thresholds(const Grid &g, int d, bool bf16, int split_req) {
    FilterParams fp{};
    fp.bf16 = bf16 ? 1 : 0;
    fp.parts = bf16 ? 3 : 2;
    // K chunks: the first k-step holds 16 - 2*parts dimensions and their norm
    // slots; the early exit after it pays when enough dimensions follow.
    const int first = 16 - 2 * fp.parts;
    int split = d;
    if (split_req > 0)
        split = std::min(split_req, d);
    else if (split_req == 0 && d >= 2 * first + 4)
        split = first;
    fp.split = split;
    fp.steps1 = (split + 2 * fp.parts + 15) / 16;
    fp.steps = fp.steps1 + (split < d ? (d - split + 2 * fp.parts + 15) / 16 : 0);
    const double u = bf16 ? std::ldexp(1.0, -8) * (1 + std::ldexp(1.0, -15))
                          : std::ldexp(1.0, -11) * (1 + std::ldexp(1.0, -12));
    const double flush = bf16 ? std::ldexp(1.0, -100) : std::ldexp(1.0, -14);
    const double M = g.scale * g.mbound;
    const double e = 2 * u * M + 2 * std::sqrt((double)d) * flush + std::ldexp(1.0, -49) * M;
    const double mu = (d + 4) * std::ldexp(1.0, -52) + std::ldexp(1.0, -50);
    const double K = fp.steps * 16.0;
    const double E2 = 4 * (K + 1) * std::ldexp(1.0, -24) * M * M;
    const double R = (bf16 ? std::ldexp(1.0, -20) : std::ldexp(1.0, -18)) * M * M +
                     d * std::ldexp(1.0, -50) * M * M + 128.0;
    const double Er = E2 + R;
    const double se = g.scale * g.eps;
    const double th_in = se * std::sqrt(1 - mu) * (1 - std::ldexp(1.0, -30)) - e;
    const double th_out = se * std::sqrt(1 + mu) * (1 + std::ldexp(1.0, -30)) + e;
    const double t_in = th_in > 0 ? -0.5 * th_in * th_in + Er : INFINITY;
    const double t_out = -0.5 * th_out * th_out - Er;
    fp.t_in = round_up_f(t_in * (t_in < 0 ? 1 - std::ldexp(1.0, -40) : 1 + std::ldexp(1.0, -40)));
    fp.t_out = round_down_f(t_out * (1 + std::ldexp(1.0, -40)));
    return fp;
}

void pack_operands(Ctx &c, SetBufs &S, const Prepared &P, const Grid &g, const FilterParams &fp) {
    hipStream_t s = c.stream;
    const size_t ntiles = std::max<size_t>(P.ntiles, 1);
    const size_t bytes = ntiles * 2 * (size_t)fp.steps * 64 * 16;
    S.frow.ensure(bytes);
    S.fcol.ensure(bytes);
    S.norms.ensure(2 * std::max<size_t>(P.n, 1) * sizeof(double));
    const double *centre = c.centre.as<double>();
    launch_op_norms(P.rows, P.n, g.d, P.dpad, fp.split, centre, g.scale, fp.bf16 != 0,
                    S.norms.as<double>(), s);
    launch_frag_pack(P.rows, P.n, g.d, P.dpad, centre, S.norms.as<double>(), g.scale, fp,
                     S.frow.p, S.fcol.p, s);
}

// ---- the whole join ----------------------------------------------------------------
struct JoinResult {
    uint64_t count = 0;
    uint64_t counters[8] = {};
    WorkList wl;
    Prepared PA, PB;
    Grid g;
};

void prep_filter(Ctx &c, SetBufs &S, const double *pts, Prepared &P, const Grid &g,
                 const int64_t *minc, const FilterParams &fp) {
    const size_t ntiles = std::max<size_t>(P.ntiles, 1);
    const size_t bytes = ntiles * 2 * (size_t)fp.steps * 64 * 16;
    S.frow.ensure(bytes);
    S.fcol.ensure(bytes);
    S.tlo.ensure(ntiles * sizeof(uint64_t));
    S.thi.ensure(ntiles * sizeof(uint64_t));
    launch_filter_prep(pts, P.perm, P.n, g.d, c.centre.as<double>(), g.scale, fp, g, minc,
                       S.tlo.as<uint64_t>(), S.thi.as<uint64_t>(), S.frow.p, S.fcol.p, c.stream);
    P.tlo = S.tlo.as<uint64_t>();
    P.thi = S.thi.as<uint64_t>();
}

// need_layout: the caller also wants the FP64 row / column-tile layouts (the
// pair-output pass reads them)
void run_join(hj_ctx *ctx, const double *A, uint64_t nA, const double *B, uint64_t nB, uint32_t d,
              double eps, const Opts &o, bool self, JoinResult &R, hj_stats *st,
              bool need_layout) {
    Ctx &c = ctx->c;
    hipStream_t s = c.stream;
    Phase ph{c.ev, s};
    ph.mark(0);
    Grid g = make_grid(c, A, nA, self ? nullptr : B, self ? 0 : nB, (int)d, eps, o.active);
    int64_t *minc = upload_minc(c, g);
    Prepared PA, PB;
    make_keys(c, c.A, A, nA, g, minc, PA);
    if (!self) make_keys(c, c.B, B, nB, g, minc, PB);
    ph.mark(1);
    sort_keys(c, c.A, g, PA);
    if (!self) sort_keys(c, c.B, g, PB);
    ph.mark(2);
    const bool filter = o.kernel != HJ_KERNEL_FP64;
    const bool fused = filter && filter_prep_fused((int)d);
    FilterParams fp{};
    if (filter) {
        fp = filter_thresholds(g, (int)d, o.kernel == HJ_KERNEL_MFMA_BF16, o.split);
        c.centre.ensure((size_t)d * sizeof(double));
        HJ_HIP(hipMemcpyAsync(c.centre.p, g.centre.data(), (size_t)d * sizeof(double),
                              hipMemcpyHostToDevice, s));
    }
    if (!fused || need_layout) {
        gather_and_bound(c, c.A, A, g, minc, PA);
        if (!self) gather_and_bound(c, c.B, B, g, minc, PB);
    }
    if (fused) {
        prep_filter(c, c.A, A, PA, g, minc, fp);
        if (!self) prep_filter(c, c.B, B, PB, g, minc, fp);
    } else if (filter) {
        pack_operands(c, c.A, PA, g, fp);
        if (!self) pack_operands(c, c.B, PB, g, fp);
    }
    const Prepared &T = self ? PA : PB;
    choose_active(c, g, T);
    WorkList wl = build_worklist(c, PA, T, g, self, o.order);
    ph.mark(3);
    c.counters.ensure(8 * sizeof(uint64_t));
    uint64_t *ctr = c.counters.as<uint64_t>();
    launch_zero_u64(ctr, 8, s);
    double eps2 = eps * eps;
    int grid = default_grid(c, o);
    ph.mark(5);
    if (!filter) {
        launch_fp64_count(PA.rows, T.cols, wl.pairs, wl.npairs, (uint32_t)nA, (uint32_t)T.n,
                          PA.dpad, eps2, self, o.kblock, ctr, grid, c.num_cus, s);
        ph.mark(6);
    } else {
        const void *fa = c.A.frow.p;
        const void *fb = self ? c.A.fcol.p : c.B.fcol.p;
        uint64_t cap = std::max<uint64_t>((uint64_t)1 << 20, c.band.bytes / sizeof(uint64_t));
        c.band.ensure(cap * sizeof(uint64_t));
        launch_mfma_count(fa, fb, wl.pairs, wl.npairs, (uint32_t)nA, (uint32_t)T.n, fp, self, ctr,
                          c.band.as<uint64_t>(), cap, grid, c.num_cus, s);
        ph.mark(6);
        uint64_t nband = 0;
        HJ_HIP(hipMemcpyAsync(&nband, ctr + 4, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HJ_HIP(hipStreamSynchronize(s));
        if (nband > cap) {
            // the queue overflowed: size it exactly and run the filter again
            cap = nband;
            c.band.ensure(cap * sizeof(uint64_t));
            launch_zero_u64(ctr, 8, s);
            launch_mfma_count(fa, fb, wl.pairs, wl.npairs, (uint32_t)nA, (uint32_t)T.n, fp, self,
                              ctr, c.band.as<uint64_t>(), cap, grid, c.num_cus, s);
        }
        launch_band_recheck(A, PA.perm, self ? A : B, T.perm, (int)d, eps2, c.band.as<uint64_t>(),
                            nband, ctr, s);
    }
    ph.mark(4);
    HJ_HIP(hipMemcpyAsync(R.counters, ctr, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HJ_HIP(hipStreamSynchronize(s));
    if (filter) {
        // K slots multiplied: the first chunk for every candidate pair, the rest for
        // the blocks that did not exit early (64 x 64 pairs each, upper bound)
        R.counters[2] = R.counters[3] * (uint64_t)(fp.steps1 * 16) +
                        (wl.npairs - R.counters[7]) * 4096ull * (uint64_t)((fp.steps - fp.steps1) * 16);
    }
    R.count = R.counters[1];
    R.wl = wl;
    R.PA = PA;
    R.PB = T;
    R.g = g;
    if (st) {
        st->grid_ms = ph.ms(0, 1);
        st->sort_ms = ph.ms(1, 2);
        st->index_ms = ph.ms(2, 3);
        st->join_ms = ph.ms(3, 4);
        st->kernel_ms = ph.ms(5, 6);
        st->tile_pairs = wl.npairs;
        st->candidate_pairs = R.counters[3];
        st->pair_dims = R.counters[2];
        st->band_pairs = R.counters[4];
        st->early_exits = R.counters[7];
        double denom = self ? (double)nA * ((double)nA - 1) / 2 : (double)nA * (double)nB;
        st->load_fraction = denom > 0 ? (double)R.counters[3] / denom : 0.0;
        st->key_dims = g.key_dims;
        st->key_bits = g.bits;
        st->active_dims = g.active;
        st->kernel = o.kernel;
    }
}

int set_error(hj_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->c.err = msg;
    return code;
}

template <class F> int guarded(hj_ctx *ctx, F &&f) {
    if (!ctx) return HJ_ERR_INVALID;
    try {
        ctx->c.err.clear();
        return f();
    } catch (const InvalidArg &e) {
        return set_error(ctx, HJ_ERR_INVALID, e.msg);
    } catch (const HipError &e) {
        std::string m = std::string(hipGetErrorString(e.code)) + " in " + e.where;
        return set_error(ctx, e.code == hipErrorOutOfMemory ? HJ_ERR_NOMEM : HJ_ERR_HIP, m);
    } catch (const std::bad_alloc &) {
        return set_error(ctx, HJ_ERR_NOMEM, "host allocation failed");
    } catch (...) {
        return set_error(ctx, HJ_ERR_INTERNAL, "unexpected exception");
    }
}

const double *upload(Ctx &c, DevBuf &buf, const double *host, uint64_t n, uint32_t d) {
    size_t bytes = (size_t)n * d * sizeof(double);
    buf.ensure(std::max<size_t>(bytes, 8));
    if (bytes) HJ_HIP(hipMemcpyAsync(buf.p, host, bytes, hipMemcpyHostToDevice, c.stream));
    return buf.as<double>();
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

void zero_stats(hj_stats *st) {
    if (st) std::memset(st, 0, sizeof(*st));
}

int count_common(hj_ctx *ctx, const double *A, uint64_t nA, const double *B, uint64_t nB,
                 uint32_t d, double eps, const hj_opts *opts, uint64_t *count, hj_stats *st,
                 bool self, bool host) {
    return guarded(ctx, [&]() -> int {
        double t0 = now_ms();
        zero_stats(st);
        if (!count) throw InvalidArg{"count must not be NULL"};
        Opts o = read_opts(opts);
        check_sizes(nA, d);
        if (!self) check_sizes(nB, d);
        if ((nA && !A) || (!self && nB && !B)) throw InvalidArg{"null point array"};
        *count = 0;
        if (nA == 0 || (!self && nB == 0) || (self && nA < 2)) {
            if (st) st->total_ms = now_ms() - t0;
            return HJ_OK;
        }
        Ctx &c = ctx->c;
        const double *dA = A, *dB = B;
        if (host) {
            dA = upload(c, c.A.raw, A, nA, d);
            if (!self) dB = upload(c, c.B.raw, B, nB, d);
        }
        JoinResult R;
        run_join(ctx, dA, nA, dB, nB, d, eps, o, self, R, st, false);
        *count = R.count;
        if (st) st->total_ms = now_ms() - t0;
        return HJ_OK;
    });
}

// Pair lists: count, then one fill pass of the FP64 pair kernel (sorted
// indices), mapped to original indices (mirrored for self-joins), radix
// sorted by (row << 32 | col), and turned into rowptr / colidx on the device.
int csr_common(hj_ctx *ctx, const double *A, uint64_t nA, const double *B, uint64_t nB,
               uint32_t d, double eps, const hj_opts *opts, int64_t *rowptr, int64_t *colidx,
               uint64_t cap, uint64_t *total, hj_stats *st, bool self) {
    return guarded(ctx, [&]() -> int {
        double t0 = now_ms();
        zero_stats(st);
        Opts o = read_opts(opts);
        check_sizes(nA, d);
        if (!self) check_sizes(nB, d);
        if (!rowptr || !total) throw InvalidArg{"rowptr and total must not be NULL"};
        if ((nA && !A) || (!self && nB && !B)) throw InvalidArg{"null point array"};
        *total = 0;
        if (nA == 0 || (!self && nB == 0) || (self && nA < 2)) {
            for (uint64_t i = 0; i <= nA; i++) rowptr[i] = 0;
            if (st) st->total_ms = now_ms() - t0;
            return HJ_OK;
        }
        Ctx &c = ctx->c;
        hipStream_t s = c.stream;
        const double *dA = upload(c, c.A.raw, A, nA, d);
        const double *dB = self ? nullptr : upload(c, c.B.raw, B, nB, d);
        JoinResult R;
        run_join(ctx, dA, nA, dB, nB, d, eps, o, self, R, st, true);
        const uint64_t np = R.count;
        const uint64_t nk = self ? 2 * np : np;
        *total = nk;
        int64_t *drow = nullptr;
        c.csr_rows.ensure((nA + 1) * sizeof(int64_t));
        drow = c.csr_rows.as<int64_t>();
        if (nk == 0) {
            launch_zero_u64(reinterpret_cast<uint64_t *>(drow), nA + 1, s);
        } else {
            c.band.ensure(np * sizeof(uint64_t));
            uint64_t *raw = c.band.as<uint64_t>();
            uint64_t *ctr = c.counters.as<uint64_t>();
            launch_zero_u64(ctr, 8, s);
            launch_fp64_pairs(R.PA.rows, R.PB.cols, R.wl.pairs, R.wl.npairs, (uint32_t)nA,
                              (uint32_t)R.PB.n, R.PA.dpad, eps * eps, self, ctr, raw, np,
                              default_grid(c, o), c.num_cus, s);
            // keys / vals ping-pong in the hkeys buffers and the B set's sort buffers
            for (int b = 0; b < 2; b++) c.hkeys[b].ensure(nk * sizeof(uint64_t));
            c.csr_cols.ensure(2 * nk * sizeof(uint32_t));
            uint64_t *kk[2] = {c.hkeys[0].as<uint64_t>(), c.hkeys[1].as<uint64_t>()};
            uint32_t *vv[2] = {c.csr_cols.as<uint32_t>(), c.csr_cols.as<uint32_t>() + nk};
            launch_map_pairs(raw, np, R.PA.perm, R.PB.perm, self, kk[0], vv[0], s);
            int rb = 1;
            while (((uint64_t)1 << rb) < std::max<uint64_t>(nA, R.PB.n)) rb++;
            int which = 0;
            radix_sort_u64_u32(kk, vv, nk, 32 + rb, &which, c.scratch_hist, c.scratch_scan, s);
            launch_csr_rowptr(kk[which], nk, nA, drow, s);
            if (colidx && nk <= cap) {
                // colidx is staged in the other key buffer
                int64_t *dcol = reinterpret_cast<int64_t *>(kk[which ^ 1]);
                launch_csr_cols(kk[which], nk, dcol, s);
                HJ_HIP(hipMemcpyAsync(colidx, dcol, nk * sizeof(int64_t), hipMemcpyDeviceToHost,
                                      s));
            }
        }
        HJ_HIP(hipMemcpyAsync(rowptr, drow, (nA + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        uint64_t written = 0;
        HJ_HIP(hipMemcpyAsync(&written, c.counters.as<uint64_t>() + 6, sizeof(uint64_t),
                              hipMemcpyDeviceToHost, s));
        HJ_HIP(hipStreamSynchronize(s));
        if (nk && written != np) throw InvalidArg{"internal: pair pass disagrees with count"};
        if (st) st->total_ms = now_ms() - t0;
        if (nk > cap && colidx) return set_error(ctx, HJ_ERR_CAPACITY, "colidx too small");
        return HJ_OK;
    });
}

}  // namespace

extern "C" {

HJ_API int hj_self_join_csr(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d, double eps,
                            const hj_opts *opts, int64_t *rowptr, int64_t *colidx,
                            uint64_t cap, uint64_t *total, hj_stats *stats) {
    return csr_common(ctx, pts, n, nullptr, 0, d, eps, opts, rowptr, colidx, cap, total, stats,
                      true);
}

HJ_API int hj_join_csr(hj_ctx *ctx, const double *A, uint64_t nA, const double *B, uint64_t nB,
                       uint32_t d, double eps, const hj_opts *opts, int64_t *rowptr,
                       int64_t *colidx, uint64_t cap, uint64_t *total, hj_stats *stats) {
    return csr_common(ctx, A, nA, B, nB, d, eps, opts, rowptr, colidx, cap, total, stats, false);
}

HJ_API void hj_opts_default(hj_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->kernel = HJ_KERNEL_FP64;
    o->active_dims = -1;
    o->order = HJ_ORDER_ROW;
    o->kblock = DEFAULT_KBLOCK;
    o->grid = 0;
}

HJ_API int hj_ctx_create(hj_ctx **out, int device) {
    if (!out) return HJ_ERR_INVALID;
    *out = nullptr;
    hj_ctx *ctx = new (std::nothrow) hj_ctx();
    if (!ctx) return HJ_ERR_NOMEM;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        delete ctx;
        return HJ_ERR_HIP;
    }
    if (device < 0 || device >= ndev) {
        delete ctx;
        return HJ_ERR_INVALID;
    }
    ctx->c.device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->c.stream) != hipSuccess) {
        delete ctx;
        return HJ_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->c.num_cus = prop.multiProcessorCount;
        ctx->c.name = prop.gcnArchName;
    }
    for (auto &e : ctx->c.ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return HJ_ERR_HIP;
        }
    }
    *out = ctx;
    return HJ_OK;
}

HJ_API int hj_ctx_destroy(hj_ctx *ctx) {
    if (!ctx) return HJ_OK;
    Ctx &c = ctx->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (SetBufs *S : {&c.A, &c.B}) {
        S->raw.release();
        for (int b = 0; b < 2; b++) {
            S->keys[b].release();
            S->vals[b].release();
        }
        S->rows.release();
        S->cols.release();
        S->frow.release();
        S->fcol.release();
        S->norms.release();
        S->tlo.release();
        S->thi.release();
        S->rowptr.release();
    }
    for (DevBuf *b : {&c.minmax, &c.scratch_hist, &c.scratch_scan, &c.flags, &c.ivals, &c.icount,
                      &c.ioff, &c.pairs[0], &c.pairs[1], &c.hkeys[0], &c.hkeys[1], &c.counters,
                      &c.band, &c.csr_rows, &c.csr_cols, &c.centre})
        b->release();
    for (auto &e : c.ev)
        if (e) (void)hipEventDestroy(e);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete ctx;
    return HJ_OK;
}

HJ_API const char *hj_last_error(const hj_ctx *ctx) {
    if (!ctx) return "null context";
    return ctx->c.err.c_str();
}

HJ_API int hj_device_info(hj_ctx *ctx, int *num_cus, char *name, int name_len) {
    if (!ctx) return HJ_ERR_INVALID;
    if (num_cus) *num_cus = ctx->c.num_cus;
    if (name && name_len > 0) {
        std::strncpy(name, ctx->c.name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    return HJ_OK;
}

HJ_API int hj_device_alloc(hj_ctx *ctx, uint64_t bytes, void **ptr) {
    return guarded(ctx, [&]() -> int {
        if (!ptr) throw InvalidArg{"ptr must not be NULL"};
        *ptr = nullptr;
        HJ_HIP(hipSetDevice(ctx->c.device));
        HJ_HIP(hipMalloc(ptr, std::max<uint64_t>(bytes, 8)));
        return HJ_OK;
    });
}

HJ_API int hj_device_free(hj_ctx *ctx, void *ptr) {
    return guarded(ctx, [&]() -> int {
        if (ptr) {
            HJ_HIP(hipStreamSynchronize(ctx->c.stream));
            HJ_HIP(hipFree(ptr));
        }
        return HJ_OK;
    });
}

HJ_API int hj_memcpy_to_device(hj_ctx *ctx, void *dst_dev, const void *src_host, uint64_t bytes) {
    return guarded(ctx, [&]() -> int {
        if (bytes && (!dst_dev || !src_host)) throw InvalidArg{"null pointer"};
        if (bytes)
            HJ_HIP(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, ctx->c.stream));
        HJ_HIP(hipStreamSynchronize(ctx->c.stream));
        return HJ_OK;
    });
}

HJ_API int hj_memcpy_to_host(hj_ctx *ctx, void *dst_host, const void *src_dev, uint64_t bytes) {
    return guarded(ctx, [&]() -> int {
        if (bytes && (!dst_host || !src_dev)) throw InvalidArg{"null pointer"};
        if (bytes)
            HJ_HIP(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
        HJ_HIP(hipStreamSynchronize(ctx->c.stream));
        return HJ_OK;
    });
}

HJ_API int hj_device_copy(hj_ctx *ctx, void *dst_dev, const void *src_dev, uint64_t bytes) {
    return guarded(ctx, [&]() -> int {
        if (bytes % 16) throw InvalidArg{"bytes must be a multiple of 16"};
        if (bytes && (!dst_dev || !src_dev)) throw InvalidArg{"null pointer"};
        launch_copy16(src_dev, dst_dev, bytes, ctx->c.num_cus, ctx->c.stream);
        HJ_HIP(hipStreamSynchronize(ctx->c.stream));
        return HJ_OK;
    });
}

HJ_API int hj_self_join_count(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d,
                              double eps, const hj_opts *opts, uint64_t *count,
                              hj_stats *stats) {
    return count_common(ctx, pts, n, nullptr, 0, d, eps, opts, count, stats, true, true);
}

HJ_API int hj_join_count(hj_ctx *ctx, const double *A, uint64_t nA, const double *B,
                         uint64_t nB, uint32_t d, double eps, const hj_opts *opts,
                         uint64_t *count, hj_stats *stats) {
    return count_common(ctx, A, nA, B, nB, d, eps, opts, count, stats, false, true);
}

HJ_API int hj_self_join_count_device(hj_ctx *ctx, const double *pts_dev, uint64_t n,
                                     uint32_t d, double eps, const hj_opts *opts,
                                     uint64_t *count, hj_stats *stats) {
    return count_common(ctx, pts_dev, n, nullptr, 0, d, eps, opts, count, stats, true, false);
}

HJ_API int hj_join_count_device(hj_ctx *ctx, const double *A_dev, uint64_t nA,
                                const double *B_dev, uint64_t nB, uint32_t d, double eps,
                                const hj_opts *opts, uint64_t *count, hj_stats *stats) {
    return count_common(ctx, A_dev, nA, B_dev, nB, d, eps, opts, count, stats, false, false);
}

HJ_API int hj_ego_sort(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d, double eps,
                       uint32_t *perm, uint64_t *keys, int32_t *key_dims, int32_t *key_bits) {
    return guarded(ctx, [&]() -> int {
        check_sizes(n, d);
        if (n && (!pts || !perm)) throw InvalidArg{"null pointer"};
        Ctx &c = ctx->c;
        const double *dp = upload(c, c.A.raw, pts, n, d);
        Grid g = make_grid(c, dp, n, nullptr, 0, (int)d, eps, -1);
        int64_t *minc = upload_minc(c, g);
        Prepared P;
        if (n) {
            make_keys(c, c.A, dp, n, g, minc, P);
            sort_keys(c, c.A, g, P);
        }
        if (n) {
            HJ_HIP(hipMemcpyAsync(perm, P.perm, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c.stream));
            if (keys)
                HJ_HIP(hipMemcpyAsync(keys, P.keys, n * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                      c.stream));
        }
        HJ_HIP(hipStreamSynchronize(c.stream));
        if (key_dims) *key_dims = g.key_dims;
        if (key_bits) *key_bits = g.bits;
        return HJ_OK;
    });
}

static int tile_pairs_common(hj_ctx *ctx, const double *A, uint64_t nA, const double *B,
                             uint64_t nB, uint32_t d, double eps, const hj_opts *opts,
                             uint32_t *pairs, uint64_t cap, uint64_t *npairs, uint32_t *permA,
                             uint32_t *permB, bool self) {
    return guarded(ctx, [&]() -> int {
        Opts o = read_opts(opts);
        check_sizes(nA, d);
        if (!self) check_sizes(nB, d);
        if (!npairs) throw InvalidArg{"npairs must not be NULL"};
        *npairs = 0;
        if (nA == 0 || (!self && nB == 0)) return HJ_OK;
        Ctx &c = ctx->c;
        const double *dA = upload(c, c.A.raw, A, nA, d);
        const double *dB = self ? dA : upload(c, c.B.raw, B, nB, d);
        Grid g = make_grid(c, dA, nA, self ? nullptr : dB, self ? 0 : nB, (int)d, eps, o.active);
        int64_t *minc = upload_minc(c, g);
        Prepared PA, PB;
        make_keys(c, c.A, dA, nA, g, minc, PA);
        sort_keys(c, c.A, g, PA);
        gather_and_bound(c, c.A, dA, g, minc, PA);
        if (!self) {
            make_keys(c, c.B, dB, nB, g, minc, PB);
            sort_keys(c, c.B, g, PB);
            gather_and_bound(c, c.B, dB, g, minc, PB);
        }
        choose_active(c, g, self ? PA : PB);
        WorkList wl = build_worklist(c, PA, self ? PA : PB, g, self, o.order);
        *npairs = wl.npairs;
        if (permA)
            HJ_HIP(hipMemcpyAsync(permA, PA.perm, nA * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c.stream));
        if (!self && permB)
            HJ_HIP(hipMemcpyAsync(permB, PB.perm, nB * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c.stream));
        if (wl.npairs > cap) {
            HJ_HIP(hipStreamSynchronize(c.stream));
            return set_error(ctx, HJ_ERR_CAPACITY, "pair buffer too small");
        }
        if (wl.npairs && pairs)
            HJ_HIP(hipMemcpyAsync(pairs, wl.pairs, wl.npairs * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, c.stream));
        HJ_HIP(hipStreamSynchronize(c.stream));
        return HJ_OK;
    });
}

HJ_API int hj_self_tile_pairs(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d,
                              double eps, const hj_opts *opts, uint32_t *pairs, uint64_t cap,
                              uint64_t *npairs, uint32_t *perm) {
    return tile_pairs_common(ctx, pts, n, nullptr, 0, d, eps, opts, pairs, cap, npairs, perm,
                             nullptr, true);
}

HJ_API int hj_join_tile_pairs(hj_ctx *ctx, const double *A, uint64_t nA, const double *B,
                              uint64_t nB, uint32_t d, double eps, const hj_opts *opts,
                              uint32_t *pairs, uint64_t cap, uint64_t *npairs, uint32_t *permA,
                              uint32_t *permB) {
    return tile_pairs_common(ctx, A, nA, B, nB, d, eps, opts, pairs, cap, npairs, permA, permB,
                             false);
}

HJ_API int hj_hilbert_index(hj_ctx *ctx, int order, const uint32_t *x, const uint32_t *y,
                            uint64_t n, uint64_t *out) {
    return guarded(ctx, [&]() -> int {
        if (order < 0 || order > 31) throw InvalidArg{"order must be in [0, 31]"};
        if (n && (!x || !y || !out)) throw InvalidArg{"null pointer"};
        if (n == 0) return HJ_OK;
        Ctx &c = ctx->c;
        c.hkeys[0].ensure(n * 8);
        c.hkeys[1].ensure(n * 8);
        uint32_t *dx = c.hkeys[0].as<uint32_t>(), *dy = dx + n;
        HJ_HIP(hipMemcpyAsync(dx, x, n * 4, hipMemcpyHostToDevice, c.stream));
        HJ_HIP(hipMemcpyAsync(dy, y, n * 4, hipMemcpyHostToDevice, c.stream));
        launch_hilbert_xy(dx, dy, n, order, c.hkeys[1].as<uint64_t>(), c.stream);
        HJ_HIP(hipMemcpyAsync(out, c.hkeys[1].p, n * 8, hipMemcpyDeviceToHost, c.stream));
        HJ_HIP(hipStreamSynchronize(c.stream));
        return HJ_OK;
    });
}

}  // extern "C"
