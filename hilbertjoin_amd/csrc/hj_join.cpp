// hj_join.cpp -- the segmented join pipeline (host orchestration).
//
// Pipeline of one join (reference: EGO_PARALLEL_TRAN egojoin.h:164-222 and
// EGO_PARALLEL_TRAN_NOSELF :395-479 set up, EGO_LOOP_TRAN :311-392 runs):
//   1. grid     per-segment coordinate range -> minimum cell, bits per field,
//               number of key dimensions; 64-bit key per point
//               (segment prefix | cells)
//   2. sort     LSD radix sort of (key, row)
//   3. index    sorted layouts (FP64 rows / transposed tiles, or the filter's
//               operand fragments), per-tile key bounds, per-tile candidate
//               intervals, tile-pair list
//   4. join     one persistent count kernel over the tile-pair list of all
//               segments, counts per segment
#include "hj_join.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <thread>

namespace hj {

thread_local ScratchArena *t_arena = nullptr;

void *ScratchArena::take_dev(size_t need) {
    const size_t at = (dev_used + 255) / 256 * 256;
    if (at > dev_cap || need > dev_cap - at)
        throw CapacityError{"the join needs more than the " + std::to_string(dev_cap >> 20) +
                            " MiB of device scratch memory of the no-allocation entry point"};
    dev_used = at + need;
    return dev + at;
}

void *ScratchArena::take_host(size_t need) {
    const size_t at = (host_used + 63) / 64 * 64;
    if (at > host_cap || need > host_cap - at)
        throw CapacityError{"the join needs more than the " + std::to_string(host_cap >> 10) +
                            " KiB of host staging memory of the no-allocation entry point"};
    host_used = at + need;
    return host + at;
}

namespace {

// Active dimensions of the stripe enumeration when the caller does not fix
// them.  A join samples query tiles and takes the a that minimizes its
// estimated work-list cost (launch_choose_active): one more active dimension
// triples the codes to resolve and cuts the tile pairs only while the tiles
// are narrow in that dimension.  The cost of one code of one query tile
// relative to one tile pair of the count kernel (measured on MI355X: a code
// costs about 0.05 ns, a tile pair about 0.2 ns with the FP16 filter, 0.5 ns
// with BF16 and 3 ns with the FP64 kernel):
float code_cost(int kernel) {
    if (kernel == HJ_KERNEL_MFMA_F16) return 0.25f;
    if (kernel == HJ_KERNEL_MFMA_BF16) return 0.1f;
    return 0.015f;
}

// sampled query tiles per segment: 64 for a single join (the estimate picks
// the same A as 256 on every measured workload at a quarter of the searches),
// fewer per segment in a batch of many
uint32_t active_samples(uint32_t nseg) {
    return std::max<uint32_t>(4, std::min<uint32_t>(64, 1024 / std::max<uint32_t>(nseg, 1)));
}

// The resident index of the serving loop fixes its active dimensions once:
// the largest a <= 6 whose key-prefix classes (points sharing their first a
// cells) still hold AUTO_CLASS_POINTS points on average -- deeper prefixes cut
// no further once a class fits in one or two tiles, while the enumeration
// grows as 3^a.
constexpr int AUTO_CLASS_POINTS = 128;
constexpr int DEFAULT_KBLOCK = 4;
// key fields kept at least (>= the largest active-dimension count + 2)
constexpr int KEY_DIMS_MIN = 8;
// dimensions whose ranges the keys need first: key fields hold at least 3
// bits and a useful key about log_c(n) + 1 fields, so no realistic grid packs
// more than 16 (grid_keys caps it there)
constexpr int KEY_RANGE_DIMS = 16;
// The FP16 filter operands' ranges of the other dimensions come, for a single
// join whose sides have at least SAMPLE_MIN_ROWS rows, from a sample of every
// SAMPLE_STEP-th row, widened by SAMPLE_MARGIN of the sample's width on each
// side; the operand kernels check every point against them, and a join with a
// point outside runs again from exact ranges.
constexpr uint32_t SAMPLE_STEP = 16;
constexpr uint64_t SAMPLE_MIN_ROWS = 8192;
constexpr double SAMPLE_MARGIN = 1.0 / 16;

int ipow3(int a) {
    int r = 1;
    for (int k = 0; k < a; k++) r *= 3;
    return r;
}

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

int bits_for(uint64_t v) {  // bits needed to hold 0..v
    int b = 0;
    while (b < 64 && (v >> b) != 0) b++;
    return b;
}

}  // namespace

Opts read_opts(const hj_opts *o) {
    Opts r;
    if (!o) return r;
    r.kernel = o->kernel;
    r.active = o->active_dims;
    r.order = o->order;
    r.kblock = o->kblock == 0 ? DEFAULT_KBLOCK : o->kblock;
    r.grid = o->grid;
    r.split = o->filter_split;
    r.tile_rows = o->tile_rows == 0 ? 8 : o->tile_rows;
    // measured at n = 1M, d = 64: FP16 0.72 ms at 8 against 0.75 at 4; BF16,
    // whose wider band sends more tile pairs down the slow path, 0.96 at 4
    // against 1.07 at 8
    r.chunk = o->mfma_chunk != 0 ? o->mfma_chunk : 4;
    r.tune = o->tune != 0;
    if (o->filter_frames != 0 && o->filter_frames != -1)
        throw InvalidArg{"filter_frames must be 0 or -1"};
    r.frames = o->filter_frames == 0;
    if (r.kernel < HJ_KERNEL_FP64 || r.kernel > HJ_KERNEL_MFMA_BF16)
        throw InvalidArg{"unknown kernel"};
    if (r.order != HJ_ORDER_HILBERT && r.order != HJ_ORDER_ROW) throw InvalidArg{"unknown order"};
    if (r.kblock != 4 && r.kblock != 8 && r.kblock != 16)
        throw InvalidArg{"kblock must be 4, 8 or 16"};
    if (r.active > 6) throw InvalidArg{"active_dims must be in [-1, 6]"};
    if (r.split < -1) throw InvalidArg{"filter_split must be >= -1"};
    if (r.grid < 0) throw InvalidArg{"grid must be >= 0"};
    if (r.tile_rows != 8 && r.tile_rows != 16) throw InvalidArg{"tile_rows must be 8 or 16"};
    if (r.chunk != 2 && r.chunk != 4 && r.chunk != 8)
        throw InvalidArg{"mfma_chunk must be 2, 4 or 8"};
    return r;
}

void check_sizes(uint64_t n, uint32_t d) {
    if (d == 0) throw InvalidArg{"d must be >= 1"};
    if (d > 65536) throw InvalidArg{"d must be <= 65536"};
    if (n > 0xFFFFFF00ull) throw InvalidArg{"n must be < 2^32 - 256"};
}

int dpad_for(int d, int kblock) {
    const int q = kblock > 8 ? kblock : 8;  // whole early-exit blocks per tile row
    return (d + q - 1) / q * q;
}

// ---- 1. grid -------------------------------------------------------------------
Grid grid_keys(const double *lo_in, const double *hi_in, int dk, int d, double eps,
               int key_budget, int active_req, uint64_t npts) {
    if (!(eps > 0) || !std::isfinite(eps)) throw InvalidArg{"eps must be finite and > 0"};
    Grid g;
    g.eps = eps;
    g.r = 1.0 / eps;
    g.d = d;
    if (!std::isfinite(g.r)) throw InvalidArg{"1/eps is not finite"};
    // a dimension without points (lo > hi) counts as [0, 0]
    auto lo_of = [&](int k) { return lo_in[k] > hi_in[k] ? 0.0 : lo_in[k]; };
    auto hi_of = [&](int k) { return lo_in[k] > hi_in[k] ? 0.0 : hi_in[k]; };
    int64_t range = 0;
    for (int k = 0; k < dk; k++) {
        const double cl = std::floor(lo_of(k) * g.r), ch = std::floor(hi_of(k) * g.r);
        if (!(std::fabs(cl) < 4e18) || !(std::fabs(ch) < 4e18) || ch - cl > 1e18)
            throw InvalidArg{"eps is too small for the coordinate range"};
        range = std::max(range, (int64_t)ch - (int64_t)cl);
    }
    if (range > ((int64_t)1 << 58)) throw InvalidArg{"eps is too small for the coordinate range"};
    // fields hold cell - (min cell - 2) in [2, range + 2]; 2^bits > range + 4
    // leaves room for +-2 without carries
    int bits = 1;
    while (((uint64_t)1 << bits) <= (uint64_t)range + 4) bits++;
    g.bits = bits;
    g.key_dims = std::min(dk, std::max(0, key_budget) / bits);
    if (npts > 1 && range >= 1) {
        // Key fields beyond the first ~log_c(n) + 1 (c cells per dimension)
        // no longer split any cell run of the sorted order, and fields beyond
        // the active ones + 1 hardly narrow a candidate interval; dropping
        // them saves radix passes without changing the work list much (one
        // more field at n = 1M, d = 64, eps 0.2 -- 44 key bits, six passes
        // instead of five -- changes 233 of its 3,986,012 tile pairs).
        const double c = (double)range + 1.0;
        const int useful = (int)std::ceil(std::log((double)npts) / std::log(c)) + 1;
        g.key_dims = std::min(g.key_dims, std::max(KEY_DIMS_MIN, useful));
        // and fewer fields (not below KEY_DIMS_MIN) where that saves a pass
        // of the 8-bit radix sort over the segment prefix and the fields: at
        // n = 1M, d = 64, eps 0.2, 8 fields instead of 10 (32 key bits, four
        // passes) add 10,830 tile pairs (0.27%) and save 0.025 ms of sort
        const int pre = 64 - std::max(0, key_budget);
        auto passes = [&](int k) { return (k * bits + pre + 7) / 8; };
        for (int k = g.key_dims - 1; k >= KEY_DIMS_MIN; k--)
            if (passes(k) < passes(g.key_dims)) {
                g.key_dims = k;
                break;
            }
    }
    g.minc.resize(g.key_dims);
    for (int k = 0; k < g.key_dims; k++) g.minc[k] = (int64_t)std::floor(lo_of(k) * g.r) - 2;
    g.active = active_req < 0 ? -1 : std::max(0, std::min(active_req, std::min(6, g.key_dims)));
    g.centre.assign(d, 0.0);
    return g;
}

void grid_operands(Grid &g, const double *lo_in, const double *hi_in, int headroom) {
    // Filter operands: y = (x - c) * s with c the per-dimension centre and s a
    // power of two such that every |y| and every norm |y| stays <= 2^14 (with
    // 2^headroom to spare).  mbound bounds |x - c| over all points (with a
    // 2^-40 relative margin for the rounding of this host computation).
    g.centre.resize(g.d);
    double m2 = 0;
    for (int k = 0; k < g.d; k++) {
        const bool empty = lo_in[k] > hi_in[k];  // no points: [0, 0]
        const double l = empty ? 0.0 : lo_in[k], h = empty ? 0.0 : hi_in[k];
        g.centre[k] = 0.5 * l + 0.5 * h;
        const double a = std::max(h - g.centre[k], g.centre[k] - l);
        m2 += a * a;
    }
    g.mbound = std::sqrt(m2) * (1.0 + std::ldexp(1.0, -40));
    g.scale = operand_scale(g.mbound, headroom);
}

Grid grid_from_bounds(const double *lo_in, const double *hi_in, int d, double eps,
                      int key_budget, int active_req, int headroom, uint64_t npts) {
    Grid g = grid_keys(lo_in, hi_in, d, d, eps, key_budget, active_req, npts);
    grid_operands(g, lo_in, hi_in, headroom);
    return g;
}

double operand_scale(double mbound, int headroom) {
    if (!(mbound > 0)) return 1.0;
    int e;
    std::frexp(mbound, &e);  // mbound in [2^(e-1), 2^e)
    return std::ldexp(1.0, 14 - headroom - e);
}

double radius_bound(const double *lo, const double *hi, const std::vector<double> &centre) {
    double m2 = 0;
    for (size_t k = 0; k < centre.size(); k++) {
        if (lo[k] > hi[k]) continue;  // no points
        double a = std::max(hi[k] - centre[k], centre[k] - lo[k]);
        m2 += a * a;
    }
    return std::sqrt(m2) * (1.0 + std::ldexp(1.0, -40));
}

namespace {

// Enqueues the per-segment ranges of dimensions [k0, k1) on stream s: mins,
// maxs and per-segment non-finite flags, written straight into the pinned
// read-back rb (or copied there from out where rb is not mapped); with
// key_flags out also holds the keys' widened flag (c.widened) and the operand
// kernels' range flag (c.opflag), cleared by the same launches.  With
// sample_step > 1 a segment of at least SAMPLE_MIN_ROWS rows contributes only
// every sample_step-th row.
void seg_bounds_enqueue(Ctx &c, const double *pts, const std::vector<uint64_t> &off, int d,
                        int k0, int k1, DevBuf &out, DevBuf &tmp, PinnedBuf &stage,
                        PinnedBuf &rb, bool key_flags, hipStream_t s, uint32_t sample_step = 1) {
    const uint32_t nseg = (uint32_t)(off.size() - 1);
    const size_t nv = (size_t)nseg * (k1 - k0);
    const size_t bytes = 2 * nv * sizeof(double) + ((size_t)nseg + 1) / 2 * sizeof(double);
    rb.ensure(bytes);
    char *dst = static_cast<char *>(rb.dev_ptr());
    const bool direct = dst != nullptr;
    const size_t flag_at = direct ? 0 : (bytes + 255) / 256 * 256;
    char *ob = static_cast<char *>(out.ensure(flag_at + 256));
    if (!direct) dst = ob;
    uint32_t *clear = nullptr;
    if (key_flags) {
        clear = reinterpret_cast<uint32_t *>(ob + flag_at);
        c.widened = clear;
        c.opflag = clear + 2;  // 8-byte aligned (clear is 256-byte aligned)
    }
    double *mm = reinterpret_cast<double *>(dst);
    launch_seg_minmax(pts, off.data(), nseg, d, k0, k1, mm, mm + nv,
                      reinterpret_cast<uint32_t *>(mm + 2 * nv), clear, tmp, stage, s,
                      sample_step, SAMPLE_MIN_ROWS);
    if (!direct) HJ_HIP(hipMemcpyAsync(rb.p, dst, bytes, hipMemcpyDeviceToHost, s));
}

// The read-back of seg_bounds_enqueue once its stream has passed it:
// lo/hi[seg * nd + k - k0]; throws InvalidArg on a non-finite coordinate.
void seg_bounds_take(const PinnedBuf &rb, uint32_t nseg, int nd, std::vector<double> &lo,
                     std::vector<double> &hi) {
    const size_t nv = (size_t)nseg * nd;
    const double *h = rb.as<double>();
    const uint32_t *bad = reinterpret_cast<const uint32_t *>(h + 2 * nv);
    for (uint32_t k = 0; k < nseg; k++)
        if (bad[k]) throw InvalidArg{"input contains a non-finite coordinate"};
    lo.assign(h, h + nv);
    hi.assign(h + nv, h + 2 * nv);
}

}  // namespace

void seg_bounds(Ctx &c, const double *pts, const std::vector<uint64_t> &off, int d,
                std::vector<double> &lo, std::vector<double> &hi) {
    seg_bounds_enqueue(c, pts, off, d, 0, d, c.minmax, c.scratch_scan, c.chunk_stage,
                       c.rb_bounds, true, c.stream);
    HJ_HIP(hipStreamSynchronize(c.stream));
    seg_bounds_take(c.rb_bounds, (uint32_t)(off.size() - 1), d, lo, hi);
}

void GridTables::upload(const std::vector<Grid> &gs, hipStream_t s) {
    host.wait();  // the staging buffer may still feed the previous upload
    size_t nminc = 0, ncentre = 0;
    for (const Grid &g : gs) {
        nminc += g.minc.size();
        ncentre += g.centre.size();
    }
    const size_t b0 = std::max<size_t>(gs.size(), 1) * sizeof(GridDev);
    const size_t b1 = std::max<size_t>(nminc, 1) * sizeof(int64_t);
    const size_t b2 = std::max<size_t>(ncentre, 1) * sizeof(double);
    const size_t b3 = b2;  // zmax, laid out as centre
    char *hp = static_cast<char *>(host.ensure(b0 + b1 + b2 + b3));
    std::memset(hp, 0, b0 + b1 + b2 + b3);
    GridDev *gd = reinterpret_cast<GridDev *>(hp);
    int64_t *mc = reinterpret_cast<int64_t *>(hp + b0);
    double *ce = reinterpret_cast<double *>(hp + b0 + b1);
    size_t om = 0, oc = 0;
    for (size_t i = 0; i < gs.size(); i++) {
        const Grid &g = gs[i];
        GridDev v{};
        v.r = g.r;
        v.scale = g.scale;
        v.key_dims = (uint32_t)g.key_dims;
        v.bits = (uint32_t)g.bits;
        v.active = g.active < 0 ? ACTIVE_AUTO : (uint32_t)g.active;
        v.minc_off = (uint32_t)om;
        v.centre_off = (uint32_t)oc;
        v.fr_q = (float)g.fr_q;
        gd[i] = v;
        std::memcpy(mc + om, g.minc.data(), g.minc.size() * sizeof(int64_t));
        std::memcpy(ce + oc, g.centre.data(), g.centre.size() * sizeof(double));
        double *zm = reinterpret_cast<double *>(hp + b0 + b1 + b2) + oc;
        for (size_t k = 0; k < g.centre.size(); k++)
            zm[k] = g.zmax.size() == g.centre.size() ? g.zmax[k] : INFINITY;
        om += g.minc.size();
        oc += g.centre.size();
    }
    char *dp = static_cast<char *>(buf.ensure(b0 + b1 + b2 + b3));
    grids.p = dp;
    minc.p = dp + b0;
    centre.p = dp + b0 + b1;
    zmax.p = dp + b0 + b1 + b2;
    HJ_HIP(hipMemcpyAsync(dp, hp, b0 + b1 + b2 + b3, hipMemcpyHostToDevice, s));
    host.record(s);
}

uint64_t *upload_segs(Ctx &c, const std::vector<SegDev> &segs,
                      const std::vector<uint64_t> &words) {
    hipStream_t s = c.stream;
    c.segs_stage.wait();  // the staging buffer may still feed the previous upload
    const size_t b0 = (segs.size() * sizeof(SegDev) + 255) / 256 * 256;
    const size_t b1 = words.size() * sizeof(uint64_t);
    char *hp = static_cast<char *>(c.segs_stage.ensure(b0 + b1));
    std::memset(hp, 0, b0 + b1);
    if (!segs.empty()) std::memcpy(hp, segs.data(), segs.size() * sizeof(SegDev));
    if (b1) std::memcpy(hp + b0, words.data(), b1);
    c.segs.ensure(std::max<size_t>(b0 + b1, 256));
    HJ_HIP(hipMemcpyAsync(c.segs.p, hp, b0 + b1, hipMemcpyHostToDevice, s));
    c.segs_stage.record(s);
    return reinterpret_cast<uint64_t *>(c.segs.as<char>() + b0);
}

// ---- reduced-precision filter: K layout and thresholds ---------------------------
FilterParams filter_layout(int d, bool bf16, int split_req) {
    FilterParams fp{};
    fp.bf16 = bf16 ? 1 : 0;
    fp.parts = bf16 ? 3 : 2;
    // K chunks: the first k-step holds 16 - 2*parts dimensions and their norm
    // slots; the early exit after it pays when enough dimensions follow.
    const int first = 16 - 2 * fp.parts;
    int split = d;
    // split when enough dimensions follow the first k-step, or when the split
    // costs no k-step at all
    const int steps_whole = (d + 2 * fp.parts + 15) / 16;
    const int steps_split = 1 + (d - first + 2 * fp.parts + 15) / 16;
    if (split_req > 0)
        split = std::min(split_req, d);
    else if (split_req == 0 && d > first && (d >= 2 * first + 4 || steps_split <= steps_whole))
        split = first;
    fp.split = split;
    fp.steps1 = (split + 2 * fp.parts + 15) / 16;
    fp.steps = fp.steps1 + (split < d ? (d - split + 2 * fp.parts + 15) / 16 : 0);
    return fp;
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
void filter_thresholds(const FilterParams &fp, int d, double scale, double M, double eps,
                       float *t_in_out, float *t_out_out) {
    const bool bf16 = fp.bf16 != 0;
    const double u = bf16 ? std::ldexp(1.0, -8) * (1 + std::ldexp(1.0, -15))
                          : std::ldexp(1.0, -11) * (1 + std::ldexp(1.0, -12));
    const double flush = bf16 ? std::ldexp(1.0, -100) : std::ldexp(1.0, -14);
    const double e = 2 * u * M + 2 * std::sqrt((double)d) * flush + std::ldexp(1.0, -49) * M;
    const double mu = (d + 4) * std::ldexp(1.0, -52) + std::ldexp(1.0, -50);
    const double K = fp.steps * 16.0;
    const double E2 = 4 * (K + 1) * std::ldexp(1.0, -24) * M * M;
    const double R = (bf16 ? std::ldexp(1.0, -20) : std::ldexp(1.0, -18)) * M * M +
                     d * std::ldexp(1.0, -50) * M * M + 128.0;
    const double Er = E2 + R;
    const double se = scale * eps;
    const double th_in = se * std::sqrt(1 - mu) * (1 - std::ldexp(1.0, -30)) - e;
    const double th_out = se * std::sqrt(1 + mu) * (1 + std::ldexp(1.0, -30)) + e;
    const double t_in = th_in > 0 ? -0.5 * th_in * th_in + Er : INFINITY;
    const double t_out = -0.5 * th_out * th_out - Er;
    *t_in_out =
        round_up_f(t_in * (t_in < 0 ? 1 - std::ldexp(1.0, -40) : 1 + std::ldexp(1.0, -40)));
    *t_out_out = round_down_f(t_out * (1 + std::ldexp(1.0, -40)));
}

// Band constants of the filter with tile frames (frame_thresholds in
// hj_frames.h): the terms of filter_thresholds per unit of the tile pair's
// radii.  The accumulated error covers the FP32 sum of the K products
// ((K + 1) 2^-24 of P, doubled), the corrections r and c (FP32 dot products of
// at most K exact products: (K + 1) 2^-24 of P), the norm split
// (filter_thresholds' R per unit of P) and the corrected norms' FP32 re-split
// (2^-21 of P); the operand error adds the FP64 rounding of z - anchor
// (2^-46 of the grid's radius bound M).
void frame_consts(const FilterParams &fp, int d, double scale, double M, double eps,
                  FrameConst *fc) {
    const bool bf16 = fp.bf16 != 0;
    const double u = bf16 ? std::ldexp(1.0, -8) * (1 + std::ldexp(1.0, -15))
                          : std::ldexp(1.0, -11) * (1 + std::ldexp(1.0, -12));
    const double flush = bf16 ? std::ldexp(1.0, -100) : std::ldexp(1.0, -14);
    const double mu = (d + 4) * std::ldexp(1.0, -52) + std::ldexp(1.0, -50);
    const double K = fp.steps * 16.0;
    const double se = scale * eps;
    std::memset(fc, 0, sizeof(*fc));
    fc->sek_in = round_down_f(se * std::sqrt(1 - mu) * (1 - std::ldexp(1.0, -30)));
    fc->sek_out = round_up_f(se * std::sqrt(1 + mu) * (1 + std::ldexp(1.0, -30)));
    fc->u = round_up_f(u);
    fc->e0 = round_up_f(2 * std::sqrt((double)d) * flush + std::ldexp(1.0, -46) * M);
    const double cR = (bf16 ? std::ldexp(1.0, -20) : std::ldexp(1.0, -18)) / 2 +
                      d * std::ldexp(1.0, -51);
    fc->cP = round_up_f(3 * (K + 1) * std::ldexp(1.0, -24) + cR + std::ldexp(1.0, -21));
    fc->c0 = 128.0f;
    fc->cd = round_up_f((d + 1) * std::ldexp(1.0, -25));
}

bool frames_possible(const FilterParams &fp, int d) {
    // the anchors come from a tile staged in FP64 in LDS, and the fast path
    // corrects one k-step (the first chunk is one k-step)
    return d <= FRAME_MAX_D && filter_prep_fused(d) && fp.steps1 == 1;
}

// Lattice step of the tile anchors: the smallest power of two q with every
// scaled coordinate |z_k| <= H = scale * max_k max(hi_k - c_k, c_k - lo_k)
// within FRAME_ANCHOR_MAX q.
double frame_lattice(const double *lo, const double *hi, const Grid &g, bool bf16) {
    double H = 0;
    for (int k = 0; k < g.d; k++) {
        if (lo[k] > hi[k]) continue;
        H = std::max(H, std::max(hi[k] - g.centre[k], g.centre[k] - lo[k]));
    }
    H *= g.scale * (1 + std::ldexp(1.0, -40));
    const double amax = bf16 ? FRAME_ANCHOR_MAX_BF16 : FRAME_ANCHOR_MAX_F16;
    if (!(H > 0)) return 1.0;
    int e;
    std::frexp(H / amax, &e);  // H / amax in [2^(e-1), 2^e)
    return std::ldexp(1.0, e);
}

// Expanded-form FP64 thresholds.
//
// The kernel computes s = Pa + Pb - 2 a.b from prefix sums of squares Pa, Pb
// and the dot product a.b, each an FMA chain over c <= d dimensions, and
// compares fma(2, a.b, -Pa) with Pb - x.  With exact norms Pa*, Pb* and
// |a_k b_k| <= (a_k^2 + b_k^2) / 2, the computed s differs from the exact
// squared distance D^2 (over those c dimensions) by at most
//   E = (2 gamma_d + 3 u) (Pa* + Pb*) + u x,   gamma_d = d u / (1 - d u),
// u = 2^-53 (the dot product's and both norms' chains, the final fma and
// subtraction).  The fixed-order predicate's own sum lies within a relative
// (d + 2) u of D^2, and eps*eps is rounded once, so with mu as the filter's,
//   s > x_out = eps2 (1 + mu) + E   =>  the predicate says outside,
//   s <= x_in = eps2 (1 - mu) - E   =>  the predicate says within,
// and a partial sum over the first c dimensions above x_out already rules the
// pair out (D^2 never decreases with c).  Pairs in between are rechecked with
// the fixed-order predicate inside the kernel.
void fp64_thresholds(int d, double eps, double pmax_sum, double *x_in, double *x_out) {
    const double u = std::ldexp(1.0, -53);
    const double eps2 = eps * eps;
    *x_in = *x_out = -1.0;
    if (!std::isfinite(pmax_sum) || pmax_sum > 1e300) return;
    const double gamma = d * u / (1.0 - d * u);
    const double mu = (d + 4) * std::ldexp(1.0, -52) + std::ldexp(1.0, -50);
    const double E = ((2 * gamma + 3 * u) * pmax_sum + 2 * u * eps2) * (1 + std::ldexp(1.0, -20));
    // a band that is not thin against eps^2 (coordinates far from the origin
    // relative to eps) would recheck most close pairs: the difference form
    // is the better kernel then
    if (!(E <= std::ldexp(eps2, -12))) return;
    *x_out = eps2 * (1 + mu) + E;
    *x_in = eps2 * (1 - mu) - E;
}

double norm2_bound(const double *lo, const double *hi, int d) {
    double s = 0;
    for (int k = 0; k < d; k++) {
        if (lo[k] > hi[k]) continue;  // no points
        const double a = std::max(std::fabs(lo[k]), std::fabs(hi[k]));
        s += a * a;
    }
    return s * (1 + std::ldexp(1.0, -40));
}

// ---- host <-> device copies ---------------------------------------------------------
namespace {

constexpr size_t STAGE_CHUNK = (size_t)64 << 20;
constexpr size_t STAGE_MIN = (size_t)32 << 20;  // smaller copies go straight through

// memcpy split over up to 16 threads (at least 4 MiB each): first-touch page
// faults of a fresh destination and the copy itself scale with threads.
void par_memcpy(void *dst, const void *src, size_t bytes) {
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t T = std::min<size_t>({16, hw, std::max<size_t>(1, bytes >> 22)});
    const size_t per = (bytes + T - 1) / T;
    std::vector<std::thread> th;
    th.reserve(T);
    for (size_t t = 1; t < T && t * per < bytes; t++) {
        const size_t o = t * per, len = std::min(per, bytes - o);
        th.emplace_back([=] {
            std::memcpy(static_cast<char *>(dst) + o, static_cast<const char *>(src) + o, len);
        });
    }
    std::memcpy(dst, src, std::min(per, bytes));
    for (auto &x : th) x.join();
}

}  // namespace

void copy_d2h(Ctx &c, void *dst, const void *src, size_t bytes) {
    hipStream_t s = c.stream;
    if (bytes < STAGE_MIN) {
        if (bytes) HJ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        HJ_HIP(hipStreamSynchronize(s));
        return;
    }
    for (PinnedBuf &b : c.stage) b.ensure(STAGE_CHUNK);
    const size_t nch = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
    auto issue = [&](size_t k) {
        const size_t o = k * STAGE_CHUNK, len = std::min(STAGE_CHUNK, bytes - o);
        PinnedBuf &b = c.stage[k & 1];
        HJ_HIP(hipMemcpyAsync(b.p, static_cast<const char *>(src) + o, len,
                              hipMemcpyDeviceToHost, s));
        b.record(s);
    };
    issue(0);
    for (size_t k = 0; k < nch; k++) {
        // the other buffer was drained by the host copy of chunk k - 1
        if (k + 1 < nch) issue(k + 1);
        PinnedBuf &b = c.stage[k & 1];
        b.wait();
        const size_t o = k * STAGE_CHUNK, len = std::min(STAGE_CHUNK, bytes - o);
        par_memcpy(static_cast<char *>(dst) + o, b.p, len);
    }
}

void copy_h2d(Ctx &c, void *dst, const void *src, size_t bytes) {
    hipStream_t s = c.stream;
    if (bytes < STAGE_MIN) {
        if (bytes) HJ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        return;
    }
    for (PinnedBuf &b : c.stage) b.ensure(STAGE_CHUNK);
    const size_t nch = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
    for (size_t k = 0; k < nch; k++) {
        PinnedBuf &b = c.stage[k & 1];
        b.wait();  // its previous DMA (chunk k - 2) is done
        const size_t o = k * STAGE_CHUNK, len = std::min(STAGE_CHUNK, bytes - o);
        par_memcpy(b.p, static_cast<const char *>(src) + o, len);
        HJ_HIP(hipMemcpyAsync(static_cast<char *>(dst) + o, b.p, len, hipMemcpyHostToDevice, s));
        b.record(s);
    }
}

// ---- 2. + 3a. keys, sort, layouts -------------------------------------------------
void side_layout(Side &S, const std::vector<uint64_t> &off, const double *pts, int d,
                 int kblock) {
    S.pts = pts;
    S.off = off;
    S.n = off.back();
    S.d = d;
    S.dpad = dpad_for(d, kblock);
    S.tile0.resize(off.size() - 1);
    size_t t = 0;
    for (size_t k = 0; k + 1 < off.size(); k++) {
        S.tile0[k] = (uint32_t)t;
        t += ntiles_of(off[k + 1] - off[k]);
    }
    S.ntiles = t;
    if (t > 0xFFFFFFF0ull / 64) throw InvalidArg{"too many points"};
}

std::vector<uint32_t> tile_owners(const Side &S) {
    std::vector<uint32_t> own(S.ntiles);
    for (size_t k = 0; k < S.tile0.size(); k++) {
        const size_t e = k + 1 < S.tile0.size() ? S.tile0[k + 1] : S.ntiles;
        for (size_t t = S.tile0[k]; t < e; t++) own[t] = (uint32_t)k;
    }
    return own;
}

void side_keys(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table, uint32_t nseg,
               const uint64_t *seg_off_dev, bool target, const GridTables &gt, int kd_max) {
    const size_t nk = std::max<size_t>(S.n, 1);
    for (int b = 0; b < 2; b++) {
        B.keys[b].ensure(nk * sizeof(uint64_t));
        B.vals[b].ensure(nk * sizeof(uint32_t));
    }
    // a query side also gets every point's widened key bounds, from which the
    // tile bounds of the work list follow after the sort
    uint64_t *plo = nullptr, *phi = nullptr;
    if (!target) {
        plo = static_cast<uint64_t *>(B.pbounds.ensure(2 * nk * sizeof(uint64_t)));
        phi = plo + nk;
    }
    if (!target && !c.widened) throw InvalidArg{"internal: keys before the coordinate ranges"};
    launch_keys(S.pts, S.n, S.d, seg_table, nseg, seg_off_dev, gt.grids.as<GridDev>(),
                gt.minc.as<int64_t>(), kd_max, !target, B.keys[0].as<uint64_t>(),
                B.vals[0].as<uint32_t>(), plo, phi, target ? nullptr : c.widened, c.stream);
    S.widened = target ? nullptr : c.widened;
    S.plo = plo;
    S.phi = phi;
    B.cur = 0;
    S.keys = B.keys[0].as<uint64_t>();
    S.perm = B.vals[0].as<uint32_t>();
}

void side_sort(Ctx &c, SetBufs &B, Side &S, int sort_bits) {
    uint64_t *kk[2] = {B.keys[0].as<uint64_t>(), B.keys[1].as<uint64_t>()};
    uint32_t *vv[2] = {B.vals[0].as<uint32_t>(), B.vals[1].as<uint32_t>()};
    int which = 0;
    if (S.n > 1 && sort_bits > 0)
        radix_sort_u64_u32(kk, vv, S.n, sort_bits, &which, c.scratch_hist, c.scratch_scan,
                           c.stream);
    B.cur = which;
    S.keys = kk[which];
    S.perm = vv[which];
}

void side_index(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table,
                const std::vector<uint32_t> &tile_owner, bool target, const SideNeeds &need) {
    side_tile_segments(c, B, S, tile_owner, c.stream);
    side_search_index(c, B, S, seg_table, target, need, c.stream);
}

void side_tile_segments(Ctx &c, SetBufs &B, Side &S, const std::vector<uint32_t> &tile_owner,
                        hipStream_t s) {
    (void)c;
    const size_t nt = std::max<size_t>(S.ntiles, 1);
    // tile -> segment-table entry
    B.tile_seg.ensure(nt * sizeof(uint32_t));
    bool all_zero = true;
    for (uint32_t v : tile_owner) all_zero &= v == 0;
    if (all_zero) {
        launch_zero_u64(B.tile_seg.as<uint64_t>(), (nt + 1) / 2, s);
    } else {
        B.tile_seg_stage.wait();  // the staging buffer may still feed the previous upload
        const size_t bytes = tile_owner.size() * sizeof(uint32_t);
        void *hp = B.tile_seg_stage.ensure(bytes);
        std::memcpy(hp, tile_owner.data(), bytes);
        HJ_HIP(hipMemcpyAsync(B.tile_seg.p, hp, bytes, hipMemcpyHostToDevice, s));
        B.tile_seg_stage.record(s);
    }
    S.tile_seg = B.tile_seg.as<uint32_t>();
}

void side_search_index(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table, bool target,
                       const SideNeeds &need, hipStream_t s) {
    (void)c;
    const size_t nt = std::max<size_t>(S.ntiles, 1);
    // first key of every tile: the level-1 table of the stripe searches (any
    // side may be searched: the target side, or a self-join's one side)
    B.tfirst.ensure(nt * sizeof(uint64_t));
    S.tfirst = B.tfirst.as<uint64_t>();
    uint64_t *tlo = nullptr, *thi = nullptr;
    if (need.bounds) {
        if (!S.plo || !S.widened) throw InvalidArg{"internal: point key bounds missing"};
        tlo = static_cast<uint64_t *>(B.tlo.ensure(nt * sizeof(uint64_t)));
        thi = static_cast<uint64_t *>(B.thi.ensure(nt * sizeof(uint64_t)));
        S.tlo = tlo;
        S.thi = thi;
    }
    launch_tile_index(S.keys, S.perm, S.plo, S.phi, seg_table, S.tile_seg, S.ntiles, target,
                      S.widened, B.tfirst.as<uint64_t>(), tlo, thi, s);
}

void side_operands(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table, bool target,
                   const GridTables &gt, const FilterParams &fp, const SideNeeds &need,
                   uint32_t *opflag, hipStream_t st) {
    hipStream_t s = st ? st : c.stream;
    const size_t nt = std::max<size_t>(S.ntiles, 1);
    const bool frag = need.frag_row || need.frag_col;
    if (need.fp64_rows || need.fp64_cols) {
        // one FP64 tile layout serves both roles: the FP64 kernels read a row
        // tile's R rows of one dimension as one contiguous scalar load; the
        // same pass writes the prefix norms
        const size_t bytes = nt * 64 * (size_t)S.dpad * sizeof(double);
        B.cols.ensure(bytes);
        B.pnorm.ensure(bytes / 4);
        launch_gather(S.pts, S.perm, seg_table, S.tile_seg, S.ntiles, target, S.d, S.dpad,
                      B.cols.as<double>(), B.pnorm.as<double>(), s);
        S.rows = B.cols.as<double>();
        S.cols = B.cols.as<double>();
        S.pn = B.pnorm.as<double>();
    }
    if (frag) {
        const size_t bytes = nt * 2 * (size_t)fp.steps * 64 * 16;
        if (need.frag_row) B.frow.ensure(bytes);
        if (need.frag_col) B.fcol.ensure(bytes);
        void *fr = need.frag_row ? B.frow.p : nullptr;
        void *fc = need.frag_col ? B.fcol.p : nullptr;
        FrameOut fo;
        S.frow2 = S.fcol2 = nullptr;
        S.fanch = nullptr;
        S.finfo = nullptr;
        S.fflag = nullptr;
        if (fp.frames) {
            // own-frame fragments, anchors per K slot, {rho, framed} per tile,
            // then the side's flag
            if (need.frag_row) fo.frow2 = B.frow2.ensure(bytes);
            if (need.frag_col) fo.fcol2 = B.fcol2.ensure(bytes);
            S.frow2 = fo.frow2;
            S.fcol2 = fo.fcol2;
            B.fanch.ensure(nt * (size_t)fp.steps * 16 * sizeof(float));
            B.finfo.ensure(nt * sizeof(float2) + sizeof(uint64_t));
            fo.anch = B.fanch.as<float>();
            fo.info = B.finfo.as<float2>();
            fo.flag = reinterpret_cast<uint32_t *>(B.finfo.as<char>() + nt * sizeof(float2));
            launch_zero_u64(reinterpret_cast<uint64_t *>(fo.flag), 1, s);
            S.fanch = fo.anch;
            S.finfo = fo.info;
            S.fflag = fo.flag;
        }
        launch_filter_prep(S.pts, S.perm, seg_table, S.tile_seg, S.ntiles, target, S.d,
                           gt.grids.as<GridDev>(), gt.centre.as<double>(), fp, fr, fc, fo,
                           gt.zmax.as<double>(), opflag, s);
        S.frow = fr;
        S.fcol = fc;
    }
}

void side_layouts(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table,
                  const std::vector<uint32_t> &tile_owner, bool target, const GridTables &gt,
                  const FilterParams &fp, const SideNeeds &need) {
    side_index(c, B, S, seg_table, tile_owner, target, need);
    side_operands(c, B, S, seg_table, target, gt, fp, need);
}

// the work-list size is read back at offset 0 of the pinned rb_list buffer, the
// grid table with the chosen active dimensions at this offset
constexpr size_t WL_READBACK_GRIDS = 64;

// ---- 3b. work list ------------------------------------------------------------------
// Tile pairs of all of a side pair's segments: no work list is longer.
uint64_t all_tile_pairs(const Side &Q, const Side &T) {
    uint64_t total = 0;
    const size_t nseg = Q.tile0.size();
    for (size_t k = 0; k < nseg; k++) {
        const uint64_t nq = (k + 1 < nseg ? Q.tile0[k + 1] : Q.ntiles) - Q.tile0[k];
        const uint64_t nt =
            k < T.tile0.size() ? (k + 1 < T.tile0.size() ? T.tile0[k + 1] : T.ntiles) - T.tile0[k]
                               : T.ntiles;
        total += nq * nt;
    }
    return total;
}

WorkListPending worklist_start(Ctx &c, const Side &Q, const Side &T, const SegDev *segs,
                               const GridDev *grids, int max_iv, int order, hipStream_t s,
                               bool defer_length) {
    WorkListPending p;
    p.s = s;
    p.order = order;
    p.max_iv = max_iv;
    p.ntq = Q.ntiles;
    p.ntt = T.ntiles;
    if (p.ntq == 0 || T.n == 0) return p;
    // a deferred length needs a capacity fixed now: the all-pairs bound, at
    // most two target tiles per stripe code of every query tile (a list
    // longer than that is emitted again at its length)
    p.defer = defer_length && order == HJ_ORDER_ROW && !t_arena;
    if (p.defer)
        p.cap_hint = std::min<uint64_t>(all_tile_pairs(Q, T), 2ull * p.ntq * (uint64_t)max_iv + 4096);
    const size_t ntq = p.ntq;
    c.ivals.ensure(ntq * (size_t)max_iv * 2 * sizeof(uint32_t));
    c.icount.ensure(ntq * sizeof(uint32_t));
    // [jcount: ntq + 1][cand: CAND_SLOTS][off: ntq + 1]; the tile-range
    // kernel writes jcount[0, ntq), so one launch clears the last length and
    // the candidate sums
    c.ioff.ensure((2 * ntq + 2 + CAND_SLOTS) * sizeof(uint64_t));
    uint64_t *jcount = c.ioff.as<uint64_t>();
    uint64_t *cand = jcount + ntq + 1;
    p.off = cand + CAND_SLOTS;
    launch_zero_u64(jcount + ntq, 1 + CAND_SLOTS, s);
    launch_tile_ranges(Q.tlo, Q.thi, Q.tile_seg, ntq, T.keys, T.tfirst, segs, grids,
                       c.ivals.as<uint32_t>(), c.icount.as<uint32_t>(), jcount, max_iv, cand, s);
    p.wl.cand = cand;
    exclusive_scan_u64(jcount, p.off, ntq + 1, c.scratch_scan, s);
    if (!p.defer) {  // a deferred list's length comes back with the join's counts
        p.hnp = static_cast<uint64_t *>(c.rb_list.ensure(WL_READBACK_GRIDS));
        readback_u64(c.rb_list, 0, p.off + ntq, 1, s);
    }
    p.pending = true;
    return p;
}

WorkList worklist_finish(Ctx &c, WorkListPending &p) {
    WorkList &wl = p.wl;
    if (!p.pending) return wl;
    p.pending = false;
    hipStream_t s = p.s;
    const size_t ntq = p.ntq;
    wl.ntq = ntq;
    wl.max_iv = p.max_iv;
    wl.off = p.off;
    wl.rows_grouped = p.order != HJ_ORDER_HILBERT;
    if (p.defer) {
        // no wait: the capacity is the larger of the buffer's and the hint
        const uint64_t have = c.pairs[0].bytes / sizeof(uint64_t);
        const uint64_t cap =
            std::max<uint64_t>(have > WORK_LIST_PAD ? have - WORK_LIST_PAD : 0, p.cap_hint);
        c.pairs[0].ensure((cap + WORK_LIST_PAD) * sizeof(uint64_t));
        launch_emit_pairs(c.ivals.as<uint32_t>(), c.icount.as<uint32_t>(), p.off, ntq, p.max_iv,
                          c.pairs[0].as<uint64_t>(), cap, s);
        wl.pairs = c.pairs[0].as<uint64_t>();
        wl.npairs = cap;
        wl.npairs_dev = p.off + ntq;
        return wl;
    }
    HJ_HIP(hipStreamSynchronize(s));
    const uint64_t np = *p.hnp;
    wl.npairs = np;
    if (np == 0) return wl;
    for (int b = 0; b < 2; b++) c.pairs[b].ensure((np + WORK_LIST_PAD) * sizeof(uint64_t));
    launch_emit_pairs(c.ivals.as<uint32_t>(), c.icount.as<uint32_t>(), p.off, ntq, p.max_iv,
                      c.pairs[0].as<uint64_t>(), np, s);
    wl.pairs = c.pairs[0].as<uint64_t>();
    if (p.order == HJ_ORDER_HILBERT && np > 1) {
        wl.rows_grouped = false;
        const size_t side = std::max(ntq, p.ntt);
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

void worklist_regrow(Ctx &c, WorkList &wl, uint64_t np, hipStream_t s) {
    c.pairs[0].ensure((np + WORK_LIST_PAD) * sizeof(uint64_t));
    launch_emit_pairs(c.ivals.as<uint32_t>(), c.icount.as<uint32_t>(), wl.off, wl.ntq, wl.max_iv,
                      c.pairs[0].as<uint64_t>(), np, s);
    wl.pairs = c.pairs[0].as<uint64_t>();
    wl.npairs = np;
}

WorkList build_worklist(Ctx &c, const Side &Q, const Side &T, const SegDev *segs,
                        const GridDev *grids, int max_iv, int order, hipStream_t s) {
    WorkListPending p = worklist_start(c, Q, T, segs, grids, max_iv, order, s);
    return worklist_finish(c, p);
}

// ---- 4. count --------------------------------------------------------------------------
// The filter's frame tables when the operands were built with tile frames
// (whether some tile took one the kernel reads from the sides' flags).
void set_frames(JoinArgs &ja, const Side &Q, const Side &T, const FilterParams &fp) {
    ja.frames = fp.frames != 0 && Q.fflag && T.fflag;
    if (fp.frames && !ja.frames && (Q.fflag || T.fflag))
        throw InvalidArg{"internal: one side of the join without its frame tables"};
    ja.fr.fa2 = Q.frow2;
    ja.fr.fb2 = T.fcol2;
    ja.fr.qinfo = Q.finfo;
    ja.fr.tinfo = T.finfo;
    ja.fr.qanch = Q.fanch;
    ja.fr.tanch = T.fanch;
    ja.fr.qflag = Q.fflag;
    ja.fr.tflag = T.fflag;
}

// The join counters followed by the filter's scheduling counters, zeroed on
// stream s by one launch.
uint64_t *zeroed_counters(Ctx &c, hipStream_t s) {
    c.counters.ensure((8 + SCHED_WORDS) * sizeof(uint64_t));
    uint64_t *ctr = c.counters.as<uint64_t>();
    launch_zero_u64(ctr, 8 + SCHED_WORDS, s);
    return ctr;
}

// The join counters and the per-segment counts, zeroed on stream s.
uint64_t *zero_count_state(Ctx &c, uint32_t nseg, hipStream_t s) {
    c.seg_counts.ensure(std::max<size_t>(nseg, 1) * sizeof(uint64_t));
    launch_zero_u64(c.seg_counts.as<uint64_t>(), nseg, s);
    return zeroed_counters(c, s);
}

void run_count(Ctx &c, const Side &Q, const Side &T, const SegDev *segs,
               const std::vector<SegDev> &segs_host, const WorkList &wl, const Opts &o,
               const FilterParams &fp, bool self, JoinOut &R, const uint32_t *opflag,
               bool zeroed) {
    const uint32_t nseg = (uint32_t)segs_host.size();
    hipStream_t s = c.stream;
    uint64_t *ctr = zeroed ? c.counters.as<uint64_t>() : zero_count_state(c, nseg, s);
    uint64_t *sc = c.seg_counts.as<uint64_t>();
    WorkList w = wl;
    JoinArgs ja{w.pairs, w.npairs, segs, Q.tile_seg, ctr, sc, o.grid, c.num_cus, o.tile_rows,
                o.chunk, reinterpret_cast<uint32_t *>(ctr + 8)};
    ja.npairs_dev = w.npairs_dev;
    ja.rows_grouped = w.rows_grouped;
    ja.opflag = opflag;
    R.ranges_missed = false;
    set_frames(ja, Q, T, fp);
    const bool filter = o.kernel != HJ_KERNEL_FP64;
    if (!filter && w.npairs_dev) throw InvalidArg{"internal: FP64 kernel without the list's length"};
    HJ_HIP(hipEventRecord(c.ev[5], s));
    if (!filter) {
        // an early-exit block never reaches past the tile layout's padding
        // (an index built for a smaller block keeps its layout)
        const int kb = Q.dpad % o.kblock == 0 && T.dpad == Q.dpad ? o.kblock : 8;
        bool expanded = Q.pn && T.pn;
        for (const SegDev &v : segs_host) expanded &= v.x_out >= 0;
        launch_fp64_count(ja, Q.rows, T.cols, Q.pn, T.pn, Q.dpad, self, kb, expanded, s);
        HJ_HIP(hipEventRecord(c.ev[6], s));
    }
    uint64_t wl_cand[CAND_SLOTS] = {};
    R.seg_counts.assign(nseg, 0);
    // pinned read-back: counters [8], candidate sums [CAND_SLOTS], segment
    // counts [nseg], the list's length [1], the operand range flag [1]
    uint64_t *hr = static_cast<uint64_t *>(
        c.rb_counts.ensure((8 + CAND_SLOTS + (size_t)nseg + 2) * sizeof(uint64_t)));
    uint64_t np_dev = 0;
    auto readback = [&] {
        // one launch for all ranges (dst: byte offsets into rb_counts)
        U64Ranges rr{};
        auto add = [&](const uint64_t *src, size_t word, size_t n) {
            rr.r[rr.nr++] = U64Range{src, reinterpret_cast<uint64_t *>(word * sizeof(uint64_t)), n};
        };
        const size_t at_np = 8 + CAND_SLOTS + nseg, at_flag = at_np + 1;
        add(ctr, 0, 8);
        if (filter && w.cand) add(w.cand, 8, CAND_SLOTS);
        if (nseg) add(sc, 8 + CAND_SLOTS, nseg);
        if (w.npairs_dev) add(w.npairs_dev, at_np, 1);
        if (opflag) add(reinterpret_cast<const uint64_t *>(opflag), at_flag, 1);
        readback_ranges(c.rb_counts, rr, s);
        HJ_HIP(hipStreamSynchronize(s));
        std::memcpy(R.counters, hr, 8 * sizeof(uint64_t));
        if (filter && w.cand) std::memcpy(wl_cand, hr + 8, sizeof(wl_cand));
        if (nseg) std::memcpy(R.seg_counts.data(), hr + 8 + CAND_SLOTS, nseg * sizeof(uint64_t));
        if (w.npairs_dev) np_dev = hr[at_np];
        if (opflag) R.ranges_missed = (uint32_t)hr[at_flag] != 0;
    };
    if (filter) {
        // filter, then the FP64 recheck of its band queue; the queue length
        // stays on the device, an overflow shows in the counters at the end
        FilterOut out;
        out.band_cap = std::max<uint64_t>((uint64_t)1 << 20, c.band.bytes / sizeof(uint64_t));
        out.band = static_cast<uint64_t *>(c.band.ensure(out.band_cap * sizeof(uint64_t)));
        launch_mfma_count(ja, Q.frow, T.fcol, fp, self, out, s);
        HJ_HIP(hipEventRecord(c.ev[6], s));
        launch_band_recheck(Q.pts, Q.perm, T.pts, T.perm, segs, Q.tile_seg, Q.d, ctr + 4, out,
                            ctr, sc, c.num_cus, s);
        readback();
        if (R.ranges_missed) return;  // nothing was counted: the caller starts again
        const bool list_short = w.npairs_dev && np_dev > w.npairs;
        if (list_short) {
            // the list was longer than its capacity: emit it whole, then the
            // length is known
            worklist_regrow(c, w, np_dev, s);
            ja.pairs = w.pairs;
            ja.npairs = w.npairs;
        }
        if (w.npairs_dev) {
            w.npairs = np_dev;
            w.npairs_dev = nullptr;
            ja.npairs = np_dev;
            ja.npairs_dev = nullptr;
        }
        if (list_short || R.counters[4] > out.band_cap) {
            // the band queue overflowed or the list was cut: size them and
            // run both again
            out.band_cap = std::max<uint64_t>(out.band_cap, R.counters[4]);
            out.band = static_cast<uint64_t *>(c.band.ensure(out.band_cap * sizeof(uint64_t)));
            zeroed_counters(c, s);
            launch_zero_u64(sc, nseg, s);
            HJ_HIP(hipEventRecord(c.ev[5], s));
            launch_mfma_count(ja, Q.frow, T.fcol, fp, self, out, s);
            HJ_HIP(hipEventRecord(c.ev[6], s));
            launch_band_recheck(Q.pts, Q.perm, T.pts, T.perm, segs, Q.tile_seg, Q.d, ctr + 4,
                                out, ctr, sc, c.num_cus, s);
            readback();
        }
    } else {
        readback();
    }
    HJ_HIP(hipEventElapsedTime(&R.kernel_ms, c.ev[5], c.ev[6]));
    if (filter) {
        // the filter takes its candidate count from the work list
        R.counters[3] = 0;
        for (uint64_t v : wl_cand) R.counters[3] += v;
        // K slots multiplied, per pair of points: each MFMA of the filter
        // multiplies 16 K slots for 32 x 32 pairs (device count of MFMAs)
        R.counters[2] = R.counters[0] * 1024ull * 16ull;
    }
    R.count = R.counters[1];
    R.wl = w;
}

void fill_stats(hj_stats *st, const JoinOut &R, const JoinPrep &prep, const Opts &o,
                double denom) {
    if (!st) return;
    st->kernel_ms = R.kernel_ms;
    st->tile_pairs = R.wl.npairs;
    st->candidate_pairs = R.counters[3];
    st->pair_dims = R.counters[2];
    st->band_pairs = R.counters[4];
    st->early_exits = R.counters[7];
    st->frames = (int32_t)R.counters[5];
    st->sampled_ranges = R.sampled_ranges;
    st->load_fraction = denom > 0 ? (double)R.counters[3] / denom : 0.0;
    if (!prep.grids.empty()) {
        const Grid &g = prep.grids[0];
        st->key_dims = g.key_dims;
        st->key_bits = g.bits;
        st->active_dims = g.active;
    }
    st->kernel = o.kernel;
}

// ---- the whole join of a problem ----------------------------------------------------
void run_problem(Ctx &c, const Problem &P, const Opts &o, JoinOut &R, JoinPrep &prep,
                 hj_stats *st, bool need_layout, bool worklist_only) {
    hipStream_t s = c.stream;
    Phase ph{c.ev, s};
    const int d = P.d;
    const bool self = P.self;
    const uint32_t nseg = (uint32_t)(P.Q.off.size() - 1);
    if (P.eps.size() != nseg || (!self && P.T.off.size() != P.Q.off.size()))
        throw InvalidArg{"internal: segment tables disagree"};
    ph.mark(0);
    // 1. grids: per join segment from the bounds of both of its sides.  The
    // key fields use only the leading KEY_RANGE_DIMS dimensions: their ranges
    // come first (the host waits for them).  For the FP16 filter on one large
    // join, the other dimensions' ranges come from a sample of the rows,
    // taken on the second stream right after them, and every point is checked
    // against them where the operands are made; a join with a point outside
    // runs again from exact ranges.  Otherwise those ranges are exact, taken
    // on the second stream beside the keys and the sort, and complete the
    // grids' operand part and the thresholds after the sort is enqueued.
    const int dk = std::min(d, KEY_RANGE_DIMS);
    const bool rest = dk < d;
    const bool filter = o.kernel != HJ_KERNEL_FP64;
    // a sample for the FP16 filter: one join of large sides (in a batch one
    // job's outlier would run every job again)
    const bool sample_try = o.kernel == HJ_KERNEL_MFMA_F16 && rest && filter_prep_fused(d) &&
                            nseg == 1 && P.Q.off[1] - P.Q.off[0] >= SAMPLE_MIN_ROWS &&
                            (self || P.T.off[1] - P.T.off[0] >= SAMPLE_MIN_ROWS);
    // nothing enqueued here may outlive a failed call (the next one reuses the
    // buffers): on an exception both streams drain first
    struct DrainOnThrow {
        Ctx &c;
        ~DrainOnThrow() {
            if (std::uncaught_exceptions() == 0) return;
            (void)hipStreamSynchronize(c.stream2);
            (void)hipStreamSynchronize(c.stream);
        }
    } drain{c};
    Ctx::RangePass *rp = c.ranges;
    auto range_pass = [&](Ctx::RangePass &r, const double *pts, const std::vector<uint64_t> &off,
                          int k0, int k1, bool key_flags, hipStream_t st, uint32_t step) {
        seg_bounds_enqueue(c, pts, off, d, k0, k1, r.out, r.tmp, r.stage, r.rb, key_flags, st,
                           step);
    };
    // the exact ranges of dimensions [dk, d) on the second stream
    auto rest_passes = [&] {
        HJ_HIP(hipEventRecord(c.fork, s));
        HJ_HIP(hipStreamWaitEvent(c.stream2, c.fork, 0));
        range_pass(rp[2], P.Q.pts, P.Q.off, dk, d, false, c.stream2, 1);
        if (!self) range_pass(rp[3], P.T.pts, P.T.off, dk, d, false, c.stream2, 1);
    };
    range_pass(rp[0], P.Q.pts, P.Q.off, 0, dk, true, s, 1);
    if (!self) range_pass(rp[1], P.T.pts, P.T.off, 0, dk, false, s, 1);
    if (sample_try) {  // every dimension of every SAMPLE_STEP-th row, after the first pass
        HJ_HIP(hipEventRecord(c.fork, s));
        HJ_HIP(hipStreamWaitEvent(c.stream2, c.fork, 0));
        range_pass(rp[2], P.Q.pts, P.Q.off, 0, d, false, c.stream2, SAMPLE_STEP);
        if (!self) range_pass(rp[3], P.T.pts, P.T.off, 0, d, false, c.stream2, SAMPLE_STEP);
    } else if (rest) {
        rest_passes();
    }
    std::vector<double> qlo, qhi, tlo, thi;  // [nseg][dk], then [nseg][d]
    HJ_HIP(hipStreamSynchronize(s));
    seg_bounds_take(rp[0].rb, nseg, dk, qlo, qhi);
    if (!self) seg_bounds_take(rp[1].rb, nseg, dk, tlo, thi);
    // the bounds of dimensions [dk, d) appended to every segment's first dk
    // (lk, hk: [nseg][dk] or [nseg][d] -> [nseg][d]), from lr, hr ([nseg][nr],
    // dimension dk at column c0); with `margin` widened by SAMPLE_MARGIN of
    // their width on each side
    auto widen = [&](std::vector<double> &lk, std::vector<double> &hk,
                     const std::vector<double> &lr, const std::vector<double> &hr, int nr, int c0,
                     bool margin) {
        const size_t w0 = lk.size() / nseg;  // dk, or d when replacing a sample
        std::vector<double> lf((size_t)nseg * d), hf((size_t)nseg * d);
        for (uint32_t k = 0; k < nseg; k++) {
            std::copy(lk.begin() + (size_t)k * w0, lk.begin() + (size_t)k * w0 + dk,
                      lf.begin() + (size_t)k * d);
            std::copy(hk.begin() + (size_t)k * w0, hk.begin() + (size_t)k * w0 + dk,
                      hf.begin() + (size_t)k * d);
            for (int j = 0; j < d - dk; j++) {
                double l = lr[(size_t)k * nr + c0 + j], h = hr[(size_t)k * nr + c0 + j];
                if (margin && l <= h) {
                    const double m = SAMPLE_MARGIN * (h - l);
                    l -= m;
                    h += m;
                }
                lf[(size_t)k * d + dk + j] = l;
                hf[(size_t)k * d + dk + j] = h;
            }
        }
        lk.swap(lf);
        hk.swap(hf);
    };
    auto widen_exact = [&](std::vector<double> &lk, std::vector<double> &hk, const PinnedBuf &rb) {
        std::vector<double> lr, hr;
        seg_bounds_take(rb, nseg, d - dk, lr, hr);
        widen(lk, hk, lr, hr, d - dk, 0, false);
    };
    // The sample stands in for the exact ranges only where its ranges of the
    // key dimensions, whose exact ranges are known, come within a quarter of
    // the margin of them: data with long tails (where the points past a
    // sample's extremes are many) take the exact ranges at once.
    bool sampled = false;
    std::vector<double> sql, sqh, stl, sth;  // the sample's ranges [nseg][d]
    if (sample_try) {
        HJ_HIP(hipStreamSynchronize(c.stream2));
        seg_bounds_take(rp[2].rb, nseg, d, sql, sqh);
        if (!self) seg_bounds_take(rp[3].rb, nseg, d, stl, sth);
        auto short_tails = [&](const std::vector<double> &el, const std::vector<double> &eh,
                               const std::vector<double> &sl, const std::vector<double> &sh) {
            for (int j = 0; j < dk; j++) {
                const double w = eh[j] - el[j], tol = 0.25 * SAMPLE_MARGIN * w;
                if (w > 0 && !(sl[j] - el[j] <= tol && eh[j] - sh[j] <= tol)) return false;
            }
            return true;
        };
        sampled = short_tails(qlo, qhi, sql, sqh) && (self || short_tails(tlo, thi, stl, sth));
        if (!sampled) rest_passes();
    }
    const int pbits = bits_for(nseg - 1);
    // the bounds of the leading nd dimensions of segment k's join: its query
    // side's, and with the target side's merged in
    auto join_bounds = [&](uint32_t k, int nd, std::vector<double> &lo, std::vector<double> &hi) {
        std::copy(qlo.begin() + (size_t)k * nd, qlo.begin() + (size_t)(k + 1) * nd, lo.begin());
        std::copy(qhi.begin() + (size_t)k * nd, qhi.begin() + (size_t)(k + 1) * nd, hi.begin());
        if (!self) {
            for (int j = 0; j < nd; j++) {
                lo[j] = std::min(lo[j], tlo[(size_t)k * nd + j]);
                hi[j] = std::max(hi[j], thi[(size_t)k * nd + j]);
            }
        }
    };
    // (per-job host work of a batch of many small joins: no allocations
    // inside the loops besides each grid's own tables)
    prep.grids.resize(nseg);
    std::vector<double> lo(d), hi(d);
    for (uint32_t k = 0; k < nseg; k++) {
        join_bounds(k, dk, lo, hi);
        prep.grids[k] = grid_keys(lo.data(), hi.data(), dk, d, P.eps[k], 64 - pbits, o.active,
                                  (P.Q.off[k + 1] - P.Q.off[k]) +
                                      (self ? 0 : P.T.off[k + 1] - P.T.off[k]));
    }
    int L = 0, kd_max = 1;
    for (const Grid &g : prep.grids) {
        L = std::max(L, g.key_dims * g.bits);
        kd_max = std::max(kd_max, g.key_dims);
    }
    const int sort_bits = L + pbits;
    const uint64_t local_mask = L >= 64 ? ~0ull : (((uint64_t)1 << L) - 1);
    // side layouts and the join segment table
    Side &Q = prep.Q;
    Side &T = prep.T;
    Q = Side{};
    T = Side{};
    side_layout(Q, P.Q.off, P.Q.pts, d, o.kblock);
    if (!self) side_layout(T, P.T.off, P.T.pts, d, o.kblock);
    FilterParams fp{};
    if (filter) fp = filter_layout(d, o.kernel == HJ_KERNEL_MFMA_BF16, o.split);
    if (filter && o.frames && frames_possible(fp, d)) fp.frames = 1;
    prep.kernel = o.kernel;
    std::vector<SegDev> &sh = prep.segs_host;
    sh.assign(nseg, SegDev{});
    for (uint32_t k = 0; k < nseg; k++) {
        SegDev &v = sh[k];
        v.qoff = Q.off[k];
        v.qn = Q.off[k + 1] - Q.off[k];
        const Side &TS = self ? Q : T;
        v.toff = TS.off[k];
        v.tn = TS.off[k + 1] - TS.off[k];
        v.qprefix = L >= 64 ? 0 : ((uint64_t)k << L);
        v.tprefix = v.qprefix;
        v.local_mask = local_mask;
        v.qtile0 = Q.tile0[k];
        v.ttile0 = TS.tile0[k];
        v.grid = k;
        v.self = self ? 1 : 0;
        v.eps2 = P.eps[k] * P.eps[k];
    }
    // the grids' operand part, the frame lattices and the thresholds, from
    // the bounds of every dimension; with zcheck each grid also gets the
    // bounds on |z| its points are checked against (Grid::zmax)
    auto operand_part = [&](bool zcheck) {
        fp.frames_fast = 0;
        for (uint32_t k = 0; k < nseg; k++) {
            const double pq =
                norm2_bound(qlo.data() + (size_t)k * d, qhi.data() + (size_t)k * d, d);
            const double pmax_sum =
                self ? 2 * pq
                     : pq + norm2_bound(tlo.data() + (size_t)k * d, thi.data() + (size_t)k * d, d);
            Grid &g = prep.grids[k];
            join_bounds(k, d, lo, hi);
            grid_operands(g, lo.data(), hi.data(), 0);
            g.zmax.clear();
            if (zcheck) {
                g.zmax.resize(d);
                for (int j = 0; j < d; j++) {
                    const bool empty = lo[j] > hi[j];
                    g.zmax[j] = empty ? 0.0
                                      : std::max(hi[j] - g.centre[j], g.centre[j] - lo[j]) *
                                            g.scale;
                }
            }
            if (fp.frames) g.fr_q = frame_lattice(lo.data(), hi.data(), g, fp.bf16 != 0);
            SegDev &v = sh[k];
            v.t_in = v.t_out = 0;
            if (filter) {
                filter_thresholds(fp, d, g.scale, g.scale * g.mbound, P.eps[k], &v.t_in,
                                  &v.t_out);
                frame_consts(fp, d, g.scale, g.scale * g.mbound, P.eps[k], &v.fc);
                // a wide grid-frame band leaves many tile pairs to the slow
                // path: the fast path then tests in the tiles' own frames too
                if (fp.frames &&
                    std::sqrt(-2.0 * v.t_out) / (g.scale * P.eps[k]) - 1 > FRAME_FAST_BAND)
                    fp.frames_fast = 1;
            }
            fp64_thresholds(d, P.eps[k], pmax_sum, &v.x_in, &v.x_out);
        }
    };
    if (sampled) {
        widen(qlo, qhi, sql, sqh, d, dk, true);
        if (!self) widen(tlo, thi, stl, sth, d, dk, true);
    }
    if (!rest || sampled) operand_part(sampled);
    std::vector<uint64_t> words(P.Q.off.begin(), P.Q.off.end());
    if (!self) words.insert(words.end(), P.T.off.begin(), P.T.off.end());
    const uint64_t *qoff_dev = upload_segs(c, sh, words);
    const uint64_t *toff_dev = self ? qoff_dev : qoff_dev + (nseg + 1);
    const SegDev *segs = c.segs.as<SegDev>();
    prep.segs = segs;
    c.gt.upload(prep.grids, s);
    // the tables go up again after a change (same sizes, same device
    // addresses), in stream order behind what read them
    auto reupload = [&] {
        if (upload_segs(c, sh, words) != qoff_dev || c.segs.as<SegDev>() != segs)
            throw InvalidArg{"internal: segment table moved"};
        c.gt.upload(prep.grids, s);
    };
    // 2. keys, sort
    side_keys(c, c.A, Q, segs, nseg, qoff_dev, false, c.gt, kd_max);
    if (!self) side_keys(c, c.B, T, segs, nseg, toff_dev, true, c.gt, kd_max);
    ph.mark(1);
    side_sort(c, c.A, Q, sort_bits);
    if (!self) side_sort(c, c.B, T, sort_bits);
    if (rest && !sampled) {
        // the exact ranges of the other dimensions, taken beside the keys
        HJ_HIP(hipStreamSynchronize(c.stream2));
        widen_exact(qlo, qhi, rp[2].rb);
        if (!self) widen_exact(tlo, thi, rp[3].rb);
        operand_part(false);
        reupload();
    }
    ph.mark(2);
    // 3. tile index, then the work list on one stream while the other writes
    // the operand layouts (neither reads the other's output)
    SideNeeds nq, nt;
    nq.bounds = true;
    nq.frag_row = filter;
    nq.frag_col = filter && self;
    nq.fp64_rows = !filter || need_layout;
    nq.fp64_cols = self && (!filter || need_layout);
    nt.frag_col = filter;
    nt.fp64_cols = !filter || need_layout;
    // the tile -> segment tables on the main stream; then the search index
    // (first keys, tile bounds) and the work list (a filter's list is emitted
    // without a host wait for its length) on one stream and the operand
    // layouts on the other.  The distance kernel runs on the main stream
    // after the one expected to end last, so that it follows it without a
    // wait across streams: the work list for one join (at n = 1M, d = 64 it
    // takes longer than the operands), the operands for a batch of many (whose
    // rows are gathered for all the jobs' operands)
    hipStream_t s2 = c.stream2;
    const bool list_main = nseg == 1;
    hipStream_t so = list_main ? s2 : s, sl = list_main ? s : s2;
    side_tile_segments(c, c.A, Q, tile_owners(Q), s);
    if (!self) side_tile_segments(c, c.B, T, tile_owners(T), s);
    HJ_HIP(hipEventRecord(c.fork, s));
    HJ_HIP(hipStreamWaitEvent(s2, c.fork, 0));
    uint32_t *opflag = sampled ? c.opflag : nullptr;
    side_operands(c, c.A, Q, segs, false, c.gt, fp, nq, opflag, so);
    if (!self) side_operands(c, c.B, T, segs, true, c.gt, fp, nt, opflag, so);
    // the count's counters are cleared behind the operands
    if (!worklist_only) zero_count_state(c, nseg, so);
    side_search_index(c, c.A, Q, segs, false, nq, sl);
    if (!self) side_search_index(c, c.B, T, segs, true, nt, sl);
    Side Tw = self ? Q : T;  // the target side's index
    if (self) T = Q;
    // active dimensions chosen on the device where the caller left them open;
    // the work-list buffers are sized for the largest possible choice
    bool any_auto = false;
    for (const Grid &g : prep.grids) any_auto |= g.active < 0;
    const GridDev *gd_host = nullptr;  // pinned, after the work-list size
    if (any_auto) {
        const uint32_t samples = active_samples(nseg);
        c.est.ensure((size_t)nseg * 8 * sizeof(uint64_t));
        launch_choose_active(Q.tlo, Q.thi, Tw.keys, Tw.tfirst, segs, nseg,
                             c.gt.grids.as<GridDev>(), samples, c.est.as<uint64_t>(),
                             code_cost(o.kernel), sl);
        // read back with the work-list size (build_worklist synchronizes)
        char *h = static_cast<char *>(
            c.rb_list.ensure(WL_READBACK_GRIDS + nseg * sizeof(GridDev)));
        gd_host = reinterpret_cast<const GridDev *>(h + WL_READBACK_GRIDS);
        static_assert(sizeof(GridDev) % sizeof(uint64_t) == 0, "grid table words");
        readback_u64(c.rb_list, WL_READBACK_GRIDS, c.gt.grids.as<uint64_t>(),
                     nseg * sizeof(GridDev) / sizeof(uint64_t), sl);
    }
    int max_iv = 1;
    for (const Grid &g : prep.grids) {
        const int a = g.active >= 0 ? g.active : std::min(6, g.key_dims);
        const int total = ipow3(a);
        max_iv = std::max(max_iv, self ? total - (total - 1) / 2 : total);
    }
    WorkListPending wp = worklist_start(c, Q, Tw, segs, c.gt.grids.as<GridDev>(), max_iv,
                                        o.order, sl, filter && !worklist_only);
    prep.wl = worklist_finish(c, wp);
    HJ_HIP(hipEventRecord(c.join, s2));
    HJ_HIP(hipStreamWaitEvent(s, c.join, 0));
    prep.fp = fp;
    ph.mark(3);
    // the chosen active dimensions are in pinned memory once the stream has
    // passed the read-back (the count ends with a synchronization)
    auto resolve_active = [&] {
        if (any_auto)
            for (uint32_t k = 0; k < nseg; k++)
                if (prep.grids[k].active < 0) prep.grids[k].active = (int)gd_host[k].active;
    };
    if (worklist_only) {
        HJ_HIP(hipStreamSynchronize(s));
        resolve_active();
        R.wl = prep.wl;
        return;
    }
    // 4. count
    run_count(c, Q, T, segs, sh, prep.wl, o, fp, self, R, opflag, true);
    R.sampled_ranges = sampled ? 1 : 0;
    if (R.ranges_missed) {
        // a point lay outside the sampled ranges: the exact ranges of the
        // other dimensions, then the operand part, the operands and the
        // count again (the keys, the sort and the work list stand)
        range_pass(rp[2], P.Q.pts, P.Q.off, dk, d, false, s, 1);
        if (!self) range_pass(rp[3], P.T.pts, P.T.off, dk, d, false, s, 1);
        HJ_HIP(hipStreamSynchronize(s));
        widen_exact(qlo, qhi, rp[2].rb);
        if (!self) widen_exact(tlo, thi, rp[3].rb);
        operand_part(false);
        reupload();
        prep.fp = fp;
        side_operands(c, c.A, Q, segs, false, c.gt, fp, nq);
        if (!self) {
            side_operands(c, c.B, T, segs, true, c.gt, fp, nt);
        } else {
            T = Q;
        }
        run_count(c, Q, T, segs, sh, prep.wl, o, fp, self, R);
        R.sampled_ranges = 2;
    }
    prep.wl = R.wl;  // with the length read back
    ph.mark(4);
    HJ_HIP(hipEventSynchronize(c.ev[4]));
    resolve_active();
    if (st) {
        st->grid_ms = ph.ms(0, 1);
        st->sort_ms = ph.ms(1, 2);
        st->index_ms = ph.ms(2, 3);
        st->join_ms = ph.ms(3, 4);
        double denom = 0;
        for (const SegDev &v : sh)
            denom += self ? (double)v.qn * ((double)v.qn - 1) / 2 : (double)v.qn * (double)v.tn;
        fill_stats(st, R, prep, o, denom);
    }
}

// ---- contexts ----------------------------------------------------------------------
int ctx_init(Ctx &c, int device) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HJ_ERR_HIP;
    if (device < 0 || device >= ndev) return HJ_ERR_INVALID;
    c.device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c.stream) != hipSuccess ||
        hipStreamCreate(&c.stream2) != hipSuccess ||
        hipEventCreateWithFlags(&c.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c.join, hipEventDisableTiming) != hipSuccess)
        return HJ_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c.num_cus = prop.multiProcessorCount;
        c.name = prop.gcnArchName;
    }
    for (auto &e : c.ev)
        if (hipEventCreate(&e) != hipSuccess) return HJ_ERR_HIP;
    return HJ_OK;
}

void ctx_release(Ctx &c) {
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    c.A.release();
    c.B.release();
    c.gt.release();
    for (DevBuf *b : {&c.minmax, &c.scratch_hist, &c.scratch_scan, &c.flags, &c.ivals, &c.icount,
                      &c.ioff, &c.pairs[0], &c.pairs[1], &c.hkeys[0], &c.hkeys[1], &c.counters,
                      &c.est, &c.seg_counts, &c.band, &c.csr_rows, &c.csr_cols, &c.csr_raw,
                      &c.segs})
        b->release();
    for (Ctx::RangePass &r : c.ranges) {
        r.out.release();
        r.tmp.release();
    }
    for (auto &e : c.ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (c.stream2) (void)hipStreamSynchronize(c.stream2);
    for (hipEvent_t *e : {&c.fork, &c.join}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    if (c.stream) (void)hipStreamDestroy(c.stream);
    if (c.stream2) (void)hipStreamDestroy(c.stream2);
    c.stream = c.stream2 = nullptr;
}

// ---- pair lists ------------------------------------------------------------------
uint64_t csr_fill(Ctx &c, const JoinPrep &prep, const SegDev *segs_emit, const Opts &o,
                  bool self, uint64_t np, int64_t *rowptr, int64_t *colidx, uint64_t cap,
                  bool dev_out) {
    hipStream_t s = c.stream;
    const Side &Q = prep.Q, &T = prep.T;
    const uint64_t nq = Q.n;
    const uint64_t nk = self ? 2 * np : np;
    int64_t *drow = rowptr;  // device outputs are written in place
    if (!dev_out) drow = static_cast<int64_t *>(c.csr_rows.ensure((nq + 1) * sizeof(int64_t)));
    uint64_t *ctr = zeroed_counters(c, s);
    if (nk == 0) {
        launch_zero_u64(reinterpret_cast<uint64_t *>(drow), nq + 1, s);
    } else {
        uint64_t *raw = static_cast<uint64_t *>(c.csr_raw.ensure(np * sizeof(uint64_t)));
        if (prep.kernel == HJ_KERNEL_FP64) {
            JoinArgs ja{prep.wl.pairs, prep.wl.npairs, segs_emit, Q.tile_seg, ctr, nullptr,
                        o.grid, c.num_cus, o.tile_rows, o.chunk};
            launch_fp64_pairs(ja, Q.rows, T.cols, Q.dpad, self, raw, np, s);
        } else {
            // the filter appends its certain pairs and the FP64 recheck its
            // accepted band pairs: the pair list at filter speed
            const size_t nseg = std::max<size_t>(prep.segs_host.size(), 1);
            c.seg_counts.ensure(nseg * sizeof(uint64_t));
            JoinArgs ja{prep.wl.pairs, prep.wl.npairs, segs_emit, Q.tile_seg, ctr,
                        c.seg_counts.as<uint64_t>(), o.grid, c.num_cus, o.tile_rows, o.chunk,
                        reinterpret_cast<uint32_t *>(ctr + 8)};
            ja.rows_grouped = prep.wl.rows_grouped;
            set_frames(ja, Q, T, prep.fp);
            FilterOut out;
            out.pairs = raw;
            out.pairs_cap = np;
            out.band_cap = std::max<uint64_t>((uint64_t)1 << 20, c.band.bytes / sizeof(uint64_t));
            for (int pass = 0; pass < 2; pass++) {
                out.band = static_cast<uint64_t *>(c.band.ensure(out.band_cap * sizeof(uint64_t)));
                zeroed_counters(c, s);
                launch_mfma_count(ja, Q.frow, T.fcol, prep.fp, self, out, s);
                launch_band_recheck(Q.pts, Q.perm, T.pts, T.perm, segs_emit, Q.tile_seg, Q.d,
                                    ctr + 4, out, ctr, nullptr, c.num_cus, s);
                uint64_t *h = static_cast<uint64_t *>(c.rb_counts.ensure(8 * sizeof(uint64_t)));
                HJ_HIP(hipMemcpyAsync(h, ctr, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
                HJ_HIP(hipStreamSynchronize(s));
                if (h[4] <= out.band_cap) break;
                out.band_cap = h[4];  // the band queue overflowed: size it and run again
            }
        }
        // keys / vals ping-pong in the hkeys and csr_cols buffers
        for (int b = 0; b < 2; b++) c.hkeys[b].ensure(nk * sizeof(uint64_t));
        c.csr_cols.ensure(2 * nk * sizeof(uint32_t));
        uint64_t *kk[2] = {c.hkeys[0].as<uint64_t>(), c.hkeys[1].as<uint64_t>()};
        uint32_t *vv[2] = {c.csr_cols.as<uint32_t>(), c.csr_cols.as<uint32_t>() + nk};
        launch_map_pairs(raw, np, segs_emit, Q.tile_seg, Q.perm, T.perm, self, kk[0], vv[0], s);
        int which = 0;
        radix_sort_u64_u32(kk, vv, nk, 32 + bits_for(nq - 1), &which, c.scratch_hist,
                           c.scratch_scan, s);
        launch_csr_rowptr(kk[which], nk, nq, drow, s);
        if (colidx && nk <= cap) {
            // host colidx is staged in the other key buffer
            int64_t *dcol = dev_out ? colidx : reinterpret_cast<int64_t *>(kk[which ^ 1]);
            launch_csr_cols(kk[which], nk, dcol, s);
            if (!dev_out) copy_d2h(c, colidx, dcol, nk * sizeof(int64_t));
        }
    }
    if (!dev_out) copy_d2h(c, rowptr, drow, (nq + 1) * sizeof(int64_t));
    uint64_t *written = static_cast<uint64_t *>(c.rb_counts.ensure(8 * sizeof(uint64_t)));
    HJ_HIP(hipMemcpyAsync(written, ctr + 6, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HJ_HIP(hipStreamSynchronize(s));
    if (nk && *written != np) throw InvalidArg{"internal: pair pass disagrees with count"};
    return nk;
}

// ---- resident index ----------------------------------------------------------------
void index_build(Ctx &c, Index &ix, const double *pts, uint64_t n, int d, double eps_max,
                 const Opts &o) {
    hipStream_t s = c.stream;
    check_sizes(n, (uint32_t)d);
    if (n == 0) throw InvalidArg{"an index needs at least one point"};
    ix.device = c.device;
    ix.d = d;
    ix.n = n;
    ix.eps_max = eps_max;
    ix.kernel = o.kernel;
    const size_t bytes = (size_t)n * d * sizeof(double);
    ix.pts.ensure(bytes);
    HJ_HIP(hipMemcpyAsync(ix.pts.p, pts, bytes, hipMemcpyDeviceToDevice, s));
    const std::vector<uint64_t> off = {0, n};
    std::vector<double> lo, hi;
    seg_bounds(c, ix.pts.as<double>(), off, d, lo, hi);
    ix.grid = grid_from_bounds(lo.data(), hi.data(), d, eps_max, INDEX_KEY_BITS, o.active,
                               INDEX_HEADROOM, n);
    ix.pmax = norm2_bound(lo.data(), hi.data(), d);
    const Grid &g = ix.grid;
    ix.local_bits = g.key_dims * g.bits;
    const bool filter = o.kernel != HJ_KERNEL_FP64;
    ix.fp = FilterParams{};
    if (filter) ix.fp = filter_layout(d, o.kernel == HJ_KERNEL_MFMA_BF16, o.split);
    std::vector<SegDev> seg(1, SegDev{});
    seg[0].qn = seg[0].tn = n;
    seg[0].local_mask = ix.local_bits >= 64 ? ~0ull : (((uint64_t)1 << ix.local_bits) - 1);
    seg[0].eps2 = eps_max * eps_max;
    const uint64_t *doff = upload_segs(c, seg, off);
    const SegDev *segs = c.segs.as<SegDev>();
    std::vector<Grid> grids = {g};
    ix.gt.upload(grids, s);
    ix.T = Side{};
    side_layout(ix.T, off, ix.pts.as<double>(), d, o.kblock);
    side_keys(c, ix.B, ix.T, segs, 1, doff, true, ix.gt, std::max(1, g.key_dims));
    side_sort(c, ix.B, ix.T, ix.local_bits);
    SideNeeds need;
    need.fp64_cols = true;  // FP64 kernel, pair lists
    need.frag_col = filter;
    side_layouts(c, ix.B, ix.T, segs, tile_owners(ix.T), true, ix.gt, ix.fp, need);
    if (ix.grid.active < 0) {
        c.counters.ensure(8 * sizeof(uint64_t));
        uint64_t *cnt = c.counters.as<uint64_t>();
        launch_prefix_classes(ix.T.keys, segs, 1, ix.gt.grids.as<GridDev>(), cnt, n, s);
        uint64_t h[8];
        HJ_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
        HJ_HIP(hipStreamSynchronize(s));
        int a = 0;
        for (int j = 1; j <= std::min(6, g.key_dims); j++)
            if (h[j] > 0 && (double)n / (double)h[j] >= AUTO_CLASS_POINTS) a = j;
        ix.grid.active = a;
        grids[0].active = a;
        ix.gt.upload(grids, s);
    }
    HJ_HIP(hipStreamSynchronize(s));
}

void run_index_query(Ctx &c, const Index &ix, const SideSpec &Qs, const std::vector<double> &eps,
                     const Opts &o, JoinOut &R, JoinPrep &prep, hj_stats *st, bool need_layout) {
    hipStream_t s = c.stream;
    Phase ph{c.ev, s};
    const int d = ix.d;
    const uint32_t nseg = (uint32_t)(Qs.off.size() - 1);
    if (eps.size() != nseg) throw InvalidArg{"internal: segment tables disagree"};
    if (nseg > INDEX_MAX_REQUESTS) throw InvalidArg{"too many requests in one batch"};
    for (double e : eps)
        if (!(e > 0) || !(e <= ix.eps_max))
            throw InvalidArg{"request eps must be in (0, eps_max of the index]"};
    ph.mark(0);
    const Grid &g = ix.grid;
    std::vector<double> qlo, qhi;
    seg_bounds(c, Qs.pts, Qs.off, d, qlo, qhi);
    // the filter needs every scaled query coordinate vector within 2^14 of
    // the centre (operand range); otherwise the batch takes the FP64 kernel
    bool filter = ix.kernel != HJ_KERNEL_FP64 && o.kernel != HJ_KERNEL_FP64;
    std::vector<double> M(nseg);
    for (uint32_t k = 0; k < nseg; k++) {
        const double rq =
            radius_bound(qlo.data() + (size_t)k * d, qhi.data() + (size_t)k * d, g.centre);
        M[k] = g.scale * std::max(g.mbound, rq);
        if (M[k] > std::ldexp(1.0, 14)) filter = false;
    }
    Opts ok = o;
    ok.kernel = filter ? ix.kernel : HJ_KERNEL_FP64;
    const int pbits = bits_for(nseg - 1);
    const int L = ix.local_bits;
    const uint64_t local_mask = L >= 64 ? ~0ull : (((uint64_t)1 << L) - 1);
    Side &Q = prep.Q;
    Q = Side{};
    side_layout(Q, Qs.off, Qs.pts, d, ix.T.dpad % 16 == 0 ? 16 : 8);
    prep.T = ix.T;
    prep.grids = {g};
    prep.fp = ix.fp;
    prep.kernel = ok.kernel;
    std::vector<SegDev> &sh = prep.segs_host;
    sh.assign(nseg, SegDev{});
    for (uint32_t k = 0; k < nseg; k++) {
        SegDev &v = sh[k];
        v.qoff = Q.off[k];
        v.qn = Q.off[k + 1] - Q.off[k];
        v.toff = 0;
        v.tn = ix.n;
        v.qprefix = (uint64_t)k << L;
        v.tprefix = 0;
        v.local_mask = local_mask;
        v.qtile0 = Q.tile0[k];
        v.ttile0 = 0;
        v.grid = 0;
        v.self = 0;
        v.eps2 = eps[k] * eps[k];
        if (filter) filter_thresholds(ix.fp, d, g.scale, M[k], eps[k], &v.t_in, &v.t_out);
        fp64_thresholds(d, eps[k],
                        ix.pmax + norm2_bound(qlo.data() + (size_t)k * d,
                                              qhi.data() + (size_t)k * d, d),
                        &v.x_in, &v.x_out);
    }
    const uint64_t *qoff_dev = upload_segs(c, sh, Qs.off);
    const SegDev *segs = c.segs.as<SegDev>();
    prep.segs = segs;
    side_keys(c, c.A, Q, segs, nseg, qoff_dev, false, ix.gt, std::max(1, g.key_dims));
    ph.mark(1);
    side_sort(c, c.A, Q, L + pbits);
    ph.mark(2);
    SideNeeds nq;
    nq.bounds = true;
    nq.frag_row = filter;
    nq.fp64_rows = !filter || need_layout;
    side_tile_segments(c, c.A, Q, tile_owners(Q), s);
    HJ_HIP(hipEventRecord(c.fork, s));
    HJ_HIP(hipStreamWaitEvent(c.stream2, c.fork, 0));
    side_operands(c, c.A, Q, segs, false, ix.gt, ix.fp, nq);
    side_search_index(c, c.A, Q, segs, false, nq, c.stream2);
    WorkListPending wp = worklist_start(c, Q, ix.T, segs, ix.gt.grids.as<GridDev>(),
                                        ipow3(g.active), o.order, c.stream2, filter);
    prep.wl = worklist_finish(c, wp);
    HJ_HIP(hipEventRecord(c.join, c.stream2));
    HJ_HIP(hipStreamWaitEvent(s, c.join, 0));
    ph.mark(3);
    run_count(c, Q, ix.T, segs, sh, prep.wl, ok, ix.fp, false, R);
    prep.wl = R.wl;  // with the length read back
    ph.mark(4);
    HJ_HIP(hipEventSynchronize(c.ev[4]));
    if (st) {
        st->grid_ms = ph.ms(0, 1);
        st->sort_ms = ph.ms(1, 2);
        st->index_ms = ph.ms(2, 3);
        st->join_ms = ph.ms(3, 4);
        double denom = 0;
        for (const SegDev &v : sh) denom += (double)v.qn * (double)v.tn;
        fill_stats(st, R, prep, ok, denom);
    }
}

}  // namespace hj
