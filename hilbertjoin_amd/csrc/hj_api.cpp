// hj_api.cpp -- C ABI of the MI355X epsilon join: contexts, device memory,
// the join entry points (single, batched, pair lists) and their helpers.
// The pipeline itself is in hj_join.cpp.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>

#include "hj_join.h"

using namespace hj;

namespace {

constexpr int DEFAULT_KBLOCK = 4;

int set_error(hj_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->c.err = msg;
    return code;
}

template <class F> int guarded(hj_ctx *ctx, F &&f) {
    if (!ctx) return HJ_ERR_INVALID;
    return guarded_call(ctx->c.err, std::forward<F>(f));
}

const double *upload(Ctx &c, DevBuf &buf, const double *host, uint64_t n, uint32_t d) {
    size_t bytes = (size_t)n * d * sizeof(double);
    buf.ensure(std::max<size_t>(bytes, 8));
    copy_h2d(c, buf.p, host, bytes);
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

Problem single_problem(const double *A, uint64_t nA, const double *B, uint64_t nB, uint32_t d,
                       double eps, bool self) {
    Problem P;
    P.d = (int)d;
    P.self = self;
    P.Q.pts = A;
    P.Q.n = nA;
    P.Q.off = {0, nA};
    if (!self) {
        P.T.pts = B;
        P.T.n = nB;
        P.T.off = {0, nB};
    }
    P.eps = {eps};
    return P;
}

// ---- autotuner -------------------------------------------------------------------
hj_opts to_c_opts(const Opts &o) {
    hj_opts r;
    std::memset(&r, 0, sizeof(r));
    r.kernel = o.kernel;
    r.active_dims = o.active;
    r.order = o.order;
    r.kblock = o.kblock;
    r.grid = o.grid;
    r.filter_split = o.split;
    r.tile_rows = o.tile_rows;
    r.mfma_chunk = o.chunk;
    r.tune = 0;
    r.filter_frames = o.frames ? 0 : -1;
    return r;
}

TuneKey tune_key(const Problem &P, const Opts &o, bool batch) {
    const uint64_t n = P.Q.n + (P.self ? 0 : P.T.n);
    int l2 = 0;
    while (l2 < 63 && ((uint64_t)1 << (l2 + 1)) <= n) l2++;
    std::vector<double> e = P.eps;
    std::nth_element(e.begin(), e.begin() + e.size() / 2, e.end());
    const int eps4 = (int)std::floor(4.0 * std::log2(e[e.size() / 2]));
    return TuneKey{o.kernel, P.self ? 1 : 0, batch ? 1 : 0, P.d, l2, eps4};
}

// Coordinate descent over the knobs, each candidate timed as the best of two
// whole joins (host wall clock around the call, after its final sync).
Opts autotune(hj_ctx *ctx, const Problem &P, const Opts &base, bool batch, hj_tune_report *rep) {
    Ctx &c = ctx->c;
    int trials = 0;
    auto time_it = [&](const Opts &o, int *active_out) -> double {
        double best = INFINITY;
        for (int r = 0; r < 2; r++) {
            hj_stats st;
            JoinOut R;
            JoinPrep prep;
            const double t0 = now_ms();
            run_problem(c, P, o, R, prep, &st, false);
            best = std::min(best, now_ms() - t0);
            if (active_out) *active_out = st.active_dims;
        }
        trials++;
        return best;
    };
    Opts cur = base;
    cur.tune = false;
    int a0 = -1;
    double cur_ms = time_it(cur, &a0);
    const double base_ms = cur_ms;
    auto consider = [&](const Opts &cand) {
        const double t = time_it(cand, nullptr);
        if (t < cur_ms) {
            cur_ms = t;
            cur = cand;
        }
    };
    // 1. active dimensions around the data-driven choice (single joins)
    if (!batch && cur.active < 0 && a0 >= 0) {
        Opts fixed = cur;
        fixed.active = a0;
        const Opts start = cur;
        for (int a : {a0 - 1, a0 + 1}) {
            if (a < 0 || a > 6) continue;
            Opts cand = start;
            cand.active = a;
            consider(cand);
        }
        (void)fixed;
    }
    // 2. persistent grid
    {
        const Opts start = cur;
        // filter: around its resident count; FP64: also far beyond it (many
        // short-lived workgroups)
        const bool fp64 = start.kernel == HJ_KERNEL_FP64;
        for (int per_cu : fp64 ? std::vector<int>{4, 64, 256} : std::vector<int>{2, 4, 8}) {
            Opts cand = start;
            cand.grid = per_cu * std::max(1, c.num_cus);
            consider(cand);
        }
    }
    // 3. tile shape of the kernel
    {
        const Opts start = cur;
        if (start.kernel == HJ_KERNEL_FP64) {
            for (int tr : {8, 16})
                for (int kb : {4, 8, 16}) {
                    if (tr == start.tile_rows && kb == start.kblock) continue;
                    Opts cand = start;
                    cand.tile_rows = tr;
                    cand.kblock = kb;
                    consider(cand);
                }
        } else {
            for (int ch : {2, 4, 8}) {
                if (ch == start.chunk) continue;
                Opts cand = start;
                cand.chunk = ch;
                consider(cand);
            }
            if (cur.split == 0) {
                Opts cand = cur;
                cand.split = -1;
                consider(cand);
            }
        }
    }
    // 4. work-list order
    {
        Opts cand = cur;
        cand.order = cur.order == HJ_ORDER_ROW ? HJ_ORDER_HILBERT : HJ_ORDER_ROW;
        consider(cand);
    }
    ctx->tuned[tune_key(P, base, batch)] = to_c_opts(cur);
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->trials = trials;
        rep->base_ms = base_ms;
        rep->best_ms = cur_ms;
        rep->best = to_c_opts(cur);
    }
    return cur;
}

// Settings for a call with opts->tune = 1: the context's tuned settings for
// the input's shape class, tuned on this input first if there are none.
Opts tuned_opts(hj_ctx *ctx, const Problem &P, const Opts &o, bool batch) {
    auto it = ctx->tuned.find(tune_key(P, o, batch));
    if (it != ctx->tuned.end()) return read_opts(&it->second);
    return autotune(ctx, P, o, batch, nullptr);
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
        if (!(eps > 0) || !std::isfinite(eps)) throw InvalidArg{"eps must be finite and > 0"};
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
        const Problem P = single_problem(dA, nA, dB, nB, d, eps, self);
        if (o.tune) o = tuned_opts(ctx, P, o, false);
        JoinOut R;
        JoinPrep prep;
        run_problem(c, P, o, R, prep, st, false);
        *count = R.count;
        if (st) st->total_ms = now_ms() - t0;
        return HJ_OK;
    });
}

// Batched joins: job k joins rows [a_off[k], a_off[k+1]) of A with itself
// (self) or with rows [b_off[k], b_off[k+1]) of B, under radius eps[k].  One
// pipeline pass and one persistent kernel launch serve all jobs.
int batch_common(hj_ctx *ctx, const double *A, const uint64_t *a_off, const double *B,
                 const uint64_t *b_off, uint32_t njobs, uint32_t d, const double *eps,
                 const hj_opts *opts, uint64_t *counts, hj_stats *st, bool self, bool host) {
    return guarded(ctx, [&]() -> int {
        double t0 = now_ms();
        zero_stats(st);
        Opts o = read_opts(opts);
        if (njobs == 0) return HJ_OK;
        if (!a_off || !eps || !counts || (!self && !b_off))
            throw InvalidArg{"null offsets, eps or counts"};
        if (njobs > (1u << 20)) throw InvalidArg{"at most 2^20 jobs per batch"};
        auto check_off = [&](const uint64_t *off) {
            if (off[0] != 0) throw InvalidArg{"offsets must start at 0"};
            for (uint32_t k = 0; k < njobs; k++)
                if (off[k + 1] < off[k]) throw InvalidArg{"offsets must not decrease"};
        };
        check_off(a_off);
        if (!self) check_off(b_off);
        const uint64_t nA = a_off[njobs], nB = self ? 0 : b_off[njobs];
        check_sizes(nA, d);
        if (!self) check_sizes(nB, d);
        if ((nA && !A) || (!self && nB && !B)) throw InvalidArg{"null point array"};
        for (uint32_t k = 0; k < njobs; k++)
            if (!(eps[k] > 0) || !std::isfinite(eps[k]))
                throw InvalidArg{"eps must be finite and > 0"};
        std::fill(counts, counts + njobs, 0);
        if (nA == 0 || (!self && nB == 0)) {
            if (st) st->total_ms = now_ms() - t0;
            return HJ_OK;
        }
        Ctx &c = ctx->c;
        const double *dA = A, *dB = B;
        if (host) {
            dA = upload(c, c.A.raw, A, nA, d);
            if (!self) dB = upload(c, c.B.raw, B, nB, d);
        }
        Problem P;
        P.d = (int)d;
        P.self = self;
        P.Q.pts = dA;
        P.Q.n = nA;
        P.Q.off.assign(a_off, a_off + njobs + 1);
        if (!self) {
            P.T.pts = dB;
            P.T.n = nB;
            P.T.off.assign(b_off, b_off + njobs + 1);
        }
        P.eps.assign(eps, eps + njobs);
        if (o.tune) o = tuned_opts(ctx, P, o, true);
        JoinOut R;
        JoinPrep prep;
        run_problem(c, P, o, R, prep, st, false);
        std::copy(R.seg_counts.begin(), R.seg_counts.end(), counts);
        if (st) st->total_ms = now_ms() - t0;
        return HJ_OK;
    });
}

// Pair lists: count, then one pass of the join's kernel that appends its
// accepted pairs as layout positions (the filter with its FP64 band recheck,
// or the FP64 pair kernel), mapped to original indices (mirrored for
// self-joins), radix sorted by (row << 32 | col), and turned into rowptr /
// colidx on the device.
// host: points, rowptr and colidx in host memory; otherwise all three are
// device pointers of the context's device and the lists are written in place.
int csr_common(hj_ctx *ctx, const double *A, uint64_t nA, const double *B, uint64_t nB,
               uint32_t d, double eps, const hj_opts *opts, int64_t *rowptr, int64_t *colidx,
               uint64_t cap, uint64_t *total, hj_stats *st, bool self, bool host) {
    return guarded(ctx, [&]() -> int {
        double t0 = now_ms();
        zero_stats(st);
        Opts o = read_opts(opts);
        check_sizes(nA, d);
        if (!self) check_sizes(nB, d);
        if (!rowptr || !total) throw InvalidArg{"rowptr and total must not be NULL"};
        if ((nA && !A) || (!self && nB && !B)) throw InvalidArg{"null point array"};
        if (!(eps > 0) || !std::isfinite(eps)) throw InvalidArg{"eps must be finite and > 0"};
        *total = 0;
        Ctx &c = ctx->c;
        if (nA == 0 || (!self && nB == 0) || (self && nA < 2)) {
            if (host) {
                for (uint64_t i = 0; i <= nA; i++) rowptr[i] = 0;
            } else {
                launch_zero_u64(reinterpret_cast<uint64_t *>(rowptr), nA + 1, c.stream);
                HJ_HIP(hipStreamSynchronize(c.stream));
            }
            if (st) st->total_ms = now_ms() - t0;
            return HJ_OK;
        }
        const double *dA = A, *dB = self ? nullptr : B;
        if (host) {
            dA = upload(c, c.A.raw, A, nA, d);
            if (!self) dB = upload(c, c.B.raw, B, nB, d);
        }
        JoinOut R;
        JoinPrep prep;
        run_problem(c, single_problem(dA, nA, dB, nB, d, eps, self), o, R, prep, st, false);
        const uint64_t nk =
            csr_fill(c, prep, prep.segs, o, self, R.count, rowptr, colidx, cap, !host);
        *total = nk;
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
                      true, true);
}

HJ_API int hj_join_csr(hj_ctx *ctx, const double *A, uint64_t nA, const double *B, uint64_t nB,
                       uint32_t d, double eps, const hj_opts *opts, int64_t *rowptr,
                       int64_t *colidx, uint64_t cap, uint64_t *total, hj_stats *stats) {
    return csr_common(ctx, A, nA, B, nB, d, eps, opts, rowptr, colidx, cap, total, stats, false,
                      true);
}

HJ_API int hj_self_join_csr_device(hj_ctx *ctx, const double *pts_dev, uint64_t n, uint32_t d,
                                   double eps, const hj_opts *opts, int64_t *rowptr_dev,
                                   int64_t *colidx_dev, uint64_t cap, uint64_t *total,
                                   hj_stats *stats) {
    return csr_common(ctx, pts_dev, n, nullptr, 0, d, eps, opts, rowptr_dev, colidx_dev, cap,
                      total, stats, true, false);
}

HJ_API int hj_join_csr_device(hj_ctx *ctx, const double *A_dev, uint64_t nA, const double *B_dev,
                              uint64_t nB, uint32_t d, double eps, const hj_opts *opts,
                              int64_t *rowptr_dev, int64_t *colidx_dev, uint64_t cap,
                              uint64_t *total, hj_stats *stats) {
    return csr_common(ctx, A_dev, nA, B_dev, nB, d, eps, opts, rowptr_dev, colidx_dev, cap, total,
                      stats, false, false);
}

HJ_API int hj_filter_band(int32_t kernel, uint32_t d, double eps, double radius,
                          int32_t filter_split, hj_filter_info *out) {
    if (!out || d == 0 || d > 65536 || !(eps > 0) || !std::isfinite(eps) || !(radius >= 0) ||
        !std::isfinite(radius) || filter_split < -1)
        return HJ_ERR_INVALID;
    if (kernel != HJ_KERNEL_MFMA_F16 && kernel != HJ_KERNEL_MFMA_BF16) return HJ_ERR_INVALID;
    const FilterParams fp = filter_layout((int)d, kernel == HJ_KERNEL_MFMA_BF16, filter_split);
    std::memset(out, 0, sizeof(*out));
    out->scale = operand_scale(radius, 0);
    filter_thresholds(fp, (int)d, out->scale, out->scale * radius, eps, &out->t_in, &out->t_out);
    out->split = fp.split;
    out->steps1 = fp.steps1;
    out->steps = fp.steps;
    out->parts = fp.parts;
    return HJ_OK;
}

HJ_API int hj_filter_pair_band(int32_t kernel, uint32_t d, double eps, double radius,
                               int32_t filter_split, float rho_a, float rho_b, float dd_bound,
                               float dd_shift, float *t_in, float *t_out) {
    if (!t_in || !t_out || d == 0 || d > 65536 || !(eps > 0) || !std::isfinite(eps) ||
        !(radius >= 0) || !std::isfinite(radius) || filter_split < -1 || !(rho_a >= 0) ||
        !(rho_b >= 0) || !(dd_bound >= 0) || !(dd_shift >= 0))
        return HJ_ERR_INVALID;
    if (kernel != HJ_KERNEL_MFMA_F16 && kernel != HJ_KERNEL_MFMA_BF16) return HJ_ERR_INVALID;
    const FilterParams fp = filter_layout((int)d, kernel == HJ_KERNEL_MFMA_BF16, filter_split);
    const double scale = operand_scale(radius, 0);
    FrameConst fc;
    frame_consts(fp, (int)d, scale, scale * radius, eps, &fc);
    frame_thresholds(fc, rho_a, rho_b, dd_bound, dd_shift, t_in, t_out);
    return HJ_OK;
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
    const int rc = ctx_init(ctx->c, device);
    if (rc != HJ_OK) {
        ctx_release(ctx->c);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return HJ_OK;
}

HJ_API int hj_ctx_destroy(hj_ctx *ctx) {
    if (!ctx) return HJ_OK;
    ctx_release(ctx->c);
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
        copy_h2d(ctx->c, dst_dev, src_host, bytes);
        HJ_HIP(hipStreamSynchronize(ctx->c.stream));
        return HJ_OK;
    });
}

HJ_API int hj_memcpy_to_host(hj_ctx *ctx, void *dst_host, const void *src_dev, uint64_t bytes) {
    return guarded(ctx, [&]() -> int {
        if (bytes && (!dst_host || !src_dev)) throw InvalidArg{"null pointer"};
        copy_d2h(ctx->c, dst_host, src_dev, bytes);
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

HJ_API int hj_batch_self_join_count(hj_ctx *ctx, const double *pts, const uint64_t *offsets,
                                    uint32_t njobs, uint32_t d, const double *eps,
                                    const hj_opts *opts, uint64_t *counts, hj_stats *stats) {
    return batch_common(ctx, pts, offsets, nullptr, nullptr, njobs, d, eps, opts, counts, stats,
                        true, true);
}

HJ_API int hj_batch_join_count(hj_ctx *ctx, const double *A, const uint64_t *a_offsets,
                               const double *B, const uint64_t *b_offsets, uint32_t njobs,
                               uint32_t d, const double *eps, const hj_opts *opts,
                               uint64_t *counts, hj_stats *stats) {
    return batch_common(ctx, A, a_offsets, B, b_offsets, njobs, d, eps, opts, counts, stats,
                        false, true);
}

HJ_API int hj_batch_self_join_count_device(hj_ctx *ctx, const double *pts_dev,
                                           const uint64_t *offsets, uint32_t njobs, uint32_t d,
                                           const double *eps, const hj_opts *opts,
                                           uint64_t *counts, hj_stats *stats) {
    return batch_common(ctx, pts_dev, offsets, nullptr, nullptr, njobs, d, eps, opts, counts,
                        stats, true, false);
}

HJ_API int hj_batch_join_count_device(hj_ctx *ctx, const double *A_dev,
                                      const uint64_t *a_offsets, const double *B_dev,
                                      const uint64_t *b_offsets, uint32_t njobs, uint32_t d,
                                      const double *eps, const hj_opts *opts, uint64_t *counts,
                                      hj_stats *stats) {
    return batch_common(ctx, A_dev, a_offsets, B_dev, b_offsets, njobs, d, eps, opts, counts,
                        stats, false, false);
}

HJ_API int hj_autotune_self(hj_ctx *ctx, const double *pts_dev, uint64_t n, uint32_t d,
                            double eps, const hj_opts *base, hj_tune_report *rep) {
    return guarded(ctx, [&]() -> int {
        Opts o = read_opts(base);
        check_sizes(n, d);
        if (n < 2 || !pts_dev) throw InvalidArg{"tuning needs at least two device points"};
        if (!(eps > 0) || !std::isfinite(eps)) throw InvalidArg{"eps must be finite and > 0"};
        autotune(ctx, single_problem(pts_dev, n, nullptr, 0, d, eps, true), o, false, rep);
        return HJ_OK;
    });
}

HJ_API int hj_autotune_join(hj_ctx *ctx, const double *A_dev, uint64_t nA, const double *B_dev,
                            uint64_t nB, uint32_t d, double eps, const hj_opts *base,
                            hj_tune_report *rep) {
    return guarded(ctx, [&]() -> int {
        Opts o = read_opts(base);
        check_sizes(nA, d);
        check_sizes(nB, d);
        if (!nA || !nB || !A_dev || !B_dev) throw InvalidArg{"tuning needs device points"};
        if (!(eps > 0) || !std::isfinite(eps)) throw InvalidArg{"eps must be finite and > 0"};
        autotune(ctx, single_problem(A_dev, nA, B_dev, nB, d, eps, false), o, false, rep);
        return HJ_OK;
    });
}

HJ_API int hj_autotune_batch_self(hj_ctx *ctx, const double *pts_dev, const uint64_t *offsets,
                                  uint32_t njobs, uint32_t d, const double *eps,
                                  const hj_opts *base, hj_tune_report *rep) {
    return guarded(ctx, [&]() -> int {
        Opts o = read_opts(base);
        if (njobs == 0 || !offsets || !eps || !pts_dev) throw InvalidArg{"empty batch"};
        if (offsets[0] != 0) throw InvalidArg{"offsets must start at 0"};
        for (uint32_t k = 0; k < njobs; k++) {
            if (offsets[k + 1] < offsets[k]) throw InvalidArg{"offsets must not decrease"};
            if (!(eps[k] > 0) || !std::isfinite(eps[k]))
                throw InvalidArg{"eps must be finite and > 0"};
        }
        check_sizes(offsets[njobs], d);
        Problem P;
        P.d = (int)d;
        P.self = true;
        P.Q.pts = pts_dev;
        P.Q.n = offsets[njobs];
        P.Q.off.assign(offsets, offsets + njobs + 1);
        P.eps.assign(eps, eps + njobs);
        autotune(ctx, P, o, true, rep);
        return HJ_OK;
    });
}

HJ_API int hj_dim_statistics(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d,
                             double eps, int32_t *reorder_dim, uint64_t *cost) {
    return guarded(ctx, [&]() -> int {
        check_sizes(n, d);
        if (!reorder_dim || (n && !pts)) throw InvalidArg{"null pointer"};
        if (!(eps > 0) || !std::isfinite(eps)) throw InvalidArg{"eps must be finite and > 0"};
        const double r = 1.0 / eps;
        std::vector<uint64_t> cst(d, 0);
        if (n) {
            Ctx &c = ctx->c;
            hipStream_t s = c.stream;
            const double *dp = upload(c, c.A.raw, pts, n, d);
            std::vector<double> lo, hi;
            seg_bounds(c, dp, {0, n}, (int)d, lo, hi);
            // per dimension: bins = cells between the smallest and largest one
            std::vector<int64_t> cmin(d);
            std::vector<uint32_t> hoff(d + 1, 0);
            uint64_t nb = 0;
            for (uint32_t k = 0; k < d; k++) {
                const double cl = std::floor(lo[k] * r), ch = std::floor(hi[k] * r);
                if (!(std::fabs(cl) < 4e18) || !(std::fabs(ch) < 4e18))
                    throw InvalidArg{"eps is too small for the coordinate range"};
                cmin[k] = (int64_t)cl;
                hoff[k] = (uint32_t)nb;
                nb += (uint64_t)((int64_t)ch - (int64_t)cl) + 1;
                if (nb > (1ull << 26))
                    throw InvalidArg{"eps is too small for the dimension statistics"};
            }
            hoff[d] = (uint32_t)nb;
            c.hkeys[1].ensure(d * (sizeof(int64_t) + sizeof(uint32_t)) + 16);
            int64_t *dcmin = c.hkeys[1].as<int64_t>();
            uint32_t *dhoff = reinterpret_cast<uint32_t *>(dcmin + d);
            HJ_HIP(hipMemcpyAsync(dcmin, cmin.data(), d * sizeof(int64_t), hipMemcpyHostToDevice,
                                  s));
            HJ_HIP(hipMemcpyAsync(dhoff, hoff.data(), d * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, s));
            c.hkeys[0].ensure(nb * sizeof(uint32_t) + 8);
            uint32_t *hist = c.hkeys[0].as<uint32_t>();
            launch_zero_u64(reinterpret_cast<uint64_t *>(hist), (nb + 1) / 2, s);
            launch_dim_hist(dp, n, (int)d, r, dcmin, dhoff, (uint32_t)nb, hist, c.num_cus, s);
            std::vector<uint32_t> h(nb);
            HJ_HIP(hipMemcpyAsync(h.data(), hist, nb * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  s));
            HJ_HIP(hipStreamSynchronize(s));
            // pairs of points whose cells in dimension k are equal or adjacent
            for (uint32_t k = 0; k < d; k++) {
                uint64_t v = 0;
                for (uint32_t b = hoff[k]; b < hoff[k + 1]; b++) {
                    v += (uint64_t)h[b] * (h[b] - (h[b] ? 1 : 0)) / 2;
                    if (b > hoff[k]) v += (uint64_t)h[b - 1] * h[b];
                }
                cst[k] = v;
            }
        }
        // most selective (fewest such pairs) first; ties keep the input order
        std::vector<int> order(d);
        for (uint32_t k = 0; k < d; k++) order[k] = (int)k;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return cst[a] < cst[b]; });
        for (uint32_t j = 0; j < d; j++) reorder_dim[order[j]] = (int32_t)j;
        if (cost) std::copy(cst.begin(), cst.end(), cost);
        return HJ_OK;
    });
}

HJ_API int hj_reorder_dimensions(hj_ctx *ctx, double *pts, uint64_t n, uint32_t d,
                                 const int32_t *reorder_dim) {
    return guarded(ctx, [&]() -> int {
        check_sizes(n, d);
        if (!reorder_dim || (n && !pts)) throw InvalidArg{"null pointer"};
        std::vector<char> seen(d, 0);
        for (uint32_t k = 0; k < d; k++) {
            const int32_t v = reorder_dim[k];
            if (v < 0 || (uint32_t)v >= d || seen[v]) throw InvalidArg{"not a permutation"};
            seen[v] = 1;
        }
        if (n == 0) return HJ_OK;
        Ctx &c = ctx->c;
        hipStream_t s = c.stream;
        const double *din = upload(c, c.A.raw, pts, n, d);
        c.B.raw.ensure((size_t)n * d * sizeof(double));
        c.hkeys[1].ensure(d * sizeof(int32_t) + 16);
        HJ_HIP(hipMemcpyAsync(c.hkeys[1].p, reorder_dim, d * sizeof(int32_t),
                              hipMemcpyHostToDevice, s));
        launch_permute_cols(din, c.B.raw.as<double>(), n, (int)d, c.hkeys[1].as<int32_t>(),
                            c.num_cus, s);
        HJ_HIP(hipMemcpyAsync(pts, c.B.raw.p, (size_t)n * d * sizeof(double),
                              hipMemcpyDeviceToHost, s));
        HJ_HIP(hipStreamSynchronize(s));
        return HJ_OK;
    });
}

HJ_API int hj_ego_sort(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d, double eps,
                       uint32_t *perm, uint64_t *keys, int32_t *key_dims, int32_t *key_bits) {
    return guarded(ctx, [&]() -> int {
        check_sizes(n, d);
        if (n && (!pts || !perm)) throw InvalidArg{"null pointer"};
        Ctx &c = ctx->c;
        const double *dp = upload(c, c.A.raw, pts, n, d);
        std::vector<uint64_t> off = {0, n};
        std::vector<double> lo, hi;
        seg_bounds(c, dp, off, (int)d, lo, hi);
        std::vector<Grid> grids = {grid_from_bounds(lo.data(), hi.data(), (int)d, eps, 64, -1, 0, 0)};
        const Grid &g = grids[0];
        std::vector<SegDev> segs(1, SegDev{});
        segs[0].qn = segs[0].tn = n;
        segs[0].local_mask = ~0ull;
        segs[0].self = 1;
        const uint64_t *doff = upload_segs(c, segs, off);
        c.gt.upload(grids, c.stream);
        Side S;
        side_layout(S, off, dp, (int)d, 8);
        if (n) {
            side_keys(c, c.A, S, c.segs.as<SegDev>(), 1, doff, false, c.gt, std::max(1, g.key_dims));
            side_sort(c, c.A, S, g.key_dims * g.bits);
            HJ_HIP(hipMemcpyAsync(perm, S.perm, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c.stream));
            if (keys)
                HJ_HIP(hipMemcpyAsync(keys, S.keys, n * sizeof(uint64_t), hipMemcpyDeviceToHost,
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
        // the work list does not depend on the distance kernel; build it with
        // the FP64 layouts
        o.kernel = HJ_KERNEL_FP64;
        JoinOut R;
        JoinPrep prep;
        run_problem(c, single_problem(dA, nA, dB, nB, d, eps, self), o, R, prep, nullptr, false,
                    true);
        const WorkList &wl = prep.wl;
        *npairs = wl.npairs;
        if (permA)
            HJ_HIP(hipMemcpyAsync(permA, prep.Q.perm, nA * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c.stream));
        if (!self && permB)
            HJ_HIP(hipMemcpyAsync(permB, prep.T.perm, nB * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c.stream));
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
