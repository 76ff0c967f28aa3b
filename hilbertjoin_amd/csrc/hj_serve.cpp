// hj_serve.cpp -- resident index, batched queries against it, and the
// serving loop (continuous batching with admission by result memory).
//
// No counterpart in the reference: it answers one join per process
// (main.cpp:102-124, mainJoin.cpp:111) and rebuilds its sort and stripes
// every time.  Here the target side of an A x B join (keys, EGO permutation,
// FP64 tile layout, MFMA column operands) is built once (index_build) and
// every batch of requests becomes the query side of one segmented join
// against it (run_index_query): one key sort over all requests (request id in
// the high key bits), one work list, one persistent kernel launch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "hj_join.h"

struct hj_index {
    hj::Index ix;
};

using namespace hj;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

template <class F> int guarded(hj_ctx *ctx, F &&f) {
    if (!ctx) return HJ_ERR_INVALID;
    return guarded_call(ctx->c.err, std::forward<F>(f));
}

int index_create_common(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d, double eps_max,
                        const hj_opts *opts, hj_index **out, bool host) {
    return guarded(ctx, [&]() -> int {
        if (!out) throw InvalidArg{"out must not be NULL"};
        *out = nullptr;
        Opts o = read_opts(opts);
        check_sizes(n, d);
        if (n == 0 || !pts) throw InvalidArg{"an index needs at least one point"};
        if (!(eps_max > 0) || !std::isfinite(eps_max))
            throw InvalidArg{"eps_max must be finite and > 0"};
        Ctx &c = ctx->c;
        const double *dp = pts;
        if (host) {
            const size_t bytes = (size_t)n * d * sizeof(double);
            c.A.raw.ensure(bytes);
            HJ_HIP(hipMemcpyAsync(c.A.raw.p, pts, bytes, hipMemcpyHostToDevice, c.stream));
            dp = c.A.raw.as<double>();
        }
        std::unique_ptr<hj_index> idx(new hj_index());
        try {
            index_build(c, idx->ix, dp, n, (int)d, eps_max, o);
        } catch (...) {
            idx->ix.release();
            throw;
        }
        *out = idx.release();
        return HJ_OK;
    });
}

void check_batch(const hj_index *idx, const uint64_t *q_off, uint32_t nreq, const double *eps) {
    if (!idx) throw InvalidArg{"null index"};
    if (nreq && (!q_off || !eps)) throw InvalidArg{"null offsets or eps"};
    if (nreq > INDEX_MAX_REQUESTS) throw InvalidArg{"at most 65536 requests per batch"};
    if (nreq && q_off[0] != 0) throw InvalidArg{"offsets must start at 0"};
    for (uint32_t k = 0; k < nreq; k++) {
        if (q_off[k + 1] < q_off[k]) throw InvalidArg{"offsets must not decrease"};
        if (!(eps[k] > 0) || !(eps[k] <= idx->ix.eps_max))
            throw InvalidArg{"request eps must be in (0, eps_max of the index]"};
    }
    if (nreq) check_sizes(q_off[nreq], (uint32_t)idx->ix.d);
}

SideSpec query_side(Ctx &c, const double *Q, const uint64_t *q_off, uint32_t nreq, int d,
                    bool host) {
    SideSpec Qs;
    Qs.n = q_off[nreq];
    Qs.off.assign(q_off, q_off + nreq + 1);
    Qs.pts = Q;
    if (host) {
        const size_t bytes = (size_t)Qs.n * d * sizeof(double);
        c.A.raw.ensure(std::max<size_t>(bytes, 8));
        if (bytes)
            HJ_HIP(hipMemcpyAsync(c.A.raw.p, Q, bytes, hipMemcpyHostToDevice, c.stream));
        Qs.pts = c.A.raw.as<double>();
    }
    return Qs;
}

int query_count_common(hj_ctx *ctx, const hj_index *idx, const double *Q, const uint64_t *q_off,
                       uint32_t nreq, const double *eps, const hj_opts *opts, uint64_t *counts,
                       hj_stats *st, bool host) {
    return guarded(ctx, [&]() -> int {
        const double t0 = now_ms();
        if (st) std::memset(st, 0, sizeof(*st));
        Opts o = read_opts(opts);
        check_batch(idx, q_off, nreq, eps);
        if (nreq && !counts) throw InvalidArg{"null counts"};
        if (nreq == 0) return HJ_OK;
        if (q_off[nreq] && !Q) throw InvalidArg{"null query points"};
        Ctx &c = ctx->c;
        if (c.device != idx->ix.device) throw InvalidArg{"index and context on different devices"};
        SideSpec Qs = query_side(c, Q, q_off, nreq, idx->ix.d, host);
        JoinOut R;
        JoinPrep prep;
        run_index_query(c, idx->ix, Qs, std::vector<double>(eps, eps + nreq), o, R, prep, st,
                        false);
        std::copy(R.seg_counts.begin(), R.seg_counts.end(), counts);
        if (st) st->total_ms = now_ms() - t0;
        return HJ_OK;
    });
}

// ---- serving loop -----------------------------------------------------------------
struct Request {
    uint64_t ticket = 0;
    std::vector<double> pts;
    uint64_t nq = 0;
    double eps = 0;
    bool pairs = false;
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    int status = HJ_OK;
    std::string err;
    uint64_t count = 0;
    uint64_t known_pairs = UINT64_MAX;  // pairs, once counted in a batch that had no room
    std::vector<int64_t> rowptr, colidx;
    uint64_t reserved = 0;  // result bytes held
    uint64_t batch = 0;
    uint32_t batch_requests = 0;

    uint64_t first_reservation() const {
        if (!pairs) return 8;
        return (nq + 1) * 8 + (known_pairs == UINT64_MAX ? 0 : known_pairs * 8);
    }
};

using ReqPtr = std::shared_ptr<Request>;

}  // namespace

struct hj_server {
    const Index *ix = nullptr;
    int device = 0;
    hj_server_opts o{};
    int kernel = HJ_KERNEL_FP64;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_init;
    int workers_up = 0, workers_failed = 0;  // worker start-up results
    std::deque<ReqPtr> pending;
    std::unordered_map<uint64_t, ReqPtr> table;
    uint64_t next_ticket = 1;
    uint64_t result_used = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::string err;
    // statistics
    std::vector<double> latencies;
    uint64_t submitted = 0, completed = 0, failed = 0, batches = 0, batch_reqs = 0;
    double first_submit = -1, last_done = 0;

    void worker_main(int wid);
    bool take_batch(std::vector<ReqPtr> &batch);
    void run_batch(Ctx &c, PinnedBuf &stage, std::vector<ReqPtr> &batch, uint64_t bid);
    void finish(std::vector<ReqPtr> &reqs, int status, const std::string &msg);
};

// Forms the next batch in arrival order (caller holds mu).  Stops at the first
// request that does not fit the point limit or the result budget, so requests
// are never overtaken.
bool hj_server::take_batch(std::vector<ReqPtr> &batch) {
    uint64_t pts = 0, reserve = 0;
    while (!pending.empty() && batch.size() < o.max_batch_requests) {
        ReqPtr r = pending.front();
        const uint64_t need = r->first_reservation();
        if (need > o.result_bytes) {  // can never be admitted
            pending.pop_front();
            r->status = HJ_ERR_CAPACITY;
            r->err = "the request's result exceeds the server's result budget";
            r->done = true;
            r->t_done = now_ms();
            failed++;
            cv_done.notify_all();
            continue;
        }
        if (!batch.empty() && pts + r->nq > o.max_batch_points) break;
        if (result_used + reserve + need > o.result_bytes) break;
        pending.pop_front();
        r->reserved = need;
        reserve += need;
        pts += r->nq;
        batch.push_back(r);
    }
    result_used += reserve;
    return !batch.empty();
}

void hj_server::finish(std::vector<ReqPtr> &reqs, int status, const std::string &msg) {
    const double t = now_ms();
    std::lock_guard<std::mutex> lk(mu);
    for (ReqPtr &r : reqs) {
        r->done = true;
        r->t_done = t;
        if (status != HJ_OK) {
            r->status = status;
            r->err = msg;
            failed++;
        } else {
            completed++;
            latencies.push_back(t - r->t_submit);
        }
        last_done = std::max(last_done, t);
    }
    cv_done.notify_all();
}

void hj_server::run_batch(Ctx &c, PinnedBuf &stage, std::vector<ReqPtr> &batch, uint64_t bid) {
    const int d = ix->d;
    const uint32_t nreq = (uint32_t)batch.size();
    std::vector<uint64_t> off(nreq + 1, 0);
    std::vector<double> eps(nreq);
    bool any_pairs = false;
    const double t_start = now_ms();
    for (uint32_t k = 0; k < nreq; k++) {
        off[k + 1] = off[k] + batch[k]->nq;
        eps[k] = batch[k]->eps;
        any_pairs |= batch[k]->pairs;
        batch[k]->t_start = t_start;
        batch[k]->batch = bid;
        batch[k]->batch_requests = nreq;
    }
    const uint64_t nq = off[nreq];
    const size_t bytes = (size_t)nq * d * sizeof(double);
    double *h = static_cast<double *>(stage.ensure(bytes));
    for (uint32_t k = 0; k < nreq; k++)
        if (batch[k]->nq)
            std::memcpy(h + (size_t)off[k] * d, batch[k]->pts.data(),
                        (size_t)batch[k]->nq * d * sizeof(double));
    c.A.raw.ensure(std::max<size_t>(bytes, 8));
    if (bytes) HJ_HIP(hipMemcpyAsync(c.A.raw.p, h, bytes, hipMemcpyHostToDevice, c.stream));
    SideSpec Qs;
    Qs.pts = c.A.raw.as<double>();
    Qs.n = nq;
    Qs.off = off;
    Opts opt;
    opt.kernel = kernel;
    opt.order = o.order;
    JoinOut R;
    JoinPrep prep;
    run_index_query(c, *ix, Qs, eps, opt, R, prep, nullptr, false);
    for (uint32_t k = 0; k < nreq; k++) batch[k]->count = R.seg_counts[k];
    std::vector<ReqPtr> ready, again;
    if (!any_pairs) {
        ready = batch;
    } else {
        // colidx memory of the pair requests, admitted in batch order
        std::vector<bool> emit(nreq, false);
        uint64_t np = 0;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (uint32_t k = 0; k < nreq; k++) {
                Request &r = *batch[k];
                if (!r.pairs) continue;
                const uint64_t extra = r.count * 8 - (r.known_pairs == UINT64_MAX ? 0 : r.count * 8);
                if (result_used + extra <= o.result_bytes) {
                    result_used += extra;
                    r.reserved += extra;
                    emit[k] = true;
                    np += r.count;
                }
            }
        }
        std::vector<SegDev> tab = prep.segs_host;
        for (uint32_t k = 0; k < nreq; k++)
            if (!emit[k]) {  // emits no pair in the fill pass (FP64 kernel or filter)
                tab[k].eps2 = -1.0;
                tab[k].t_in = tab[k].t_out = INFINITY;
            }
        upload_segs(c, tab, off);
        std::vector<int64_t> rowptr(nq + 1);
        std::vector<int64_t> colidx(np);
        csr_fill(c, prep, c.segs.as<SegDev>(), opt, false, np, rowptr.data(), colidx.data(), np);
        for (uint32_t k = 0; k < nreq; k++) {
            Request &r = *batch[k];
            if (!r.pairs) {
                ready.push_back(batch[k]);
                continue;
            }
            if (!emit[k]) {
                again.push_back(batch[k]);
                continue;
            }
            const int64_t base = rowptr[off[k]];
            r.rowptr.resize(r.nq + 1);
            for (uint64_t i = 0; i <= r.nq; i++) r.rowptr[i] = rowptr[off[k] + i] - base;
            r.colidx.assign(colidx.begin() + base, colidx.begin() + base + (int64_t)r.count);
            r.pts.clear();
            r.pts.shrink_to_fit();
            ready.push_back(batch[k]);
        }
    }
    for (ReqPtr &r : ready) {
        r->pts.clear();
        r->pts.shrink_to_fit();
    }
    if (!again.empty()) {
        // counted, but their pairs did not fit the result budget: they wait
        // at the head of the queue with their exact size
        std::lock_guard<std::mutex> lk(mu);
        for (auto it = again.rbegin(); it != again.rend(); ++it) {
            Request &r = **it;
            result_used -= r.reserved;
            r.reserved = 0;
            r.known_pairs = r.count;
            pending.push_front(*it);
        }
        cv_work.notify_all();
    }
    finish(ready, HJ_OK, "");
}

void hj_server::worker_main(int /*wid*/) {
    Ctx c;
    if (ctx_init(c, device) != HJ_OK) {
        ctx_release(c);
        std::lock_guard<std::mutex> lk(mu);
        err = "a worker could not create its HIP stream";
        workers_failed++;
        cv_init.notify_all();
        return;
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        workers_up++;
        cv_init.notify_all();
    }
    PinnedBuf stage;
    for (;;) {
        std::vector<ReqPtr> batch;
        uint64_t bid = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return (!pending.empty() && take_batch(batch)) || stop; });
            if (batch.empty()) {
                if (!stop) continue;
                // stopping: what cannot be admitted any more fails
                const double t = now_ms();
                for (ReqPtr &r : pending) {
                    r->status = HJ_ERR_CAPACITY;
                    r->err = "the server stopped before the request fitted its result budget";
                    r->done = true;
                    r->t_done = t;
                    failed++;
                }
                pending.clear();
                cv_done.notify_all();
                break;
            }
            bid = ++batches;
            batch_reqs += batch.size();
        }
        std::string msg;
        const int rc = guarded_call(msg, [&]() -> int {
            run_batch(c, stage, batch, bid);
            return HJ_OK;
        });
        if (rc != HJ_OK) {
            {
                std::lock_guard<std::mutex> lk(mu);
                for (ReqPtr &r : batch) {
                    result_used -= r->reserved;
                    r->reserved = 0;
                }
            }
            std::vector<ReqPtr> undone;
            for (ReqPtr &r : batch)
                if (!r->done) undone.push_back(r);
            finish(undone, rc, msg);
            cv_work.notify_all();
        }
    }
    ctx_release(c);
}

extern "C" {

HJ_API int hj_index_create(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d,
                           double eps_max, const hj_opts *opts, hj_index **out) {
    return index_create_common(ctx, pts, n, d, eps_max, opts, out, true);
}

HJ_API int hj_index_create_device(hj_ctx *ctx, const double *pts_dev, uint64_t n, uint32_t d,
                                  double eps_max, const hj_opts *opts, hj_index **out) {
    return index_create_common(ctx, pts_dev, n, d, eps_max, opts, out, false);
}

HJ_API int hj_index_destroy(hj_index *idx) {
    if (!idx) return HJ_OK;
    (void)hipSetDevice(idx->ix.device);
    (void)hipDeviceSynchronize();
    idx->ix.release();
    delete idx;
    return HJ_OK;
}

HJ_API int hj_index_info(const hj_index *idx, uint64_t *n, uint32_t *d, double *eps_max,
                         int32_t *key_dims, int32_t *key_bits, int32_t *active_dims) {
    if (!idx) return HJ_ERR_INVALID;
    const Index &ix = idx->ix;
    if (n) *n = ix.n;
    if (d) *d = (uint32_t)ix.d;
    if (eps_max) *eps_max = ix.eps_max;
    if (key_dims) *key_dims = ix.grid.key_dims;
    if (key_bits) *key_bits = ix.grid.bits;
    if (active_dims) *active_dims = ix.grid.active;
    return HJ_OK;
}

HJ_API int hj_index_query_count(hj_ctx *ctx, const hj_index *idx, const double *Q,
                                const uint64_t *q_off, uint32_t nreq, const double *eps,
                                const hj_opts *opts, uint64_t *counts, hj_stats *stats) {
    return query_count_common(ctx, idx, Q, q_off, nreq, eps, opts, counts, stats, true);
}

HJ_API int hj_index_query_count_device(hj_ctx *ctx, const hj_index *idx, const double *Q_dev,
                                       const uint64_t *q_off, uint32_t nreq, const double *eps,
                                       const hj_opts *opts, uint64_t *counts, hj_stats *stats) {
    return query_count_common(ctx, idx, Q_dev, q_off, nreq, eps, opts, counts, stats, false);
}

HJ_API int hj_index_query_csr(hj_ctx *ctx, const hj_index *idx, const double *Q,
                              const uint64_t *q_off, uint32_t nreq, const double *eps,
                              const hj_opts *opts, int64_t *rowptr, int64_t *colidx,
                              uint64_t cap, uint64_t *total, hj_stats *stats) {
    return guarded(ctx, [&]() -> int {
        const double t0 = now_ms();
        if (stats) std::memset(stats, 0, sizeof(*stats));
        Opts o = read_opts(opts);
        check_batch(idx, q_off, nreq, eps);
        if (!rowptr || !total) throw InvalidArg{"rowptr and total must not be NULL"};
        *total = 0;
        if (nreq == 0) {
            rowptr[0] = 0;
            return HJ_OK;
        }
        if (q_off[nreq] && !Q) throw InvalidArg{"null query points"};
        Ctx &c = ctx->c;
        if (c.device != idx->ix.device) throw InvalidArg{"index and context on different devices"};
        SideSpec Qs = query_side(c, Q, q_off, nreq, idx->ix.d, true);
        JoinOut R;
        JoinPrep prep;
        run_index_query(c, idx->ix, Qs, std::vector<double>(eps, eps + nreq), o, R, prep, stats,
                        false);
        const uint64_t nk = csr_fill(c, prep, prep.segs, o, false, R.count, rowptr, colidx, cap);
        *total = nk;
        if (stats) stats->total_ms = now_ms() - t0;
        if (nk > cap && colidx) {
            ctx->c.err = "colidx too small";
            return HJ_ERR_CAPACITY;
        }
        return HJ_OK;
    });
}

HJ_API void hj_server_opts_default(hj_server_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->workers = 2;
    o->kernel = -1;
    o->max_batch_requests = 4096;
    o->max_batch_points = 1ull << 18;
    o->result_bytes = 1ull << 30;
    o->order = HJ_ORDER_ROW;
}

HJ_API int hj_server_create(const hj_index *idx, int device, const hj_server_opts *opts,
                            hj_server **out) {
    if (!idx || !out) return HJ_ERR_INVALID;
    *out = nullptr;
    hj_server_opts o;
    hj_server_opts_default(&o);
    if (opts) o = *opts;
    if (o.workers < 1 || o.workers > 8) return HJ_ERR_INVALID;
    if (o.max_batch_requests == 0 || o.max_batch_requests > INDEX_MAX_REQUESTS)
        return HJ_ERR_INVALID;
    if (o.max_batch_points == 0 || o.result_bytes < 8) return HJ_ERR_INVALID;
    if (o.order != HJ_ORDER_ROW && o.order != HJ_ORDER_HILBERT) return HJ_ERR_INVALID;
    if (device != idx->ix.device) return HJ_ERR_INVALID;
    int kernel = idx->ix.kernel;
    if (o.kernel == HJ_KERNEL_FP64)
        kernel = HJ_KERNEL_FP64;
    else if (o.kernel >= 0 && o.kernel != idx->ix.kernel)
        return HJ_ERR_INVALID;  // the index holds no operands for that kernel
    hj_server *s = new (std::nothrow) hj_server();
    if (!s) return HJ_ERR_NOMEM;
    s->ix = &idx->ix;
    s->device = device;
    s->o = o;
    s->kernel = kernel;
    auto shut_down = [s] {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (auto &t : s->workers) t.join();
        delete s;
    };
    try {
        for (int w = 0; w < o.workers; w++) s->workers.emplace_back(&hj_server::worker_main, s, w);
    } catch (...) {
        shut_down();
        return HJ_ERR_INTERNAL;
    }
    // every worker reports whether it could set up its context; a server
    // without any working worker would accept requests it never serves
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_init.wait(lk, [&] { return s->workers_up + s->workers_failed == o.workers; });
        if (s->workers_up == 0) {
            lk.unlock();
            shut_down();
            return HJ_ERR_HIP;
        }
    }
    *out = s;
    return HJ_OK;
}

HJ_API int hj_server_submit(hj_server *srv, const double *q, uint64_t nq, double eps,
                            int want_pairs, uint64_t *ticket) {
    if (!srv || !ticket || (nq && !q)) return HJ_ERR_INVALID;
    if (!(eps > 0) || !(eps <= srv->ix->eps_max)) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->err = "request eps must be in (0, eps_max of the index]";
        return HJ_ERR_INVALID;
    }
    if (nq > 0xFFFFFF00ull) return HJ_ERR_INVALID;
    ReqPtr r;
    try {
        r = std::make_shared<Request>();
        r->pts.assign(q, q + (size_t)nq * srv->ix->d);
    } catch (const std::bad_alloc &) {
        return HJ_ERR_NOMEM;
    }
    r->nq = nq;
    r->eps = eps;
    r->pairs = want_pairs != 0;
    r->t_submit = now_ms();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) return HJ_ERR_INVALID;
        r->ticket = srv->next_ticket++;
        srv->table[r->ticket] = r;
        srv->pending.push_back(r);
        srv->submitted++;
        if (srv->first_submit < 0) srv->first_submit = r->t_submit;
        *ticket = r->ticket;
    }
    srv->cv_work.notify_one();
    return HJ_OK;
}

HJ_API int hj_server_wait(hj_server *srv, uint64_t ticket, double timeout_ms, hj_result *res) {
    if (!srv || !res) return HJ_ERR_INVALID;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->table.find(ticket);
    if (it == srv->table.end()) {
        srv->err = "unknown ticket";
        return HJ_ERR_INVALID;
    }
    ReqPtr r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) {
        srv->cv_done.wait(lk, ready);
    } else if (!srv->cv_done.wait_for(
                   lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
        return HJ_ERR_TIMEOUT;
    }
    std::memset(res, 0, sizeof(*res));
    res->ticket = r->ticket;
    res->status = r->status;
    res->pairs = r->pairs && r->status == HJ_OK ? 1 : 0;
    res->nq = r->nq;
    res->count = r->count;
    res->rowptr = res->pairs ? r->rowptr.data() : nullptr;
    res->colidx = res->pairs ? r->colidx.data() : nullptr;
    res->queue_ms = r->t_start > 0 ? r->t_start - r->t_submit : 0;
    res->run_ms = r->t_start > 0 ? r->t_done - r->t_start : 0;
    res->latency_ms = r->t_done - r->t_submit;
    res->batch = r->batch;
    res->batch_requests = r->batch_requests;
    if (r->status != HJ_OK) srv->err = r->err;
    return HJ_OK;
}

HJ_API int hj_server_release(hj_server *srv, uint64_t ticket) {
    if (!srv) return HJ_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->table.find(ticket);
        if (it == srv->table.end() || !it->second->done) return HJ_ERR_INVALID;
        srv->result_used -= it->second->reserved;
        srv->table.erase(it);
    }
    srv->cv_work.notify_all();  // memory came free: waiting requests may fit
    return HJ_OK;
}

HJ_API int hj_server_stats_get(hj_server *srv, hj_server_stats *st) {
    if (!srv || !st) return HJ_ERR_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = srv->submitted;
    st->completed = srv->completed;
    st->failed = srv->failed;
    st->batches = srv->batches;
    st->pending = srv->pending.size();
    st->result_bytes_used = srv->result_used;
    st->mean_batch_requests = srv->batches ? (double)srv->batch_reqs / srv->batches : 0.0;
    if (!srv->latencies.empty()) {
        std::vector<double> v = srv->latencies;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            // nearest-rank percentile
            size_t k = (size_t)std::ceil(p * v.size());
            return v[std::min(v.size() - 1, k > 0 ? k - 1 : 0)];
        };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        st->mean_ms = sum / v.size();
        st->max_ms = v.back();
        const double span = srv->last_done - srv->first_submit;
        st->requests_per_s = span > 0 ? srv->completed / (span * 1e-3) : 0.0;
    }
    return HJ_OK;
}

HJ_API int hj_server_stats_reset(hj_server *srv) {
    if (!srv) return HJ_ERR_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->latencies.clear();
    srv->completed = srv->failed = srv->batches = srv->batch_reqs = 0;
    srv->submitted = srv->pending.size();
    srv->first_submit = -1;
    srv->last_done = 0;
    return HJ_OK;
}

HJ_API int hj_server_destroy(hj_server *srv) {
    if (!srv) return HJ_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &t : srv->workers) t.join();
    delete srv;
    return HJ_OK;
}

HJ_API const char *hj_server_last_error(const hj_server *srv) {
    if (!srv) return "null server";
    // a copy taken under the lock (workers write the message concurrently);
    // valid until this thread's next call
    static thread_local std::string copy;
    hj_server *m = const_cast<hj_server *>(srv);
    std::lock_guard<std::mutex> lk(m->mu);
    copy = m->err;
    return copy.c_str();
}

}  // extern "C"
