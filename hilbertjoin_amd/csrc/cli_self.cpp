// hilbertSelfJoinCardinality -- self-join count on one MI355X (main.cpp).
//
// Same flags and result line as the reference:
//   N;D;JPPP;THREADS;EPSILON;STRIPES;KBLOCK;TIME;ALGTIME;SORTTIME;INDEXTIME;REORDERTIME;COUNTS;LOADPERCENT;WH
// THREADS reports the GPU's compute units; the dimension reorder pre-pass of
// the reference (main.cpp:83-91) runs by default (--no-reorder skips it and
// REORDERTIME is then 0); STRIPES follows -s (default 3 -> 14 stripes, as
// main.cpp:44,55); LOADPERCENT is the candidate fraction of the reference's
// formula (egojoin.cpp:797) counted over 64 x 64 tiles instead of 8 x 8
// blocks; WH is 0 as in the reference (main.cpp:49).
#include "cli_common.h"

int main(int argc, char **argv) {
    hjcli::Args a;
    a.n = 4000000;
    a.d = 41;
    a.eps = 0.034;
    if (!hjcli::parse(argc, argv, false, a)) return 1;
    std::vector<double> pts;
    if (!hjcli::load(a.fileA, a.n, a.d, 1, pts)) return 1;
    hj_ctx *ctx = nullptr;
    if (hj_ctx_create(&ctx, a.device) != HJ_OK) {
        std::fprintf(stderr, "no usable HIP device %d\n", a.device);
        return 1;
    }
    int cus = 0;
    hj_device_info(ctx, &cus, nullptr, 0);
    hj_opts o;
    hj_opts_default(&o);
    o.kernel = a.kernel;
    o.order = a.order;
    o.kblock = a.kblock;
    o.active_dims = a.actdim;
    double reorder_time = 0.0;
    if (a.reorder) {
        double tr = hjcli::now_s();
        std::vector<double> *sets[1] = {&pts};
        const size_t ns[1] = {a.n};
        if (!hjcli::reorder_sets(ctx, a, sets, 1, ns)) {
            std::fprintf(stderr, "reorder failed: %s\n", hj_last_error(ctx));
            hj_ctx_destroy(ctx);
            return 1;
        }
        reorder_time = hjcli::now_s() - tr;
    }
    hj_stats st;
    uint64_t count = 0;
    double t0 = hjcli::now_s();
    int rc = hj_self_join_count(ctx, pts.data(), a.n, (uint32_t)a.d, a.eps, &o, &count, &st);
    double algtime = hjcli::now_s() - t0;
    if (rc != HJ_OK) {
        std::fprintf(stderr, "self-join failed: %s\n", hj_last_error(ctx));
        hj_ctx_destroy(ctx);
        return 1;
    }
    // STRIPES as the reference derives it from the active dimensions (main.cpp:55)
    int stripes = ((int)std::pow(3, st.active_dims) + 1) / 2;
    double sort_time = (st.grid_ms + st.sort_ms) * 1e-3;
    double index_time = st.index_ms * 1e-3;
    double wh = 0.0;
    double jppp = count == 0 ? 0.0 : (double)count / (double)a.n;
    std::printf("%zu;%zu;%f;%d;%2.14f;%d;%d;%f;%f;%f;%f;%f;%zu;%f;%f\n", a.n, a.d, jppp, cus, a.eps,
                stripes, a.kblock, algtime + reorder_time, algtime - sort_time, sort_time,
                index_time, reorder_time, (size_t)count, st.load_fraction, wh);
    hj_ctx_destroy(ctx);
    return 0;
}
