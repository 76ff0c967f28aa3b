// hilbertJoinCardinality -- A x B join count on one MI355X (mainJoin.cpp).
//
// Same flags and result line as the reference:
//   N;M;D;JPPP;THREADS;EPSILON;STRIPES;KBLOCK;TIME;ALGTIME;SORTTIME;INDEXTIME;REORDERTIME;COUNTS;LOADPERCENT;WH
// Differences: set B is read from -g (the reference reads A's file again,
// mainJoin.cpp:77), and -s sets the active dimensions of the join (the
// reference hard-codes 2, egojoin.cpp:847; default 3 as mainJoin.cpp:46).
// The reorder pre-pass runs by default as in mainJoin.cpp:97-100
// (--no-reorder skips it).  LOADPERCENT uses the reference's formula
// (egojoin.cpp:865) over 64 x 64 tiles instead of 8 x 8 blocks.
#include "cli_common.h"

int main(int argc, char **argv) {
    hjcli::Args a;
    a.n = 200000;
    a.m = 200000;
    a.d = 64;
    a.eps = 0.034;
    if (!hjcli::parse(argc, argv, true, a)) return 1;
    std::vector<double> A, B;
    if (!hjcli::load(a.fileA, a.n, a.d, 1, A)) return 1;
    if (!hjcli::load(a.fileB, a.m, a.d, 2, B)) return 1;
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
        std::vector<double> *sets[2] = {&A, &B};
        const size_t ns[2] = {a.n, a.m};
        if (!hjcli::reorder_sets(ctx, a, sets, 2, ns)) {
            std::fprintf(stderr, "reorder failed: %s\n", hj_last_error(ctx));
            hj_ctx_destroy(ctx);
            return 1;
        }
        reorder_time = hjcli::now_s() - tr;
    }
    hj_stats st;
    uint64_t count = 0;
    double t0 = hjcli::now_s();
    int rc = hj_join_count(ctx, A.data(), a.n, B.data(), a.m, (uint32_t)a.d, a.eps, &o, &count,
                           &st);
    double algtime = hjcli::now_s() - t0;
    if (rc != HJ_OK) {
        std::fprintf(stderr, "join failed: %s\n", hj_last_error(ctx));
        hj_ctx_destroy(ctx);
        return 1;
    }
    // STRIPES as the reference derives it from the active dimensions (main.cpp:55)
    int stripes = ((int)std::pow(3, st.active_dims) + 1) / 2;
    double sort_time = (st.grid_ms + st.sort_ms) * 1e-3;
    double index_time = st.index_ms * 1e-3;
    double wh = 0.0;
    double jppp = count == 0 ? 0.0 : (double)count / (double)a.n;
    double blocks = (double)(a.n / 8);
    double loadpercent =
        blocks > 1 ? ((double)st.candidate_pairs / 64.0) / blocks / (blocks - 1) * 200.0 : 0.0;
    std::printf("%zu;%zu;%zu;%f;%d;%2.14f;%d;%d;%f;%f;%f;%f;%f;%ld;%f;%f\n", a.n, a.m, a.d, jppp,
                cus, a.eps, stripes, a.kblock, algtime + reorder_time, algtime - sort_time,
                sort_time, index_time, reorder_time, (long)count, loadpercent, wh);
    hj_ctx_destroy(ctx);
    return 0;
}
