// hj_compat.cpp -- the reference's join entry points (egojoin.h:56,64) on top of
// the C ABI.  One process-wide context, created on first use.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "../../include/egojoin_compat.h"
#include "../../include/hilbertjoin_amd.h"

namespace {

hj_ctx *g_ctx = nullptr;
std::mutex g_mu;

[[noreturn]] void die(const char *what, const char *msg) {
    std::fprintf(stderr, "hilbertjoin_amd: %s: %s\n", what, msg ? msg : "");
    std::exit(1);
}

hj_ctx *context() {
    if (g_ctx) return g_ctx;
    const char *dev = std::getenv("HJ_DEVICE");
    int device = dev ? std::atoi(dev) : 0;
    if (hj_ctx_create(&g_ctx, device) != HJ_OK) die("context", "no usable HIP device");
    return g_ctx;
}

void env_opts(hj_opts *o) {
    hj_opts_default(o);
    const char *k = std::getenv("HJ_KERNEL");
    if (k) {
        if (!std::strcmp(k, "f16"))
            o->kernel = HJ_KERNEL_MFMA_F16;
        else if (!std::strcmp(k, "bf16"))
            o->kernel = HJ_KERNEL_MFMA_BF16;
        else if (std::strcmp(k, "fp64"))
            die("HJ_KERNEL", "expected fp64, f16 or bf16");
    }
}

// prepareStripes (egojoin.cpp:276-282): the active dimensions are the
// smallest a with 3^a >= 2 * stripes.
int active_for_stripes(int stripes) {
    int a = 0, p = 1;
    while (p < 2 * stripes) {
        p *= 3;
        a++;
    }
    return a > 6 ? 6 : a;
}

}  // namespace

void test_ego_loop3_macro(size_t n, size_t d, double epsilon, double *array,
                          size_t *countresult, int stripes, double *sortTime, double *indextime,
                          double *loadpercent) {
    std::lock_guard<std::mutex> lk(g_mu);
    hj_ctx *ctx = context();
    hj_opts o;
    env_opts(&o);
    o.active_dims = active_for_stripes(stripes);
    hj_stats st;
    uint64_t cnt = 0;
    int rc = hj_self_join_count(ctx, array, n, (uint32_t)d, epsilon, &o, &cnt, &st);
    if (rc != HJ_OK) die("test_ego_loop3_macro", hj_last_error(ctx));
    if (countresult) *countresult = (size_t)cnt;
    if (sortTime) *sortTime = (st.grid_ms + st.sort_ms) * 1e-3;
    if (indextime) *indextime = st.index_ms * 1e-3;
    if (loadpercent) *loadpercent = st.load_fraction;
}

void test_ego_loop3_noself(const size_t nA, const size_t nB, const int d, const double epsilon,
                           double *arrayA, double *arrayB, size_t *countresult,
                           const int activedims, double *sortTime, double *indextime,
                           double *loadpercent) {
    std::lock_guard<std::mutex> lk(g_mu);
    hj_ctx *ctx = context();
    hj_opts o;
    env_opts(&o);
    o.active_dims = activedims < 0 ? -1 : (activedims > 6 ? 6 : activedims);
    hj_stats st;
    uint64_t cnt = 0;
    int rc = hj_join_count(ctx, arrayA, nA, arrayB, nB, (uint32_t)d, epsilon, &o, &cnt, &st);
    if (rc != HJ_OK) die("test_ego_loop3_noself", hj_last_error(ctx));
    if (countresult) *countresult = (size_t)cnt;
    if (sortTime) *sortTime = (st.grid_ms + st.sort_ms) * 1e-3;
    if (indextime) *indextime = st.index_ms * 1e-3;
    if (loadpercent) {
        double b = (double)(nA / 8);
        *loadpercent = b > 1 ? ((double)st.candidate_pairs / 64.0) / b / (b - 1) * 200.0 : 0.0;
    }
}

void outputStatistics(int n, int d, double epsilon, double *array, int *reorder_dim) {
    std::lock_guard<std::mutex> lk(g_mu);
    hj_ctx *ctx = context();
    if (n < 0 || d < 1) die("outputStatistics", "bad n or d");
    int rc = hj_dim_statistics(ctx, array, (uint64_t)n, (uint32_t)d, epsilon, reorder_dim,
                               nullptr);
    if (rc != HJ_OK) die("outputStatistics", hj_last_error(ctx));
    // the reference's array has 8 padding entries, mapped to themselves
    for (int j = d; j < d + 8; j++) reorder_dim[j] = j;
}

void reorder_dimensions(int n, int d, double *array, int *reorder_dim) {
    std::lock_guard<std::mutex> lk(g_mu);
    hj_ctx *ctx = context();
    if (n < 0 || d < 1) die("reorder_dimensions", "bad n or d");
    int rc = hj_reorder_dimensions(ctx, array, (uint64_t)n, (uint32_t)d, reorder_dim);
    if (rc != HJ_OK) die("reorder_dimensions", hj_last_error(ctx));
}
