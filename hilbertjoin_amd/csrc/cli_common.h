// cli_common.h -- shared pieces of the two command-line tools.
//
// The tools keep the reference's flags (util/arguments.cpp:24,100) and its
// ';'-separated result line (main.cpp:122, mainJoin.cpp:139).  Extra long
// options select the MI355X kernel and traversal; they do not clash with the
// reference's single-letter flags.
#pragma once

#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hilbertjoin_amd.h"

namespace hjcli {

struct Args {
    size_t n = 0, m = 0, d = 0;
    double eps = 0.0;
    std::string fileA, fileB;
    int actdim = 3;   // -s, default 3 as in the reference (main.cpp:44, mainJoin.cpp:46);
                      // --auto-dims: chosen from the data
    int kernel = HJ_KERNEL_FP64;
    int order = HJ_ORDER_ROW;
    int kblock = 4;
    int device = 0;
    bool reorder = true;  // dimension statistics + reorder pre-pass, as the reference always
                          // runs it (main.cpp:83-87, mainJoin.cpp:97-100); --no-reorder skips it
};

inline void usage(const char *prog, bool join) {
    std::fprintf(stderr, "There are obligatory parameters.\n");
    std::fprintf(stderr, "Usage: %s\n", prog);
    std::fprintf(stderr, "Obligatory parameters:\n");
    if (join)
        std::fprintf(stderr,
                     "n (number of objects in set A)\nm (number of objects in set B)\n"
                     "e (epsilon)\nd (dimensionality)\n");
    else
        std::fprintf(stderr, "n (number of objects)\ne (epsilon)\nd (dimensionality)\n");
    std::fprintf(stderr, "Optional parameters:\n");
    std::fprintf(stderr, "s number of active dimensions, 0..6 (default 3)\n");
    std::fprintf(stderr,
                 "f (filename) if there is no filename we use random generated data [0.0, 1.0)\n");
    if (join)
        std::fprintf(stderr, "g (filename set B) if there is no filename we use random "
                             "generated data [0.0, 1.0)\n");
    std::fprintf(stderr, "--kernel=fp64|f16|bf16  --order=hilbert|row  --kblock=4|8|16  "
                         "--device=N  --no-reorder  --auto-dims\n");
}

inline bool parse(int argc, char **argv, bool join, Args &a) {
    if (argc < 5) {
        usage(argv[0], join);
        return false;
    }
    static struct option longopts[] = {{"kernel", required_argument, nullptr, 1000},
                                       {"order", required_argument, nullptr, 1001},
                                       {"kblock", required_argument, nullptr, 1002},
                                       {"device", required_argument, nullptr, 1003},
                                       {"reorder", no_argument, nullptr, 1004},
                                       {"no-reorder", no_argument, nullptr, 1005},
                                       {"auto-dims", no_argument, nullptr, 1006},
                                       {nullptr, 0, nullptr, 0}};
    const char *shortopts = join ? "n:m:e:d:t:f:g:k:a:s:" : "n:e:d:t:f:k:s:";
    int c;
    while ((c = getopt_long(argc, argv, shortopts, longopts, nullptr)) != -1) {
        switch (c) {
        case 'n': a.n = (size_t)std::atol(optarg); break;
        case 'm': a.m = (size_t)std::atol(optarg); break;
        case 'd': a.d = (size_t)std::atoi(optarg); break;
        case 'e': a.eps = std::atof(optarg); break;
        case 's':
            a.actdim = std::atoi(optarg);
            if (a.actdim < 0) a.actdim = 7;  // rejected below, as in the reference
            break;
        case 'f': a.fileA = optarg; break;
        case 'g': a.fileB = optarg; break;
        case 't':
        case 'k':
        case 'a': break;  // accepted and ignored, as in the reference
        case 1000:
            if (!std::strcmp(optarg, "fp64")) a.kernel = HJ_KERNEL_FP64;
            else if (!std::strcmp(optarg, "f16")) a.kernel = HJ_KERNEL_MFMA_F16;
            else if (!std::strcmp(optarg, "bf16")) a.kernel = HJ_KERNEL_MFMA_BF16;
            else {
                std::fprintf(stderr, "unknown kernel '%s'\n", optarg);
                return false;
            }
            break;
        case 1001:
            if (!std::strcmp(optarg, "hilbert")) a.order = HJ_ORDER_HILBERT;
            else if (!std::strcmp(optarg, "row")) a.order = HJ_ORDER_ROW;
            else {
                std::fprintf(stderr, "unknown order '%s'\n", optarg);
                return false;
            }
            break;
        case 1002: a.kblock = std::atoi(optarg); break;
        case 1003: a.device = std::atoi(optarg); break;
        case 1004: a.reorder = true; break;
        case 1005: a.reorder = false; break;
        case 1006: a.actdim = -1; break;
        default: return false;  // getopt already printed the unknown option
        }
    }
    if (a.actdim < -1 || a.actdim > 6) {
        std::fprintf(stderr, "parameter active dimensions is typical in the range of [0-5]\n");
        return false;
    }
    if (a.kblock != 4 && a.kblock != 8 && a.kblock != 16) {
        std::fprintf(stderr, "KBLOCK must be 4, 8 or 16\n");
        return false;
    }
    return true;
}

inline bool load(const std::string &file, size_t n, size_t d, uint64_t seed,
                 std::vector<double> &out) {
    out.assign(n * d, 0.0);
    if (file.empty()) {
        hj_gen_uniform(seed, n, (uint32_t)d, out.data());
        return true;
    }
    if (hj_read_points(file.c_str(), n, (uint32_t)d, out.data()) != HJ_OK) {
        std::fprintf(stderr, "File '%s' does not exist or is too short!\n", file.c_str());
        return false;
    }
    return true;
}

// The reference's pre-pass (main.cpp:83-91): dimension statistics of the first
// set, then the same column permutation applied to every set.  Returns false
// on failure.
inline bool reorder_sets(hj_ctx *ctx, const Args &a, std::vector<double> *sets[], int nsets,
                         const size_t *ns) {
    std::vector<int32_t> rd(a.d);
    if (hj_dim_statistics(ctx, sets[0]->data(), ns[0], (uint32_t)a.d, a.eps, rd.data(),
                          nullptr) != HJ_OK)
        return false;
    for (int k = 0; k < nsets; k++)
        if (hj_reorder_dimensions(ctx, sets[k]->data(), ns[k], (uint32_t)a.d, rd.data()) !=
            HJ_OK)
            return false;
    return true;
}

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace hjcli
