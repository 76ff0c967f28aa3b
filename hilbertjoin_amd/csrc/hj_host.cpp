// hj_host.cpp -- host-side data services of the library: the point reader
// (replacing read_file, dataIo.cpp:113-165) and the seeded synthetic-data
// generators (replacing the MKL generator random_init_unif, dataIo.cpp:4-33).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hilbertjoin_amd.h"

namespace {

constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// element k of stream s, uniform in [0, 1)
inline double uniform(uint64_t s, uint64_t k) {
    return (double)(mix64(mix64(s) + (k + 1) * GOLDEN) >> 11) * 0x1.0p-53;
}

inline uint64_t substream(uint64_t s, uint64_t tag) {
    return mix64(s ^ (tag * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
}

// Box-Muller from uniforms 2k and 2k+1 of stream s
inline double normal(uint64_t s, uint64_t k) {
    double u1 = 1.0 - uniform(s, 2 * k);  // (0, 1]
    double u2 = uniform(s, 2 * k + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// Gamma(a) by Marsaglia-Tsang; a < 1 via Gamma(a+1) * U^(1/a).  *ctr advances
// through stream s.
double gamma_sample(double a, uint64_t s, uint64_t *ctr) {
    double boost = 1.0;
    if (a < 1.0) {
        boost = std::pow(uniform(s, (*ctr)++), 1.0 / a);
        a += 1.0;
    }
    const double dd = a - 1.0 / 3.0, cc = 1.0 / std::sqrt(9.0 * dd);
    for (;;) {
        double x = normal(s, (*ctr)++);
        double v = 1.0 + cc * x;
        if (v <= 0) continue;
        v = v * v * v;
        double u = uniform(s, (*ctr)++);
        if (u < 1.0 - 0.0331 * x * x * x * x) return dd * v * boost;
        if (std::log(u) < 0.5 * x * x + dd * (1.0 - v + std::log(v))) return dd * v * boost;
    }
}

template <class F> void parallel_rows(uint64_t n, F &&f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        uint64_t a = n * t / nt, b = n * (t + 1) / nt;
        th.emplace_back([=, &f]() {
            for (uint64_t i = a; i < b; i++) f(i);
        });
    }
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

HJ_API void hj_gen_uniform(uint64_t seed, uint64_t n, uint32_t d, double *out) {
    if (!out) return;
    parallel_rows(n, [&](uint64_t i) {
        for (uint32_t j = 0; j < d; j++) out[i * d + j] = uniform(seed, i * d + j);
    });
}

HJ_API void hj_gen_mixture(uint64_t seed, uint64_t n, uint32_t d, uint32_t ncomp, double sigma,
                           double alpha, double *out) {
    if (!out || ncomp == 0) return;
    const uint64_t s_centre = substream(seed, 1), s_weight = substream(seed, 2),
                   s_pick = substream(seed, 3), s_coord = substream(seed, 4);
    std::vector<double> centre((size_t)ncomp * d), cdf(ncomp);
    for (uint64_t k = 0; k < (uint64_t)ncomp * d; k++) centre[k] = uniform(s_centre, k);
    uint64_t ctr = 0;
    double tot = 0;
    for (uint32_t c = 0; c < ncomp; c++) {
        tot += gamma_sample(alpha, s_weight, &ctr);
        cdf[c] = tot;
    }
    for (uint32_t c = 0; c < ncomp; c++) cdf[c] /= tot;
    const double top = std::nextafter(1.0, 0.0);
    parallel_rows(n, [&](uint64_t i) {
        double u = uniform(s_pick, i);
        uint32_t c = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (c >= ncomp) c = ncomp - 1;
        for (uint32_t j = 0; j < d; j++) {
            double x = centre[(size_t)c * d + j] + sigma * normal(s_coord, i * d + j);
            out[i * d + j] = std::min(std::max(x, 0.0), top);
        }
    });
}

HJ_API int hj_read_points(const char *path, uint64_t n, uint32_t d, double *out) {
    if (!path || (!out && n && d)) return HJ_ERR_INVALID;
    size_t len = std::strlen(path);
    bool bin = len >= 4 && std::strcmp(path + len - 4, ".bin") == 0;
    if (bin) {
        FILE *fp = std::fopen(path, "rb");
        if (!fp) return HJ_ERR_INVALID;
        size_t want = (size_t)n * d;
        size_t got = want ? std::fread(out, sizeof(double), want, fp) : 0;
        std::fclose(fp);
        return got == want ? HJ_OK : HJ_ERR_INVALID;
    }
    std::ifstream in(path);
    if (!in) return HJ_ERR_INVALID;
    std::string line;
    uint64_t i = 0;
    while (i < n && std::getline(in, line)) {
        if (!line.empty() && line[0] == '%') continue;
        // tokens as strtok(line, ",") gives them on the line fgets returns,
        // newline included (read_file dataIo.cpp:146-158): runs of ','
        // separate, there are no empty tokens, and the newline ends the last
        // token or, after a trailing ',' (or on an empty line), is one; atof
        // semantics (0.0 when not a number)
        std::string buf = line + "\n";
        char *save = nullptr;
        uint32_t j = 0;
        for (char *tok = strtok_r(&buf[0], ",", &save); tok && j < d;
             tok = strtok_r(nullptr, ",", &save))
            out[i * d + j++] = std::atof(tok);
        for (; j < d; j++) out[i * d + j] = 0.0;
        i++;
    }
    return i == n ? HJ_OK : HJ_ERR_INVALID;
}

}  // extern "C"
