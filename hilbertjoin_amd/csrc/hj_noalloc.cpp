// hj_noalloc.cpp -- the self-join count that allocates nothing.
//
// hj_self_join_count_device_noalloc runs the same pipeline and kernels as
// hj_self_join_count_device, but every internal device buffer (keys, row
// permutations, tile layouts, filter operands, work list, band queue,
// counters, sort histograms) comes from `scratch`, a file-scope device array of
// HJ_NOALLOC_SCRATCH_BYTES, and every host staging buffer from a file-scope
// host array.  It calls no hipMalloc, hipHostMalloc, hipFree or other
// allocating HIP routine: for callers that cannot allocate (code under graph
// capture, tight serving loops, harnesses that forbid it).
//
// Each GPU holds its own copy of `scratch` from the time the library's code
// object is loaded on it until the process ends.  Its device address is looked
// up once per device (hipGetSymbolAddress, with that device current) and kept;
// host code never passes the array's name itself to a kernel or a copy.  Every
// call on one GPU shares that GPU's array (and that GPU's slice of the host
// staging array), so calls that overlap on one GPU -- from different threads
// or on different contexts of that device -- race: the caller serializes them.  The
// contents left by the previous call are never relied on: counters and flags
// are cleared by kernels at the start of each use.
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "hj_join.h"

using namespace hj;

namespace {

constexpr size_t SCRATCH_BYTES = HJ_NOALLOC_SCRATCH_BYTES;
constexpr size_t HOST_STAGE_BYTES = (size_t)256 << 10;  // per device
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
alignas(256) unsigned char host_stage[MAX_DEVICES][HOST_STAGE_BYTES];

std::mutex scratch_mu;
void *scratch_ptr[MAX_DEVICES] = {};

// Device address of this device's copy of `scratch` (the current device is
// `device`).
void *scratch_on(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw InvalidArg{"device id out of range"};
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[device]) {
        void *p = nullptr;
        HJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw HipError{hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)"};
        scratch_ptr[device] = p;
    }
    return scratch_ptr[device];
}

// Device bytes the join needs before its work list is known (the work list
// itself comes on top): sort keys and rows, the tile layout of the kernel, tile
// bounds, the candidate intervals of every tile for the largest stripe
// enumeration, the band queue, the counters.
size_t fixed_need(uint64_t n, uint32_t d, const Opts &o) {
    const size_t nt = (n + 63) / 64;
    size_t b = 24 * (size_t)n + 2 * (2 * (size_t)d + 1) * 8;
    b += nt * (4 + 8 + 16);
    if (o.kernel == HJ_KERNEL_FP64) {
        b += nt * 64 * (size_t)dpad_for((int)d, o.kblock) * 8;
    } else {
        const FilterParams fp = filter_layout((int)d, o.kernel == HJ_KERNEL_MFMA_BF16, o.split);
        b += 2 * nt * 2 * (size_t)fp.steps * 64 * 16 + ((size_t)8 << 20);
        if (o.frames && frames_possible(fp, (int)d))  // anchors and tile radii
            b += nt * ((size_t)fp.steps * 16 * sizeof(float) + sizeof(float2)) + 8;
    }
    b += nt * 365 * 8 + nt * 20;
    b += (8 + SCHED_WORDS) * sizeof(uint64_t);  // join and scheduling counters
    return b;
}

}  // namespace

extern "C" {

HJ_API int hj_self_join_count_device_noalloc(hj_ctx *ctx, const double *pts_dev, uint64_t n,
                                             uint32_t d, double eps, const hj_opts *opts,
                                             uint64_t *count, hj_stats *stats) {
    if (!ctx) return HJ_ERR_INVALID;
    return guarded_call(ctx->c.err, [&]() -> int {
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (!count) throw InvalidArg{"count must not be NULL"};
        Opts o = read_opts(opts);
        if (o.tune) throw InvalidArg{"the no-allocation entry point does not autotune"};
        check_sizes(n, d);
        if (n && !pts_dev) throw InvalidArg{"null point array"};
        if (!(eps > 0) || !std::isfinite(eps)) throw InvalidArg{"eps must be finite and > 0"};
        *count = 0;
        if (n < 2) return HJ_OK;
        const size_t need = fixed_need(n, d, o);
        if (need > SCRATCH_BYTES)
            throw CapacityError{"n = " + std::to_string(n) + ", d = " + std::to_string(d) +
                                " needs at least " + std::to_string(need >> 20) +
                                " MiB of scratch; the no-allocation entry point has " +
                                std::to_string(SCRATCH_BYTES >> 20) +
                                " MiB (use hj_self_join_count_device)"};
        Ctx &base = ctx->c;
        HJ_HIP(hipSetDevice(base.device));
        ScratchArena arena;
        arena.dev = static_cast<char *>(scratch_on(base.device));
        arena.dev_cap = SCRATCH_BYTES;
        arena.host = reinterpret_cast<char *>(host_stage[base.device]);
        arena.host_cap = HOST_STAGE_BYTES;
        // a context of its own whose buffers are all borrowed from the arena;
        // the stream and the timing events are the caller's context's
        Ctx c;
        c.device = base.device;
        c.num_cus = base.num_cus;
        c.stream = base.stream;
        c.stream2 = base.stream2;
        c.fork = base.fork;
        c.join = base.join;
        for (int k = 0; k < 8; k++) c.ev[k] = base.ev[k];
        struct ArenaScope {
            explicit ArenaScope(ScratchArena *a) { t_arena = a; }
            ~ArenaScope() { t_arena = nullptr; }
        } scope(&arena);
        const auto t0 = std::chrono::steady_clock::now();
        Problem P;
        P.d = (int)d;
        P.self = true;
        P.Q.pts = pts_dev;
        P.Q.n = n;
        P.Q.off = {0, n};
        P.eps = {eps};
        JoinOut R;
        JoinPrep prep;
        try {
            run_problem(c, P, o, R, prep, stats, false);
        } catch (...) {
            // nothing enqueued may outlive the call: the next one reuses the
            // same scratch memory
            (void)hipStreamSynchronize(c.stream2);
            (void)hipStreamSynchronize(c.stream);
            throw;
        }
        *count = R.count;
        if (stats)
            stats->total_ms = std::chrono::duration<double, std::milli>(
                                  std::chrono::steady_clock::now() - t0)
                                  .count();
        return HJ_OK;
    });
}

}  // extern "C"
