// Wavefront-level output queues staged in LDS, shared by the pair-appending
// kernels (the MFMA filter's band and pair queues, the FP64 pair kernel).
//
// A kernel appends (row | col << 32) entries to arrays in HBM whose cursors
// are counters in device memory.  On clustered data most tile pairs append
// something, and one atomic per entry (or per row or block) on one counter
// serialises.  Each wavefront instead gathers its entries in LDS and writes
// them out in coalesced runs with one atomic per run.  All queue state (fill
// counts, destinations, capacities, cursors) lives in LDS, not in registers:
// the kernels that use it run at their register limit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hj {

// Number of set bits of m in the lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// The 64-bit value v of lane `lane` (wave-uniform), in every lane.
__device__ __forceinline__ uint64_t read_lane_u64(uint64_t v, uint32_t lane) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

__device__ __forceinline__ uint64_t read_first_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// x, opaque to the optimiser: what is computed from it stays where it is used
// instead of being hoisted out of a persistent loop into a register held for
// the whole kernel.
__device__ __forceinline__ uint32_t opaque_v(uint32_t x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Orders the LDS accesses of one wavefront's lanes before and after the
// point (LDS executes one wavefront's instructions in issue order; this keeps
// the compiler from moving them across).
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Entries per wavefront and queue staged in LDS.
constexpr uint32_t QCAP = 256;

// The NQ queues of one workgroup of four wavefronts.
template <uint32_t NQ>
struct QueueLds {
    uint64_t slots[4][NQ][QCAP];
    uint32_t fill[4][NQ];
    uint64_t dst[NQ];     // uint64_t * destination array
    uint64_t cap[NQ];     // entries past cap are counted but not stored
    uint64_t cursor[NQ];  // unsigned long long * append counter
};

// Sets queue q up; called by one thread before the workgroup's barrier
// (every fill count is cleared by queue_clear_fills).
template <uint32_t NQ>
__device__ __forceinline__ void queue_setup(QueueLds<NQ> &qs, uint32_t q, uint64_t *dst,
                                            uint64_t cap, uint64_t *cursor) {
    qs.dst[q] = reinterpret_cast<uint64_t>(dst);
    qs.cap[q] = cap;
    qs.cursor[q] = reinterpret_cast<uint64_t>(cursor);
}

template <uint32_t NQ>
__device__ __forceinline__ void queue_clear_fills(QueueLds<NQ> &qs) {
    if (threadIdx.x < 4 * NQ) (&qs.fill[0][0])[threadIdx.x] = 0;
}

// Writes the f staged entries of queue Q of wavefront wv out.  Called by the
// whole wavefront.
template <uint32_t NQ, uint32_t Q>
__device__ __forceinline__ void queue_flush(QueueLds<NQ> &qs, uint32_t wv, uint32_t f) {
    const uint32_t lane = opaque_v(threadIdx.x & 63);
    const uint32_t z = opaque_v(0);
    wave_lds_fence();
    uint64_t *dst = reinterpret_cast<uint64_t *>(
        read_first_u64(*reinterpret_cast<volatile uint64_t *>(&qs.dst[Q + z])));
    const uint64_t cap = read_first_u64(*reinterpret_cast<volatile uint64_t *>(&qs.cap[Q + z]));
    uint64_t base = 0;
    if (lane == 0) {
        auto *cursor = reinterpret_cast<unsigned long long *>(
            *reinterpret_cast<volatile uint64_t *>(&qs.cursor[Q + z]));
        base = atomicAdd(cursor, (unsigned long long)f);
    }
    base = read_lane_u64(base, 0);
#pragma unroll
    for (uint32_t k = 0; k < QCAP / 64; k++) {
        const uint32_t i = lane + 64 * k;
        if (i < f && base + i < cap) dst[base + i] = qs.slots[wv][Q][i];
    }
    wave_lds_fence();
}

// Appends v of the lanes with sel to queue Q.  Called by the whole wavefront.
template <uint32_t NQ, uint32_t Q>
__device__ __forceinline__ void queue_push(QueueLds<NQ> &qs, bool sel, uint64_t v) {
    const uint64_t hits = __ballot(sel);
    if (!hits) return;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    volatile uint32_t *pf = &qs.fill[wv + opaque_v(0)][Q];
    uint32_t f = __builtin_amdgcn_readfirstlane(*pf);
    if (f > QCAP - 64) {
        queue_flush<NQ, Q>(qs, wv, f);
        f = 0;
    }
    if (sel) qs.slots[wv][Q][f + lanes_below(hits)] = v;
    *pf = f + (uint32_t)__popcll(hits);
}

// Writes what queue Q of this wavefront still holds.  Called by the whole
// wavefront before it exits.
template <uint32_t NQ, uint32_t Q>
__device__ __forceinline__ void queue_drain(QueueLds<NQ> &qs) {
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t f =
        __builtin_amdgcn_readfirstlane(*reinterpret_cast<volatile uint32_t *>(&qs.fill[wv][Q]));
    if (f) queue_flush<NQ, Q>(qs, wv, f);
}

}  // namespace hj
