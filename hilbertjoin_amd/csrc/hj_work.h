// hj_work.h -- the persistent schedules shared by the join kernels.
//
// The tile-pair list (in Hilbert or row order) is cut into chunks of CHUNK
// consecutive pairs; the grid is sized to the kernel's occupancy
// (persistent_grid), so all wavefronts are resident at once.
//
//   WorkCursor   no atomics: wavefront g of G takes chunks g, g + G, ... over
//                the whole list, and reaches the end of its share.
//   XcdCursor    the chunks are cut into XCD_COUNT contiguous segments, one
//                per XCD (workgroup b runs on XCD b % 8), and each segment is
//                handed out in order by XCD_COUNTERS counters of that XCD:
//                counter r's k-th fetch covers the FETCH chunks starting at
//                chunk (k XCD_COUNTERS + r) FETCH of the segment.  The
//                resident wavefronts of an XCD then work on one narrow window
//                of its segment however unevenly their chunks cost (a static
//                deal drifts apart over a long list, and the column tiles of
//                the window no longer fit the XCD's L2: 74% L2 hits against
//                93% for the filter at n = 1M, d = 64).  Several counters per
//                XCD, each fetch covering a few chunks, keep the atomics off
//                the critical path: fetches on one address serialize at about
//                20 ns each, and a fetch's value is read FETCH chunks after it
//                was issued (a device-scope atomic takes microseconds to
//                return under load).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hj_internal.h"

namespace hj {

constexpr uint32_t XCD_COUNT = 8;

// Real points of a 64-point tile with `rem` points left in its segment: an
// integer select (the generic min overloads for 64-bit integers go through
// FP64 conversions).
__device__ __forceinline__ uint32_t tile_points(uint64_t rem) {
    return rem < 64 ? (uint32_t)rem : 64u;
}

// Chunk indices are 32-bit (scalar arithmetic throughout); the host keeps a
// list below WORK_MAX_CHUNKS chunks (work_list_fits).
constexpr uint64_t WORK_MAX_CHUNKS = (uint64_t)1 << 31;

template <uint32_t CHUNK> struct WorkCursor {
    uint32_t chunk, end_chunk, stride, last_chunk, last_len;
    uint32_t in_chunk;

    __device__ __forceinline__ explicit WorkCursor(uint64_t n) : in_chunk(0) {
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const uint32_t waves_per_block = blockDim.x >> 6;
        const uint32_t nchunks = (uint32_t)((n + CHUNK - 1) / CHUNK);
        last_chunk = nchunks ? nchunks - 1 : 0;
        last_len = nchunks ? (uint32_t)(n - (uint64_t)last_chunk * CHUNK) : 0;
        end_chunk = nchunks;
        chunk = blockIdx.x * waves_per_block + wave;
        stride = gridDim.x * waves_per_block;
    }

    // next tile-pair index of this wavefront, or false at the end of its share
    __device__ __forceinline__ bool next(uint64_t &w) {
        if (chunk >= end_chunk) return false;
        const uint32_t len = chunk == last_chunk ? last_len : CHUNK;
        if (in_chunk >= len) return false;
        w = (uint64_t)chunk * CHUNK + in_chunk;
        if (++in_chunk == len) {
            in_chunk = 0;
            chunk += stride;
        }
        return true;
    }
};

// Scheduling counters of XcdCursor: XCD_COUNTERS per XCD, one per 128-byte
// line, zeroed before every launch (SCHED_WORDS 64-bit words, hj_internal.h).
constexpr uint32_t XCD_COUNTERS = 8;
constexpr uint32_t XCD_FETCH = 2;  // default chunks per fetch
constexpr uint32_t SCHED_LINE_WORDS = 32;  // 32-bit words per counter line
static_assert((size_t)XCD_COUNT * XCD_COUNTERS * SCHED_LINE_WORDS / 2 == SCHED_WORDS,
              "scheduling counter area");

template <uint32_t CHUNK, uint32_t FETCH = XCD_FETCH> struct XcdCursor {
    uint32_t base, end, last_chunk, last_len, r, cur, j, pending;
    uint32_t *cnt;

    __device__ __forceinline__ XcdCursor(uint64_t n, uint32_t *sched) {
        const uint32_t nchunks = (uint32_t)((n + CHUNK - 1) / CHUNK);
        last_chunk = nchunks ? nchunks - 1 : 0;
        last_len = nchunks ? (uint32_t)(n - (uint64_t)last_chunk * CHUNK) : 0;
        // a grid smaller than the XCD count keeps one segment, so every chunk
        // has a workgroup
        const uint32_t nx = gridDim.x < XCD_COUNT ? 1u : XCD_COUNT;
        const uint32_t xcd = blockIdx.x % nx;
        base = (uint32_t)((uint64_t)nchunks * xcd / nx);
        end = (uint32_t)((uint64_t)nchunks * (xcd + 1) / nx);
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        r = ((blockIdx.x / nx) * (blockDim.x >> 6) + wave) % XCD_COUNTERS;
        cnt = sched + SCHED_LINE_WORDS * (xcd * XCD_COUNTERS + r);
        j = FETCH;  // nothing taken yet
        pending = fetch();
    }

    // one lane fetches; the value is read (and waited for) at the next refill
    __device__ __forceinline__ uint32_t fetch() {
        uint32_t v = 0;
        if ((threadIdx.x & 63) == 0) v = atomicAdd(cnt, 1u);
        return v;
    }

    // next whole chunk: its first tile-pair index and its length (<= CHUNK)
    __device__ __forceinline__ bool next_chunk(uint64_t &w0, uint32_t &len) {
        if (j >= FETCH) {
            const uint32_t k = __builtin_amdgcn_readfirstlane(pending);
            const uint64_t c0 =
                base + ((uint64_t)k * XCD_COUNTERS + r) * FETCH;  // no 32-bit wrap
            j = 0;
            if (c0 >= end) return false;
            cur = (uint32_t)c0;
            pending = fetch();  // in flight while this fetch's chunks run
        }
        const uint32_t c = cur + j++;
        if (c >= end) return false;
        w0 = (uint64_t)c * CHUNK;
        len = c == last_chunk ? last_len : CHUNK;
        return true;
    }
};

// The host checks the list size against the 32-bit chunk indices.
inline bool work_list_fits(uint64_t npairs, uint32_t chunk) {
    return (npairs + chunk - 1) / chunk < WORK_MAX_CHUNKS;
}

// Blocks that fit on the device at once for `kernel` with `block` threads.
template <class K> int persistent_grid(K kernel, int block, int num_cus, int requested) {
    if (requested > 0) return requested;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return per_cu * (num_cus > 0 ? num_cus : 1);
}

}  // namespace hj
