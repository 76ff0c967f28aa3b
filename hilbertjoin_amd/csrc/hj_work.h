// hj_work.h -- the persistent schedules shared by the join kernels.
//
// The tile-pair list (in Hilbert or row order) is cut into chunks of CHUNK
// consecutive pairs.  No atomics: every wavefront owns a fixed set of chunks
// and reaches the end of its share; the grid is sized to the kernel's
// occupancy (persistent_grid), so all wavefronts are resident at once.
//
//   XCD = false  wavefront g of G takes chunks g, g + G, ... over the whole
//                list: the resident wavefronts sweep one window of the list,
//                and per-chunk costs average out over all wavefronts.
//   XCD = true   the chunks are cut into XCD_COUNT contiguous segments, one
//                per XCD.  Workgroups are dealt to the XCDs round-robin
//                (workgroup b runs on XCD b % 8), so the wavefronts of XCD x
//                are those of workgroups x, x + 8, ...; wavefront g of the G
//                on XCD x takes chunks seg_x + g, seg_x + g + G, ...  Tile
//                pairs that share tiles then meet in the same XCD's L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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

template <uint32_t CHUNK, bool XCD> struct WorkCursor {
    uint32_t chunk, end_chunk, stride, last_chunk, last_len;
    uint32_t in_chunk;

    __device__ __forceinline__ explicit WorkCursor(uint64_t n) : in_chunk(0) {
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const uint32_t waves_per_block = blockDim.x >> 6;
        const uint32_t nchunks = (uint32_t)((n + CHUNK - 1) / CHUNK);
        last_chunk = nchunks ? nchunks - 1 : 0;
        last_len = nchunks ? (uint32_t)(n - (uint64_t)last_chunk * CHUNK) : 0;
        if (XCD) {
            // a grid smaller than the XCD count keeps one segment, so every
            // chunk has a workgroup
            const uint32_t nx = gridDim.x < XCD_COUNT ? 1u : XCD_COUNT;
            const uint32_t xcd = blockIdx.x % nx;
            const uint32_t local_block = blockIdx.x / nx;
            // workgroups on this XCD: b = xcd + 8 k < gridDim.x
            const uint32_t blocks_here = (gridDim.x - xcd + nx - 1) / nx;
            end_chunk = (uint32_t)((uint64_t)nchunks * (xcd + 1) / nx);
            chunk = (uint32_t)((uint64_t)nchunks * xcd / nx) +
                    local_block * waves_per_block + wave;
            stride = blocks_here * waves_per_block;
        } else {
            end_chunk = nchunks;
            chunk = blockIdx.x * waves_per_block + wave;
            stride = gridDim.x * waves_per_block;
        }
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

    // next whole chunk: its first tile-pair index and its length (<= CHUNK)
    __device__ __forceinline__ bool next_chunk(uint64_t &w0, uint32_t &len) {
        if (chunk >= end_chunk) return false;
        w0 = (uint64_t)chunk * CHUNK;
        len = chunk == last_chunk ? last_len : CHUNK;
        chunk += stride;
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
