// hj_work.h -- the persistent schedule shared by the join kernels.
//
// The tile-pair list (in Hilbert or row order) is cut into chunks of CHUNK
// consecutive pairs.  Wavefront g of G resident wavefronts takes chunks g,
// g + G, g + 2G, ...: no atomics, every wavefront reaches the end of the list,
// and at any moment the resident wavefronts work on a window of neighbouring
// chunks, so consecutive tile pairs (which share tiles) meet in L2.  The grid
// is sized to the kernel's occupancy (persistent_grid), so all wavefronts are
// resident at once.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hj {

constexpr uint64_t WORK_CHUNK = 4;

struct WorkCursor {
    uint64_t chunk, stride, npairs;
    uint32_t in_chunk;

    __device__ __forceinline__ explicit WorkCursor(uint64_t n) : npairs(n), in_chunk(0) {
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const uint64_t waves_per_block = blockDim.x >> 6;
        chunk = (uint64_t)blockIdx.x * waves_per_block + wave;
        stride = (uint64_t)gridDim.x * waves_per_block;
    }

    // next tile-pair index of this wavefront, or false at the end of its share
    __device__ __forceinline__ bool next(uint64_t &w) {
        w = chunk * WORK_CHUNK + in_chunk;
        if (w >= npairs) return false;
        if (++in_chunk == WORK_CHUNK) {
            in_chunk = 0;
            chunk += stride;
        }
        return true;
    }
};

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
