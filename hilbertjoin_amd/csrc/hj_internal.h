// hj_internal.h -- shared declarations of the MI355X epsilon-join library.
//
// Everything here is internal; the public surface is include/hilbertjoin_amd.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/hilbertjoin_amd.h"

namespace hj {

// Points per tile.  One wavefront (64 lanes) owns one tile of columns; a
// tile pair (I, J) is the unit of work of every join kernel.
constexpr int TILE = 64;

// Pad coordinates.  Row-side pads sit at +PAD, column-side pads at -PAD, so a
// pad never lies within any finite eps of a real point or of another pad; the
// kernels also mask pads explicitly when counting.
constexpr double PAD = 1e150;

struct HipError {
    hipError_t code;
    std::string where;
};

#define HJ_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) throw ::hj::HipError{e_, std::string(#call) + " @ " +   \
                                                         __FILE__ + ":" +             \
                                                         std::to_string(__LINE__)};   \
    } while (0)

struct InvalidArg {
    std::string msg;
};

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void *ensure(size_t need) {
        if (need <= bytes && p) return p;
        if (p) HJ_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
        size_t alloc = need < 256 ? 256 : need;
        HJ_HIP(hipMalloc(&p, alloc));
        bytes = alloc;
        return p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Grid description shared by both sides of a join.
struct Grid {
    double eps = 0, r = 0;      // r = 1.0 / eps (FP64, as the cell rule)
    int d = 0;                  // dimensionality
    int key_dims = 0;           // D': number of leading dimensions packed in the key
    int bits = 0;               // b: bits per packed dimension
    int active = 0;             // active dimensions of the stripe enumeration
    std::vector<int64_t> minc;  // per key dimension: minimum cell
    // centring / scaling for the reduced-precision filter
    std::vector<double> centre;  // per dimension
    double scale = 1.0;          // power of two
    double mbound = 0.0;         // bound on |x - centre| over all points
};

// Device-side description of a prepared (EGO-sorted) point set.
struct SetBufs {
    DevBuf raw;        // n x d row-major input copy (when the input came from the host)
    DevBuf keys[2];    // u64 keys, ping-pong for the sort
    DevBuf vals[2];    // u32 permutation, ping-pong for the sort
    DevBuf rows;       // sorted rows: [ntiles*64][dpad] FP64, pads +PAD
    DevBuf cols;       // sorted transposed tiles: [ntiles][dpad][64] FP64, pads -PAD
    DevBuf frow, fcol; // reduced-precision MFMA operand fragments (row / column role)
    DevBuf norms;      // FP64 squared norms of the rounded operands
    DevBuf tlo, thi;   // per-tile min/max widened keys (u64)
    DevBuf rowptr;     // CSR helpers
    int cur = 0;       // which keys/vals buffer holds the sorted result
};

struct Ctx {
    int device = 0;
    int num_cus = 0;
    std::string name;
    hipStream_t stream = nullptr;
    std::string err;
    SetBufs A, B;
    DevBuf minmax, scratch_hist, scratch_scan, flags;
    DevBuf ivals, icount, ioff, pairs[2], hkeys[2];
    DevBuf counters;  // u64 counters for the join kernels
    DevBuf band;      // band-pair queue of the reduced-precision filter
    DevBuf csr_rows, csr_cols;
    DevBuf centre;    // per-dimension centre of the filter's operands
    hipEvent_t ev[8] = {};
};

// ---- scan / sort (hj_sort.hip) ------------------------------------------------
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, DevBuf &tmp,
                        hipStream_t s);
void exclusive_scan_u64(const uint64_t *in, uint64_t *out, size_t n, DevBuf &tmp,
                        hipStream_t s);
// LSD radix sort of (u64 key, V value) over key bits [0, bits).  The sorted
// result ends in buffer index *which (0 or 1) of keys/vals.
void radix_sort_u64_u32(uint64_t *keys[2], uint32_t *vals[2], size_t n, int bits, int *which,
                        DevBuf &hist, DevBuf &scan_tmp, hipStream_t s);
void radix_sort_u64_u64(uint64_t *keys[2], uint64_t *vals[2], size_t n, int bits, int *which,
                        DevBuf &hist, DevBuf &scan_tmp, hipStream_t s);

// ---- preparation kernels (hj_prep.hip) --------------------------------------------
void launch_minmax(const double *pts, size_t n, int d, double *mins, double *maxs,
                   uint32_t *bad, DevBuf &tmp, hipStream_t s);
void launch_keys(const double *pts, size_t n, int d, const Grid &g, const int64_t *minc_dev,
                 uint64_t *keys, uint32_t *vals, hipStream_t s);
void launch_gather(const double *pts, const uint32_t *perm, size_t n, int d, int dpad,
                   double *rows, double *cols, hipStream_t s);
void launch_tile_bounds(const double *cols, size_t n, int dpad, const Grid &g,
                        const int64_t *minc_dev, uint64_t *tlo, uint64_t *thi, hipStream_t s);
// per query tile: merged candidate tile intervals against the sorted target keys
void launch_tile_ranges(const uint64_t *tlo, const uint64_t *thi, size_t ntiles_q,
                        const uint64_t *tkeys, size_t nt, const Grid &g, bool self,
                        uint32_t *ivals, uint32_t *icount, uint64_t *jcount, int max_iv,
                        hipStream_t s);
void launch_emit_pairs(const uint32_t *ivals, const uint32_t *icount, const uint64_t *off,
                       size_t ntiles_q, int max_iv, uint64_t *pairs, hipStream_t s);
void launch_hilbert_keys(const uint64_t *pairs, size_t npairs, int order, uint64_t *hkeys,
                         hipStream_t s);
void launch_hilbert_xy(const uint32_t *x, const uint32_t *y, size_t n, int order,
                       uint64_t *out, hipStream_t s);
void launch_zero_u64(uint64_t *p, size_t n, hipStream_t s);
void launch_copy16(const void *src, void *dst, size_t bytes, int num_cus, hipStream_t s);
// counts[a] = distinct prefixes of the first a key fields (a = 1..6) of sorted keys
void launch_prefix_classes(const uint64_t *keys, size_t n, int key_dims, int bits,
                           uint64_t *counts, int num_cus, hipStream_t s);

// ---- FP64 join kernel (hj_fp64.hip) -------------------------------------------
struct JoinCounters {
    // [0] unused, [1] count, [2] pair-dims evaluated, [3] candidate pairs,
    // [4] band pairs, [5] unused, [6] pair-output cursor, [7] filter early exits
    uint64_t *c;
};
void launch_fp64_count(const double *rows, const double *cols, const uint64_t *pairs,
                       uint64_t npairs, uint32_t nA, uint32_t nB, int dpad, double eps2,
                       bool self, int kblock, uint64_t *counters, int grid, int num_cus,
                       hipStream_t s);
// pair output: appends (i | j << 32) of sorted indices through counter [6]
void launch_fp64_pairs(const double *rows, const double *cols, const uint64_t *pairs,
                       uint64_t npairs, uint32_t nA, uint32_t nB, int dpad, double eps2,
                       bool self, uint64_t *counters, uint64_t *out, uint64_t out_cap, int grid,
                       int num_cus, hipStream_t s);
// CSR post-processing: sorted-index pairs -> original-index keys (row << 32 | col),
// mirrored for self-joins; rowptr from sorted keys
void launch_map_pairs(const uint64_t *in, uint64_t np, const uint32_t *permA,
                      const uint32_t *permB, bool mirror, uint64_t *out_keys, uint32_t *out_vals,
                      hipStream_t s);
void launch_csr_rowptr(const uint64_t *keys, uint64_t nk, uint64_t nrows, int64_t *rowptr,
                       hipStream_t s);
void launch_csr_cols(const uint64_t *keys, uint64_t nk, int64_t *colidx, hipStream_t s);

// ---- reduced-precision filter (hj_mfma.hip) ------------------------------------
struct FilterParams {
    float t_out;   // acc < t_out  => certainly outside eps
    float t_in;    // acc >= t_in  => certainly inside eps
    int split;     // data dimensions of the first K chunk (d: no early exit)
    int steps1;    // k-steps (16 slots) of the first chunk
    int steps;     // k-steps in total
    int parts;     // norm parts per side folded into each chunk
    int bf16;      // operand type
};
void launch_op_norms(const double *rows, size_t n, int d, int dpad, int split,
                     const double *centre, double scale, bool bf16, double *norms,
                     hipStream_t s);
void launch_frag_pack(const double *rows, size_t n, int d, int dpad, const double *centre,
                      const double *norms, double scale, const FilterParams &fp, void *frow,
                      void *fcol, hipStream_t s);
void launch_mfma_count(const void *fragA, const void *fragB, const uint64_t *pairs,
                       uint64_t npairs, uint32_t nA, uint32_t nB, const FilterParams &fp,
                       bool self, uint64_t *counters, uint64_t *band, uint64_t band_cap,
                       int grid, int num_cus, hipStream_t s);
void launch_band_recheck(const double *ptsA, const uint32_t *permA, const double *ptsB,
                         const uint32_t *permB, int d, double eps2, const uint64_t *band,
                         uint64_t nband, uint64_t *counters, hipStream_t s);
// fused tile bounds + operand fragments straight from the input (d small enough)
bool filter_prep_fused(int d);
void launch_filter_prep(const double *pts, const uint32_t *perm, size_t n, int d,
                        const double *centre, double scale, const FilterParams &fp,
                        const Grid &g, const int64_t *minc, uint64_t *tlo, uint64_t *thi,
                        void *frow, void *fcol, hipStream_t s);

}  // namespace hj
