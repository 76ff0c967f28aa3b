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
#include "hj_frames.h"

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

// A buffer does not fit the caller-provided scratch memory of a call that
// allocates nothing (reported as HJ_ERR_CAPACITY).
struct CapacityError {
    std::string msg;
};

// Scratch memory of a call that must not allocate: device and host bump
// arenas.  While a thread runs such a call, t_arena points at its arena and
// every DevBuf / PinnedBuf that grows takes its memory from it instead of
// hipMalloc / hipHostMalloc (and never frees it).  Host staging memory from the
// arena is pageable, so a PinnedBuf in it synchronizes its stream instead of
// recording an event.
struct ScratchArena {
    char *dev = nullptr;
    size_t dev_cap = 0, dev_used = 0;
    char *host = nullptr;
    size_t host_cap = 0, host_used = 0;
    void *take_dev(size_t need);
    void *take_host(size_t need);
};
extern thread_local ScratchArena *t_arena;

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;  // p lies in a ScratchArena (never freed)
    void *ensure(size_t need) {
        if (need <= bytes && p) return p;
        if (t_arena) {
            p = t_arena->take_dev(need < 256 ? 256 : need);
            bytes = need < 256 ? 256 : need;
            borrowed = true;
            return p;
        }
        if (p && !borrowed) HJ_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
        borrowed = false;
        size_t alloc = need < 256 ? 256 : need;
        HJ_HIP(hipMalloc(&p, alloc));
        bytes = alloc;
        return p;
    }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        borrowed = false;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Pinned host memory (hipHostMalloc) for the small tables a join uploads and
// the results it reads back: copies from and to it are truly asynchronous,
// where pageable memory would go through the runtime's staging one copy at a
// time.
// A staging buffer records an event after the copy that reads it (record) and
// waits for that event before it is written again (wait), so a stream never
// has to drain just to reuse it.
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipEvent_t done = nullptr;  // after the last copy from or to the buffer
    bool borrowed = false;      // p lies in a ScratchArena (pageable, never freed)
    void *dptr = nullptr;       // device address of p (pinned memory only)
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    void *ensure(size_t need) {
        if (need <= bytes && p) return p;
        wait();
        dptr = nullptr;
        const size_t alloc = need < 4096 ? 4096 : need;
        if (t_arena) {
            p = t_arena->take_host(alloc);
            bytes = alloc;
            borrowed = true;
            return p;
        }
        if (p && !borrowed) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        borrowed = false;
        HJ_HIP(hipHostMalloc(&p, alloc, hipHostMallocDefault));
        bytes = alloc;
        return p;
    }
    void wait() {
        if (done && !borrowed) HJ_HIP(hipEventSynchronize(done));
    }
    void record(hipStream_t s) {
        if (borrowed) {  // pageable arena memory: the copy must be complete now
            HJ_HIP(hipStreamSynchronize(s));
            return;
        }
        if (!done) HJ_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        HJ_HIP(hipEventRecord(done, s));
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
    // The address at which a kernel may store into the buffer (pinned memory
    // is mapped into the device's address space), or null for pageable arena
    // memory, which only copies may reach.
    void *dev_ptr() {
        if (borrowed || !p) return nullptr;
        if (!dptr) HJ_HIP(hipHostGetDevicePointer(&dptr, p, 0));
        return dptr;
    }
    ~PinnedBuf() {
        if (done) (void)hipEventSynchronize(done);
        if (p && !borrowed) (void)hipHostFree(p);
        if (done) (void)hipEventDestroy(done);
    }
};

// Non-owning view of device memory inside a DevBuf.
struct DevView {
    void *p = nullptr;
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Host description of one cell grid (the EGO grid of one join).
struct Grid {
    double eps = 0, r = 0;      // r = 1.0 / eps (FP64, as the cell rule)
    int d = 0;                  // dimensionality
    int key_dims = 0;           // D': number of leading dimensions packed in the key
    int bits = 0;               // b: bits per packed dimension
    int active = -1;            // active dimensions of the stripe enumeration (-1: choose)
    std::vector<int64_t> minc;  // per key dimension: minimum cell minus the margin
    // centring / scaling for the reduced-precision filter
    std::vector<double> centre;  // per dimension
    // per dimension, when the ranges of some dimensions came from a sample:
    // the bound on |z_k| = |x_k - centre_k| * scale that every point must meet
    // for the scale, the radius bound and the lattice to hold (empty: the
    // ranges are exact)
    std::vector<double> zmax;
    double scale = 1.0;          // power of two
    double mbound = 0.0;         // bound on |x - centre| over all points
    double fr_q = 0.0;           // lattice step of tile anchors, scaled units (0: no frames)
};

// Device copy of a grid.  Packed keys hold, in the local key bits, field k
// (dimension k, k < key_dims, dimension 0 most significant) = floor(x_k * r)
// - minc[k]; see point_key in hj_device.h.
// GridDev.active of a grid whose active dimensions are still to be chosen
constexpr uint32_t ACTIVE_AUTO = 0xFFFFFFFFu;

struct GridDev {
    double r;             // 1 / eps of the grid
    double scale;         // filter operand scale (power of two)
    uint32_t key_dims;    // packed dimensions
    uint32_t bits;        // bits per packed dimension
    uint32_t active;      // active dimensions of the stripe enumeration
    uint32_t minc_off;    // minc[key_dims] at minc + minc_off
    uint32_t centre_off;  // centre[d] at centre + centre_off
    float fr_q;           // filter: lattice step of tile anchors (0: no tile frames)
};

// One query segment and the target segment it is joined with.  A plain join
// has one segment; a batch of independent joins has one per join; a serving
// batch has one per request, all against the resident index.  Keys carry the
// segment in their high bits (prefix) and the grid cells below (local bits).
struct SegDev {
    uint64_t qoff, qn;     // query points [qoff, qoff + qn): input rows = sorted positions
    uint64_t toff, tn;     // target points [toff, toff + tn) of the target set
    uint64_t qprefix;      // key prefix of the query segment
    uint64_t tprefix;      // key prefix of the target segment
    uint64_t local_mask;   // local (cell) key bits
    uint32_t qtile0;       // first layout tile of the query segment
    uint32_t ttile0;       // first layout tile of the target segment
    uint32_t grid;         // GridDev index
    uint32_t self;         // 1: pairs inside one set, counted once with i < j
    double eps2;           // FP64 threshold (eps * eps, rounded once)
    float t_in, t_out;     // filter thresholds (see filter_thresholds)
    FrameConst fc;         // filter with tile frames: band constants (hj_frames.h)
    // expanded-form FP64 kernel (see fp64_thresholds): a computed
    // |a|^2 + |b|^2 - 2 a.b above x_out is certainly outside eps, at or below
    // x_in certainly within; x_out < 0 selects the difference-form kernel
    double x_in, x_out;
};

// Host mirror of one segment.
struct SegHost {
    uint64_t qoff = 0, qn = 0, toff = 0, tn = 0;
    uint32_t qtile0 = 0, ttile0 = 0;
    int grid = 0;
    bool self = true;
    double eps = 0;
};

// Device buffers of a prepared (EGO-sorted) point set.
struct SetBufs {
    DevBuf raw;        // n x d row-major input copy (when the input came from the host)
    DevBuf keys[2];    // u64 keys, ping-pong for the sort
    DevBuf vals[2];    // u32 permutation, ping-pong for the sort
    DevBuf cols;       // sorted transposed tiles: [ntiles][dpad][64] FP64, pads -PAD
                       // (the FP64 kernels read rows and columns from it)
    DevBuf pnorm;      // [ntiles][dpad / 4][64] FP64 halved negative prefix squared
                       // norms -P/2 after every 4 dimensions, -inf for pads
                       // (expanded-form FP64 kernel)
    DevBuf frow, fcol; // reduced-precision MFMA operand fragments (row / column role)
    DevBuf frow2, fcol2;  // tile frames: the fragments of every tile in its own frame
    DevBuf fanch;      // tile frames: [ntiles][steps][16] FP32 anchor per K slot (0 off data)
    DevBuf finfo;      // tile frames: per tile {rho, framed} (float2), then the side's
                       // "some tile framed" word
    DevBuf tlo, thi;   // per-tile min/max widened keys (u64)
    DevBuf pbounds;    // query side: per-point widened key bounds (lo[n], hi[n], u64)
    DevBuf tile_seg;   // per layout tile: entry of the side's segment table
    DevBuf tfirst;     // per layout tile: its first sorted key
    PinnedBuf tile_seg_stage;  // staging of tile_seg
    int cur = 0;       // which keys/vals buffer holds the sorted result
    void release() {
        raw.release();
        for (int b = 0; b < 2; b++) {
            keys[b].release();
            vals[b].release();
        }
        for (DevBuf *d : {&cols, &pnorm, &frow, &fcol, &frow2, &fcol2, &fanch, &finfo, &tlo,
                          &thi, &pbounds, &tile_seg, &tfirst})
            d->release();
    }
};

// Device copies of a set of grids: the GridDev table, the minc and the centre
// arrays in one allocation, uploaded with one copy.
struct GridTables {
    DevBuf buf;
    DevView grids, minc, centre;  // views into buf
    DevView zmax;                 // per grid like centre (+inf where a grid has none)
    PinnedBuf host;               // staging of the last upload
    void upload(const std::vector<Grid> &g, hipStream_t s);
    void release() {
        buf.release();
        grids = minc = centre = zmax = DevView{};
    }
};

struct Ctx {
    int device = 0;
    int num_cus = 0;
    std::string name;
    hipStream_t stream = nullptr;
    // a second stream of default priority: the work list is built on it while
    // the main stream writes the kernels' operand layouts
    hipStream_t stream2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::string err;
    SetBufs A, B;
    DevBuf minmax, scratch_hist, scratch_scan, flags;
    // set by launch_keys when some query point's key bounds were widened
    // (beside the non-finite flag in minmax, zeroed with it)
    uint32_t *widened = nullptr;
    // set by the filter operand kernels when a point lies outside the ranges
    // its grid's operands were scaled for (Grid::zmax); zeroed with widened,
    // the low half of an aligned 64-bit word (read back as one)
    uint32_t *opflag = nullptr;
    DevBuf ivals, icount, ioff, pairs[2], hkeys[2];
    DevBuf counters;    // u64 counters for the join kernels
    DevBuf est;         // work-list estimates of the active-dimension choice
    DevBuf seg_counts;  // per-segment counts
    DevBuf band;        // band-pair queue of the reduced-precision filter
    DevBuf csr_rows, csr_cols, csr_raw;  // CSR output: row pointers, columns, raw pairs
    DevBuf segs;        // segment tables and offsets of the current join
    PinnedBuf segs_stage;  // staging of segs
    PinnedBuf chunk_stage;  // staging of the coordinate-range chunk table
    // pinned destinations of the read-backs: segment bounds, the work-list
    // size and active dimensions, the counters of a count
    PinnedBuf rb_bounds, rb_list, rb_counts;
    PinnedBuf stage[2];  // staging of large host <-> device copies (copy_d2h / copy_h2d)
    // the coordinate-range passes of a join: [0] / [1] the key candidate
    // dimensions of the query / target side, [2] / [3] the dimensions past
    // them (on stream2, beside the keys and the sort)
    struct RangePass {
        DevBuf out, tmp;       // ranges and flags; chunk table and partial ranges
        PinnedBuf stage, rb;   // staging of the chunk table; read-back
    } ranges[4];
    GridTables gt;      // grids of the current join
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
// Per-segment min / max of dimensions [k0, k1) of rows of d coordinates:
// mins/maxs[seg * (k1 - k0) + k - k0] and segbad[seg] = 1 when a coordinate of
// the segment's range is not finite, else 0 (device addresses, mapped pinned
// memory included); seg_off: host array of nseg + 1 row offsets.  Clears
// clear[0 .. 3] (when not null) in stream order.  A segment of at
// least sample_min_rows rows is sampled: only every sample_step-th row counts.
void launch_seg_minmax(const double *pts, const uint64_t *seg_off, uint32_t nseg, int d,
                       int k0, int k1, double *mins, double *maxs, uint32_t *segbad,
                       uint32_t *clear, DevBuf &tmp, PinnedBuf &stage, hipStream_t s,
                       uint32_t sample_step = 1, uint64_t sample_min_rows = 0);
// Packed keys (segment prefix | cells) and the identity permutation; with plo
// != nullptr also every point's widened local key bounds (point_key_bounds)
// in plo / phi, by input row.
void launch_keys(const double *pts, size_t n, int d, const SegDev *segs, uint32_t nseg,
                 const uint64_t *seg_off, const GridDev *grids, const int64_t *minc,
                 int kd_max, bool query_prefix, uint64_t *keys, uint32_t *vals, uint64_t *plo,
                 uint64_t *phi, uint32_t *widened, hipStream_t s);
// FP64 layout of the sorted points, per layout tile: tiles [ntiles][dpad][64]
// (pads -PAD) and the halved negative prefix squared norms -P/2 after every 4
// dimensions, [ntiles][dpad / 4][64] (pads -inf).
void launch_gather(const double *pts, const uint32_t *perm, const SegDev *segs,
                   const uint32_t *tile_seg, size_t ntiles, bool target_side, int d, int dpad,
                   double *cols, double *pnorm, hipStream_t s);
// Per layout tile: its first sorted key (tfirst) and, with tlo / thi, its
// widened key bounds from the per-point bounds of launch_keys (read through
// the EGO permutation perm; the tile's first and last keys when launch_keys
// left *widened == 0: no point of the set near a cell boundary).
void launch_tile_index(const uint64_t *keys, const uint32_t *perm, const uint64_t *plo,
                       const uint64_t *phi, const SegDev *segs, const uint32_t *tile_seg,
                       size_t ntiles, bool target_side, const uint32_t *widened,
                       uint64_t *tfirst, uint64_t *tlo, uint64_t *thi, hipStream_t s);
// counts[seg * 8 + a] = distinct prefixes of the first a key fields (a = 1..6)
// over the target points of each segment (sorted keys)
void launch_prefix_classes(const uint64_t *keys, const SegDev *segs, uint32_t nseg,
                           const GridDev *grids, uint64_t *counts, uint64_t max_tn,
                           hipStream_t s);
// Sets GridDev.active of segment k's grid (grid k) where it is ACTIVE_AUTO:
// the choice that minimizes the estimated work-list cost, from `samples`
// query tiles per segment (tlo / thi: query tile bounds; est: scratch of
// nseg * 8 u64; code_cost: cost of one stripe code of one query
// tile relative to one tile pair of the count kernel).
void launch_choose_active(const uint64_t *tlo, const uint64_t *thi, const uint64_t *tkeys,
                          const uint64_t *tfirst, const SegDev *segs, uint32_t nseg,
                          GridDev *grids, uint32_t samples, uint64_t *est, float code_cost,
                          hipStream_t s);
// per query tile: merged candidate tile intervals against the sorted target
// keys; the candidate point pairs of all tiles are added to cand_total[0 ..
// CAND_SLOTS) (tile I into slot I % CAND_SLOTS)
constexpr uint32_t CAND_SLOTS = 64;
// (tkeys: sorted target keys; tfirst: first key of every target layout tile)
void launch_tile_ranges(const uint64_t *tlo, const uint64_t *thi, const uint32_t *tile_seg,
                        size_t ntiles_q, const uint64_t *tkeys, const uint64_t *tfirst,
                        const SegDev *segs, const GridDev *grids, uint32_t *ivals,
                        uint32_t *icount, uint64_t *jcount, int max_iv, uint64_t *cand_total,
                        hipStream_t s);
// First sorted key of every layout tile of a side (query or target spans).
// (entries at or past cap are not written)
void launch_emit_pairs(const uint32_t *ivals, const uint32_t *icount, const uint64_t *off,
                       size_t ntiles_q, int max_iv, uint64_t *pairs, uint64_t cap,
                       hipStream_t s);
void launch_hilbert_keys(const uint64_t *pairs, size_t npairs, int order, uint64_t *hkeys,
                         hipStream_t s);
void launch_hilbert_xy(const uint32_t *x, const uint32_t *y, size_t n, int order,
                       uint64_t *out, hipStream_t s);
void launch_zero_u64(uint64_t *p, size_t n, hipStream_t s);
// n words from device memory into pinned host memory (hipHostMalloc, mapped
// into the device's address space) by a kernel: unlike a small
// hipMemcpyAsync to the host, which waits on the host until the stream
// reaches it, it returns at once; the host reads dst after an event or a
// synchronization on s.
void launch_store_u64(const uint64_t *src, uint64_t *dst, size_t n, hipStream_t s);
// n words from device memory into buf at byte offset off: by launch_store_u64
// where the buffer is pinned, by a copy where it is pageable (arena memory).
void readback_u64(PinnedBuf &buf, size_t off, const uint64_t *src, size_t n, hipStream_t s);
// Several such read-backs in one launch: range i copies r[i].n words from
// r[i].src to byte offset (size_t)r[i].dst of buf.
constexpr int READBACK_RANGES = 5;
struct U64Range {
    const uint64_t *src;
    uint64_t *dst;
    size_t n;
};
struct U64Ranges {
    U64Range r[READBACK_RANGES];
    int nr;
};
void readback_ranges(PinnedBuf &buf, const U64Ranges &rs, hipStream_t s);
void launch_copy16(const void *src, void *dst, size_t bytes, int num_cus, hipStream_t s);
// per-dimension cell histograms: hist[hoff[k] + floor(x_k r) - cmin[k]] (u32
// counts, zeroed by the caller); LDS-private copies when all bins fit
constexpr size_t DIM_HIST_LDS_BYTES = 64 * 1024;
void launch_dim_hist(const double *pts, size_t n, int d, double r, const int64_t *cmin,
                     const uint32_t *hoff, uint32_t nbins_total, uint32_t *hist, int num_cus,
                     hipStream_t s);
// out[i][perm[k]] = in[i][k] over n rows of d columns
void launch_permute_cols(const double *in, double *out, size_t n, int d, const int32_t *perm,
                         int num_cus, hipStream_t s);

// ---- FP64 join kernel (hj_fp64.hip) -------------------------------------------
// 64-bit words of the filter's scheduling counters (XcdCursor in hj_work.h:
// 8 XCDs x 8 counters, one per 128-byte line), zeroed with the join counters.
constexpr size_t SCHED_WORDS = 1024;
// Join counters: [0] MFMAs issued (filter), [1] count, [2] pair-dims evaluated,
// [3] candidate pairs, [4] band pairs, [5] tile frames used by the filter (0,
// 1: slow path, 2: both paths), [6] pair-output cursor, [7] filter early
// exits.  Per-segment counts go to seg_counts[seg].
struct JoinArgs {
    const uint64_t *pairs;
    uint64_t npairs;  // the list's length, or its capacity when npairs_dev is set
    const SegDev *segs;
    const uint32_t *tile_seg;  // query layout tile -> segment
    uint64_t *ctr;
    uint64_t *seg_counts;
    int grid;      // persistent workgroups (0: from occupancy)
    int num_cus;
    int tile_rows = 8;   // FP64 kernel: rows per early-exit vote (8 or 16)
    int chunk = 4;       // MFMA filter: tile pairs per operand prefetch chunk (2, 4 or 8)
    uint32_t *sched = nullptr;  // MFMA filter: SCHED_WORDS zeroed words (XcdCursor)
    // MFMA filter: tiles may have their own frames (hj_frames.h; the sides'
    // flags say whether some tile took one); the sides' frame tables
    bool frames = false;
    struct {
        const void *fa2 = nullptr, *fb2 = nullptr;  // own-frame fragments (row, column role)
        const float2 *qinfo = nullptr, *tinfo = nullptr;
        const float *qanch = nullptr, *tanch = nullptr;
        const uint32_t *qflag = nullptr, *tflag = nullptr;
    } fr;
    // MFMA filter: the list's length on the device (null: npairs is it)
    const uint64_t *npairs_dev = nullptr;
    // the pairs of a query tile are consecutive in the list (row order)
    bool rows_grouped = false;
    // MFMA filter: set when the operands' ranges did not hold (the launch
    // then does nothing)
    const uint32_t *opflag = nullptr;
};
// rows: the query side's FP64 tile layout, cols: the target side's (the same
// array for a self-join); rnorm / cnorm their prefix norms.  expanded: the
// expanded-form kernel (every segment's x_out >= 0), else the difference form.
void launch_fp64_count(const JoinArgs &ja, const double *rows, const double *cols,
                       const double *rnorm, const double *cnorm, int dpad, bool self, int kblock,
                       bool expanded, hipStream_t s);
// pair output: appends (layout row | layout col << 32) through counter [6]
void launch_fp64_pairs(const JoinArgs &ja, const double *rows, const double *cols, int dpad,
                       bool self, uint64_t *out, uint64_t out_cap, hipStream_t s);
// pair output post-processing: layout (row, col) -> keys (query input row)
// << 32 | (target input row - toff of its segment), with the pair mirrored
// for self segments; the segment id goes to out_seg
void launch_map_pairs(const uint64_t *in, uint64_t np, const SegDev *segs,
                      const uint32_t *tile_seg, const uint32_t *qperm, const uint32_t *tperm,
                      bool mirror, uint64_t *out_keys, uint32_t *out_seg, hipStream_t s);
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
    int frames;    // tiles may take their own operand frames (hj_frames.h)
    int frames_fast;  // and the fast path tests in them (the grid-frame band is wide)
};
// Operand fragments per layout tile straight from the input rows through the
// EGO permutation.  filter_prep_fused(d): the tile is staged in LDS
// (row-contiguous loads).
bool filter_prep_fused(int d);
// Tile frames (hj_frames.h): own-frame fragments of the framed tiles
// (frow2 / fcol2 where frow / fcol are built), per-slot anchors, per-tile
// {rho, framed} and the side's "some tile framed" word (zeroed by the
// caller); a null anch builds the grid-frame fragments only.
struct FrameOut {
    void *frow2 = nullptr, *fcol2 = nullptr;
    float *anch = nullptr;
    float2 *info = nullptr;
    uint32_t *flag = nullptr;
};
void launch_filter_prep(const double *pts, const uint32_t *perm, const SegDev *segs,
                        const uint32_t *tile_seg, size_t ntiles, bool target_side, int d,
                        const GridDev *grids, const double *centre, const FilterParams &fp,
                        void *frow, void *fcol, const FrameOut &fo, const double *zmax,
                        uint32_t *opflag, hipStream_t s);

// Outputs of the filter: the band queue (ctr[4] counts its entries), and
// optionally the accepted pairs themselves (ctr[6] counts them; a null pairs
// only counts).  Both hold layout positions row | col << 32; entries past a
// capacity are counted but not stored.
// Entries allocated past the end of a work list: the MFMA filter loads a whole
// chunk of pairs at once and discards what lies past the end.
constexpr uint64_t WORK_LIST_PAD = 8;

struct FilterOut {
    uint64_t *band = nullptr;
    uint64_t band_cap = 0;
    uint64_t *pairs = nullptr;
    uint64_t pairs_cap = 0;
};
void launch_mfma_count(const JoinArgs &ja, const void *fragA, const void *fragB,
                       const FilterParams &fp, bool self, const FilterOut &out, hipStream_t s);
// FP64 recheck of the band queue; its length is read from *nband_dev on the
// device (entries beyond out.band_cap are skipped); accepted band pairs are
// counted and, with out.pairs, appended
void launch_band_recheck(const double *ptsA, const uint32_t *permA, const double *ptsB,
                         const uint32_t *permB, const SegDev *segs, const uint32_t *tile_seg,
                         int d, const uint64_t *nband_dev, const FilterOut &out, uint64_t *ctr,
                         uint64_t *seg_counts, int num_cus, hipStream_t s);

}  // namespace hj
