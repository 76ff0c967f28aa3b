// hj_join.h -- the segmented join pipeline shared by the single join, the
// batched joins and the resident-index queries.
//
// A problem is a query side Q and a target side T (T = Q for self-joins),
// each cut into segments of consecutive rows.  Join segment k pairs query
// segment k with target segment segs[k].t under grid segs[k].grid and radius
// segs[k].eps.  Sorting the keys (segment prefix | grid cells) leaves every
// segment's points contiguous and in EGO order, so one work list and one
// persistent kernel launch cover all segments.
#pragma once

#include <map>
#include <new>
#include <utility>
#include <vector>

#include "hj_internal.h"

namespace hj {
// Autotuned settings of a context, per shape class of input.
struct TuneKey {
    int kernel, self, batch, d, log2n, eps4;
    bool operator<(const TuneKey &o) const {
        const int a[6] = {kernel, self, batch, d, log2n, eps4};
        const int b[6] = {o.kernel, o.self, o.batch, o.d, o.log2n, o.eps4};
        for (int i = 0; i < 6; i++)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
};
}  // namespace hj

// The opaque context of the C ABI.
struct hj_ctx {
    hj::Ctx c;
    std::map<hj::TuneKey, hj_opts> tuned;
};

namespace hj {

// Runs f, turning the library's exceptions into status codes and a message.
template <class F> int guarded_call(std::string &err, F &&f) {
    try {
        err.clear();
        return f();
    } catch (const InvalidArg &e) {
        err = e.msg;
        return HJ_ERR_INVALID;
    } catch (const CapacityError &e) {
        err = e.msg;
        return HJ_ERR_CAPACITY;
    } catch (const HipError &e) {
        err = std::string(hipGetErrorString(e.code)) + " in " + e.where;
        return e.code == hipErrorOutOfMemory ? HJ_ERR_NOMEM : HJ_ERR_HIP;
    } catch (const std::bad_alloc &) {
        err = "host allocation failed";
        return HJ_ERR_NOMEM;
    } catch (...) {
        err = "unexpected exception";
        return HJ_ERR_INTERNAL;
    }
}

struct Opts {
    int kernel = HJ_KERNEL_FP64;
    int active = -1;
    int order = HJ_ORDER_ROW;
    int kblock = 4;
    int grid = 0;
    int split = 0;  // filter: first-chunk dimensions (0 auto, -1 no early exit)
    int tile_rows = 8;   // FP64 kernel: rows per early-exit vote (8 or 16)
    int chunk = 4;       // MFMA filter: tile pairs per operand prefetch chunk (2, 4, 8)
    bool tune = false;
    bool frames = true;  // MFMA filter: tile frames allowed (hj_frames.h)
};

Opts read_opts(const hj_opts *o);
void check_sizes(uint64_t n, uint32_t d);
// Dimensions per row of the FP64 tile layout: d rounded up to whole
// early-exit blocks of the FP64 kernel (multiples of 8, or of kblock = 16).
int dpad_for(int d, int kblock);
inline size_t ntiles_of(uint64_t n) { return (size_t)((n + 63) / 64); }

// Grid of one join from the per-dimension bounds of its points: cells of width
// eps, bits per key field, key dimensions within key_budget bits (and, for
// npts > 0 points, no more than the sorted order of that many points can use),
// filter centre and scale (the scale leaves 2^headroom room for points beyond
// the bounds: serving queries).
Grid grid_from_bounds(const double *lo, const double *hi, int d, double eps, int key_budget,
                      int active_req, int headroom, uint64_t npts);
// The same in two parts: the key part (cells, bits, key dimensions, cell
// minima, active dimensions) from the bounds of the leading dk <= d
// dimensions, which then hold every key field, and the operand part (centre,
// radius bound, scale) from the bounds of all d (the centre is sized by
// grid_keys, so a grid's tables keep their layout).
Grid grid_keys(const double *lo, const double *hi, int dk, int d, double eps, int key_budget,
               int active_req, uint64_t npts);
void grid_operands(Grid &g, const double *lo, const double *hi, int headroom);

// Per-segment min / max of each dimension, read back: lo/hi[seg * d + k].
// Throws InvalidArg on a non-finite coordinate.
void seg_bounds(Ctx &c, const double *pts, const std::vector<uint64_t> &off, int d,
                std::vector<double> &lo, std::vector<double> &hi);

// Power-of-two scale of the filter operands for points within mbound of the
// centre: every scaled coordinate vector stays within 2^(14 - headroom).
double operand_scale(double mbound, int headroom);

// Largest |x - centre| over the points whose bounds are lo/hi (d values each).
double radius_bound(const double *lo, const double *hi, const std::vector<double> &centre);

// Filter layout of the K axis (split, k-steps, norm parts) for d dimensions.
FilterParams filter_layout(int d, bool bf16, int split_req);
// Thresholds x_in / x_out of the expanded-form FP64 kernel for radius eps in
// d dimensions, when the squared norms of a row point and of a column point
// add up to at most pmax_sum; both -1 when the rounding band would not be
// thin against eps^2 (the difference-form kernel is used then).
void fp64_thresholds(int d, double eps, double pmax_sum, double *x_in, double *x_out);
// Largest squared norm of the points whose bounds are lo / hi (d values each).
double norm2_bound(const double *lo, const double *hi, int d);

// Classification thresholds for radius eps with every scaled coordinate
// vector within M of the centre (M = scale * radius bound).
void filter_thresholds(const FilterParams &fp, int d, double scale, double M, double eps,
                       float *t_in, float *t_out);
// Band constants of the filter with tile frames for one segment (M: the
// grid's bound on the scaled operand radius).
void frame_consts(const FilterParams &fp, int d, double scale, double M, double eps,
                  FrameConst *fc);
// Lattice step of tile anchors for grid g and the bounds lo, hi of its points.
double frame_lattice(const double *lo, const double *hi, const Grid &g, bool bf16);
// Whether a join of dimension d with layout fp may use tile frames.
bool frames_possible(const FilterParams &fp, int d);

struct SideSpec {
    const double *pts = nullptr;  // device rows, n x d
    uint64_t n = 0;
    std::vector<uint64_t> off;  // segment row offsets (nseg + 1)
};

struct JoinSeg {
    uint32_t t = 0;       // target segment
    uint32_t grid = 0;    // grid index
    double eps = 0;       // radius (<= the grid's eps)
    double radius = -1;   // bound on |x - centre| of both sides (< 0: the grid's)
};

// A prepared side: sorted keys, EGO permutation, layouts, tile tables.
struct Side {
    const double *pts = nullptr;
    uint64_t n = 0;
    size_t ntiles = 0;
    int d = 0, dpad = 0;
    std::vector<uint64_t> off;    // segment row offsets
    std::vector<uint32_t> tile0;  // first layout tile of each segment
    const uint64_t *keys = nullptr;
    const uint32_t *perm = nullptr;
    const double *rows = nullptr, *cols = nullptr;  // FP64 tile layout (both the same array)
    const double *pn = nullptr;                      // its prefix norms
    const void *frow = nullptr, *fcol = nullptr;
    void *frow2 = nullptr, *fcol2 = nullptr;  // tile frames: own-frame fragments
    const float *fanch = nullptr;    // tile frames: anchor slots (hj_frames.h)
    const float2 *finfo = nullptr;   // tile frames: per tile {rho, framed}
    const uint32_t *fflag = nullptr; // tile frames: set when some tile is framed
    const uint64_t *tlo = nullptr, *thi = nullptr;
    const uint64_t *plo = nullptr, *phi = nullptr;  // per-point key bounds (query side)
    const uint32_t *widened = nullptr;  // some point's bounds were widened (query side)
    const uint32_t *tile_seg = nullptr;  // layout tile -> entry of the side's segment table
    const uint64_t *tfirst = nullptr;    // layout tile -> its first sorted key
};

// Which layouts a side needs.
struct SideNeeds {
    bool fp64_rows = false, fp64_cols = false;  // FP64 layouts
    bool frag_row = false, frag_col = false;    // filter fragments (row / column role)
    bool bounds = false;                        // per-tile key bounds (query side)
};

// Tile layout of a side: first tile per segment and the total.
void side_layout(Side &S, const std::vector<uint64_t> &off, const double *pts, int d,
                 int kblock);

// Keys of one side (identity permutation).  seg_table: the side's segment
// table (one entry per segment of the side, device), seg_off_dev: its row
// offsets on the device.
void side_keys(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table, uint32_t nseg,
               const uint64_t *seg_off_dev, bool target, const GridTables &gt, int kd_max);
// LSD radix sort of the side's keys over their low sort_bits bits.
void side_sort(Ctx &c, SetBufs &B, Side &S, int sort_bits);
// Tile index of a sorted side, what the work list reads: tile -> segment
// (tile_owner[t]: the segment-table entry of tile t), first key per tile and
// (need.bounds) the tiles' widened key bounds.
void side_index(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table,
                const std::vector<uint32_t> &tile_owner, bool target, const SideNeeds &need);
// side_index in its two parts, each on a stream of the caller's choice: the
// tile -> segment table, then (once that is written) the first keys and the
// tile bounds
void side_tile_segments(Ctx &c, SetBufs &B, Side &S, const std::vector<uint32_t> &tile_owner,
                        hipStream_t s);
void side_search_index(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table, bool target,
                       const SideNeeds &need, hipStream_t s);
// Operand layouts of a sorted side, what the distance kernels read: the FP64
// tile layout and prefix norms, or the filter fragments.
// With opflag the filter operand kernels check every point against the
// grids' bounds (Grid::zmax) and set *opflag when one lies outside.  st: the
// stream (null: the context's main stream).
void side_operands(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table, bool target,
                   const GridTables &gt, const FilterParams &fp, const SideNeeds &need,
                   uint32_t *opflag = nullptr, hipStream_t st = nullptr);
// Both of the above.
void side_layouts(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table,
                  const std::vector<uint32_t> &tile_owner, bool target, const GridTables &gt,
                  const FilterParams &fp, const SideNeeds &need);
// tile_owner of a side whose segment k owns its tiles tile0[k] ..
std::vector<uint32_t> tile_owners(const Side &S);

struct WorkList {
    const uint64_t *pairs = nullptr;
    uint64_t npairs = 0;
    const uint64_t *cand = nullptr;  // device: candidate point pairs, CAND_SLOTS partial sums
    // A list built without a host wait for its length: the length stays on
    // the device (npairs_dev), npairs is the buffer's capacity until the
    // length has been read back, and entries past the capacity were not
    // written (the join is then run again with a list of the full length).
    const uint64_t *npairs_dev = nullptr;
    // what emitting the list again needs: the query tiles' runs and offsets
    size_t ntq = 0;
    int max_iv = 0;
    const uint64_t *off = nullptr;
    bool rows_grouped = false;  // row order: a query tile's pairs are consecutive
};

// Built on stream s (which it synchronizes once, for the list's length).
WorkList build_worklist(Ctx &c, const Side &Q, const Side &T, const SegDev *segs,
                        const GridDev *grids, int max_iv, int order, hipStream_t s);
// The same in two halves: worklist_start enqueues the stripe intervals, their
// runs and the scan of the run lengths, with the list's length read back into
// pinned memory; worklist_finish waits for that length, sizes the list and
// enqueues its emission (and the Hilbert order).  Work the caller enqueues on
// other streams in between runs beside the searches.
// With defer_length, worklist_finish does not wait: the list's length stays
// on the device (WorkList::npairs_dev) and the list holds up to a capacity
// sized from the tile counts (row order only; the caller reads the length
// back with its results and calls worklist_regrow when it exceeded the
// capacity).
struct WorkListPending {
    WorkList wl;
    hipStream_t s = nullptr;
    size_t ntq = 0, ntt = 0;
    int max_iv = 0, order = 0;
    uint64_t *off = nullptr;        // device: exclusive scan of the run lengths
    const uint64_t *hnp = nullptr;  // pinned: the list's length
    uint64_t cap_hint = 0;          // deferred length: capacity to provide
    bool pending = false;
    bool defer = false;
};
WorkListPending worklist_start(Ctx &c, const Side &Q, const Side &T, const SegDev *segs,
                               const GridDev *grids, int max_iv, int order, hipStream_t s,
                               bool defer_length = false);
WorkList worklist_finish(Ctx &c, WorkListPending &p);
// The list again with room for np entries (np: the length read back), on
// stream s.
void worklist_regrow(Ctx &c, WorkList &wl, uint64_t np, hipStream_t s);

// Result of one segmented join.
struct JoinOut {
    uint64_t count = 0;
    uint64_t counters[8] = {};
    std::vector<uint64_t> seg_counts;
    WorkList wl;
    float kernel_ms = 0;
    // the operand kernels found a point outside the sampled ranges: the
    // filter counted nothing, and the join runs again from exact ranges
    bool ranges_missed = false;
    int sampled_ranges = 0;  // hj_stats.sampled_ranges
};

// Runs the count kernel (and the filter's band recheck) over a work list.
// segs: the segment table on the device, segs_host its host copy.  With
// opflag (the filter's operand range flag) it returns at the first read-back
// with R.ranges_missed set when the flag was set.  zeroed: the caller has
// enqueued zero_count_state on the main stream already.
void run_count(Ctx &c, const Side &Q, const Side &T, const SegDev *segs,
               const std::vector<SegDev> &segs_host, const WorkList &wl, const Opts &o,
               const FilterParams &fp, bool self, JoinOut &R, const uint32_t *opflag = nullptr,
               bool zeroed = false);
// The join counters and per-segment counts of run_count, zeroed on stream s;
// returns the counters.
uint64_t *zero_count_state(Ctx &c, uint32_t nseg, hipStream_t s);

// The whole join of a problem built from device rows: grids from the data,
// keys, sort, layouts, work list, count.  need_layout keeps the FP64 layouts
// with a filter kernel too; worklist_only stops before the count.
struct Problem {
    int d = 0;
    bool self = true;
    SideSpec Q, T;
    std::vector<double> eps;  // per join segment (segment k: query k with target k)
};

struct JoinPrep {
    Side Q, T;
    std::vector<Grid> grids;
    std::vector<SegDev> segs_host;
    const SegDev *segs = nullptr;  // join segment table (device)
    FilterParams fp{};
    int kernel = 0;  // HJ_KERNEL_* the join ran with (an index query may fall back to FP64)
    WorkList wl;
};

// Host <-> device copies of large buffers on the context's stream through the
// two pinned staging buffers: the DMA of one 64 MiB chunk overlaps a
// multi-threaded host copy of the other (pageable memory moves at 9-16 GB/s
// through the runtime's own staging).  copy_d2h returns when dst holds the
// data; copy_h2d returns when src may be reused (the DMAs stay stream-ordered).
void copy_d2h(Ctx &c, void *dst, const void *src, size_t bytes);
void copy_h2d(Ctx &c, void *dst, const void *src, size_t bytes);

void run_problem(Ctx &c, const Problem &P, const Opts &o, JoinOut &R, JoinPrep &prep,
                 hj_stats *st, bool need_layout, bool worklist_only = false);

// Pair lists of a prepared join (run_problem or run_index_query): one pass of
// the join's kernel over the work list that appends the accepted pairs (the
// FP64 pair kernel, or the filter with its FP64 band recheck), mapped
// to (global query row, target row - toff of the segment) keys, mirrored for
// self segments, radix sorted, turned into rowptr (Q.n + 1 entries, device
// rows of all segments) and colidx.  segs_emit: the join segment table to use
// for the pass (a segment with eps2 < 0 and t_in = t_out = +inf emits
// nothing under either kernel); np: pairs expected.
// Writes rowptr (always) and colidx (when it fits cap) to the host, or in
// place when dev_out (both then device pointers); returns the number of
// colidx entries.
uint64_t csr_fill(Ctx &c, const JoinPrep &prep, const SegDev *segs_emit, const Opts &o,
                  bool self, uint64_t np, int64_t *rowptr, int64_t *colidx, uint64_t cap,
                  bool dev_out = false);

// ---- resident index ------------------------------------------------------------
// A point set sorted once into HBM under a grid of cell width eps_max; query
// batches (sets of requests, each a query point set with its own eps <=
// eps_max) are joined against it.  Keys of the index use at most
// INDEX_KEY_BITS bits, leaving the high bits for the request prefix.
constexpr int INDEX_KEY_BITS = 48;
constexpr uint32_t INDEX_MAX_REQUESTS = 1u << (64 - INDEX_KEY_BITS);
// filter operands of the index leave 2^INDEX_HEADROOM of room for query points
// beyond the index's own bounds
constexpr int INDEX_HEADROOM = 2;

struct Index {
    int device = 0;
    int d = 0;
    uint64_t n = 0;
    double eps_max = 0;
    int kernel = HJ_KERNEL_FP64;  // filter operands held for this kernel
    Grid grid;
    double pmax = 0;  // largest squared norm of an index point (norm2_bound)
    int local_bits = 0;
    GridTables gt;
    SetBufs B;
    DevBuf pts;  // own copy of the points (n x d)
    Side T;
    FilterParams fp{};
    void release() {
        gt.release();
        B.release();
        pts.release();
    }
};

// Builds ix from device rows pts (copied into the index).
void index_build(Ctx &c, Index &ix, const double *pts, uint64_t n, int d, double eps_max,
                 const Opts &o);

// Joins a batch of requests with the index: request k is rows [Qs.off[k],
// Qs.off[k+1]) of Qs.pts under radius eps[k].  Per-request counts in
// R.seg_counts.  The filter is used when the index holds filter operands, the
// options ask for it and every query lies within the operands' range;
// otherwise the FP64 kernel.
void run_index_query(Ctx &c, const Index &ix, const SideSpec &Qs, const std::vector<double> &eps,
                     const Opts &o, JoinOut &R, JoinPrep &prep, hj_stats *st, bool need_layout);

// Context set-up / tear-down (the C ABI's hj_ctx wraps one).
int ctx_init(Ctx &c, int device);
void ctx_release(Ctx &c);

// Segment-table upload: the join segments followed by extra u64 words (row
// offsets); returns the device address of the words.
uint64_t *upload_segs(Ctx &c, const std::vector<SegDev> &segs, const std::vector<uint64_t> &words);

// Sets the stats fields common to every join.
void fill_stats(hj_stats *st, const JoinOut &R, const JoinPrep &prep, const Opts &o, double denom);

}  // namespace hj
