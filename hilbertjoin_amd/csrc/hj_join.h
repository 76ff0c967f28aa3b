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

#include <vector>

#include "hj_internal.h"

namespace hj {

struct Opts {
    int kernel = HJ_KERNEL_FP64;
    int active = -1;
    int order = HJ_ORDER_ROW;
    int kblock = 8;
    int grid = 0;
    int split = 0;  // filter: first-chunk dimensions (0 auto, -1 no early exit)
};

Opts read_opts(const hj_opts *o);
void check_sizes(uint64_t n, uint32_t d);
int dpad_for(int d);
inline size_t ntiles_of(uint64_t n) { return (size_t)((n + 63) / 64); }

// Grid of one join from the per-dimension bounds of its points: cells of width
// eps, bits per key field, key dimensions within key_budget bits, filter
// centre and scale (the scale leaves 2^headroom room for points beyond the
// bounds: serving queries).
Grid grid_from_bounds(const double *lo, const double *hi, int d, double eps, int key_budget,
                      int active_req, int headroom);

// Per-segment min / max of each dimension, read back: lo/hi[seg * d + k].
// Throws InvalidArg on a non-finite coordinate.
void seg_bounds(Ctx &c, const double *pts, const std::vector<uint64_t> &off, int d,
                std::vector<double> &lo, std::vector<double> &hi);

// Largest |x - centre| over the points whose bounds are lo/hi (d values each).
double radius_bound(const double *lo, const double *hi, const std::vector<double> &centre);

// Filter layout of the K axis (split, k-steps, norm parts) for d dimensions.
FilterParams filter_layout(int d, bool bf16, int split_req);
// Classification thresholds for radius eps with every scaled coordinate
// vector within M of the centre (M = scale * radius bound).
void filter_thresholds(const FilterParams &fp, int d, double scale, double M, double eps,
                       float *t_in, float *t_out);

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
    const double *rows = nullptr, *cols = nullptr;
    const void *frow = nullptr, *fcol = nullptr;
    const uint64_t *tlo = nullptr, *thi = nullptr;
    const uint32_t *tile_seg = nullptr;  // layout tile -> entry of the side's segment table
};

// Which layouts a side needs.
struct SideNeeds {
    bool fp64_rows = false, fp64_cols = false;  // FP64 layouts
    bool frag = false;                          // filter fragments (both roles)
    bool bounds = false;                        // per-tile key bounds (query side)
};

// Tile layout of a side: first tile per segment and the total.
void side_layout(Side &S, const std::vector<uint64_t> &off, const double *pts, int d);

// Keys of one side (identity permutation).  seg_table: the side's segment
// table (one entry per segment of the side, device), seg_off_dev: its row
// offsets on the device.
void side_keys(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table, uint32_t nseg,
               const uint64_t *seg_off_dev, bool target, const GridTables &gt, int kd_max);
// LSD radix sort of the side's keys over their low sort_bits bits.
void side_sort(Ctx &c, SetBufs &B, Side &S, int sort_bits);
// Layouts of a sorted side; tile_owner[t]: the segment-table entry of tile t.
void side_layouts(Ctx &c, SetBufs &B, Side &S, const SegDev *seg_table,
                  const std::vector<uint32_t> &tile_owner, bool target, const GridTables &gt,
                  const FilterParams &fp, const SideNeeds &need);
// tile_owner of a side whose segment k owns its tiles tile0[k] ..
std::vector<uint32_t> tile_owners(const Side &S);

struct WorkList {
    const uint64_t *pairs = nullptr;
    uint64_t npairs = 0;
};

WorkList build_worklist(Ctx &c, const Side &Q, const Side &T, const SegDev *segs,
                        const GridDev *grids, int max_iv, int order);

// Result of one segmented join.
struct JoinOut {
    uint64_t count = 0;
    uint64_t counters[8] = {};
    std::vector<uint64_t> seg_counts;
    WorkList wl;
    float kernel_ms = 0;
};

// Runs the count kernel (and the filter's band recheck) over a work list.
void run_count(Ctx &c, const Side &Q, const Side &T, const SegDev *segs, uint32_t nseg,
               const WorkList &wl, const Opts &o, const FilterParams &fp, bool self, JoinOut &R);

// The whole join of a problem built from device rows: grids from the data,
// keys, sort, layouts, work list, count.  need_layout keeps the FP64 layouts
// (pair output); worklist_only stops before the count.
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
    WorkList wl;
};

void run_problem(Ctx &c, const Problem &P, const Opts &o, JoinOut &R, JoinPrep &prep,
                 hj_stats *st, bool need_layout, bool worklist_only = false);

// Segment-table upload: the join segments followed by extra u64 words (row
// offsets); returns the device address of the words.
uint64_t *upload_segs(Ctx &c, const std::vector<SegDev> &segs, const std::vector<uint64_t> &words);

// Sets the stats fields common to every join.
void fill_stats(hj_stats *st, const JoinOut &R, const JoinPrep &prep, const Opts &o, double denom);

}  // namespace hj
