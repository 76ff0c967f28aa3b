/*
 * hilbertjoin_amd.h -- C ABI of the MI355X epsilon similarity join.
 *
 * Plain C types only (pointers, sizes, status codes).  Every entry point
 * returns HJ_OK (0) on success or an HJ_ERR_* code; hj_last_error() then
 * describes the failure.  A context owns one HIP stream on one device and a
 * set of grow-only device buffers; calls on one context are serialised by the
 * caller, calls on different contexts may run concurrently.
 *
 * Counting semantics (identical to the reference's count, minus its defects):
 *   self-join : number of unordered pairs i < j with ||x_i - x_j||_2 <= eps
 *   A x B join: number of pairs (a, b) with ||a - b||_2 <= eps
 * where the distance test is the fixed-order FP64 predicate
 *   s = 0; for k in 0..d-1: t = a_k - b_k; s = fma(t, t, s);  s <= eps*eps.
 * Every point takes part (the reference drops the n % 8 tail after its sort,
 * egojoin.h:325,496), and inputs are never modified (the reference sorts its
 * input in place and restores it incorrectly, egojoin.cpp:476-511).
 */
#ifndef HILBERTJOIN_AMD_H
#define HILBERTJOIN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HJ_API __attribute__((visibility("default")))

enum {
    HJ_OK = 0,
    HJ_ERR_INVALID = 1,  /* bad argument (null pointer, eps <= 0, non-finite input, ...) */
    HJ_ERR_HIP = 2,      /* a HIP runtime call failed */
    HJ_ERR_NOMEM = 3,    /* device or host allocation failed */
    HJ_ERR_CAPACITY = 4, /* an output buffer is too small; the required size is reported */
    HJ_ERR_INTERNAL = 5,
    HJ_ERR_TIMEOUT = 6   /* hj_server_wait: the request is not done yet */
};

/* distance kernels */
enum {
    HJ_KERNEL_FP64 = 0,      /* plain FP64 tiles, early exit every kblock dimensions */
    HJ_KERNEL_MFMA_F16 = 1,  /* FP16 MFMA filter (FP32 accumulate) + FP64 band recheck */
    HJ_KERNEL_MFMA_BF16 = 2  /* BF16 MFMA filter (FP32 accumulate) + FP64 band recheck */
};

/* order in which tile pairs are drained */
enum { HJ_ORDER_HILBERT = 0, HJ_ORDER_ROW = 1 };

typedef struct hj_opts {
    int32_t kernel;      /* HJ_KERNEL_* (default FP64) */
    int32_t active_dims; /* active dimensions of the stripe enumeration, 0..6; -1 = chosen from
                            the data: the largest a <= 6 whose a-cell key prefixes hold >= 128
                            target points on average */
    int32_t order;       /* HJ_ORDER_* (default row order) */
    int32_t kblock;      /* FP64 early-exit cadence in dimensions: 4, 8 or 16 (0 = default 4;
                            the reference's compile-time KBLOCK, CMakeLists.txt:63-71) */
    int32_t grid;        /* distance-kernel workgroups; 0 = chosen from the device and the
                            work list */
    int32_t filter_split; /* MFMA filter: dimensions in the first K chunk, after which a
                             64x64 block that is certainly out exits; 0 = auto, -1 = none */
    int32_t tile_rows;   /* FP64 kernel: rows per early-exit vote group, 8 or 16 (0 = 8) */
    int32_t mfma_chunk;  /* MFMA filter: tile pairs whose column operands are loaded ahead
                            together, 2, 4 or 8 (0 = 4) */
    int32_t tune;        /* 1: use the context's autotuned settings for inputs of this shape
                            class, tuning on this input first if there are none (see
                            hj_autotune); the fields above are the starting point */
    int32_t filter_frames; /* MFMA filter: 0 = tiles whose points lie close together take
                              their own operand frame, which narrows the error band to the
                              tiles' own extent (d <= 64); -1 = one frame per join */
    int32_t reserved[2];
} hj_opts;

typedef struct hj_stats {
    double total_ms;          /* wall time of the call (host clock, after the final sync) */
    double grid_ms;           /* coordinate ranges + grid keys            (GPU events) */
    double sort_ms;           /* LSD radix sort of the keys              (GPU events) */
    double index_ms;          /* gather, tile bounds, work list, ordering (GPU events) */
    double join_ms;           /* distance kernels (+ band recheck)        (GPU events) */
    double kernel_ms;         /* the distance kernel's launch alone       (GPU events) */
    uint64_t tile_pairs;      /* 64 x 64 tile pairs in the work list */
    uint64_t candidate_pairs; /* point pairs inside those tiles (self: i<j on the diagonal) */
    uint64_t pair_dims;       /* FP64 kernel: pair-dimensions evaluated (incl. masked lanes);
                                 MFMA filter: K slots multiplied per pair of points, summed
                                 (16 x 32 x 32 per MFMA issued) */
    uint64_t band_pairs;      /* MFMA filter: pairs rechecked in FP64 */
    double load_fraction;     /* candidate_pairs / (n(n-1)/2) or / (nA*nB) */
    int32_t key_dims;         /* leading dimensions packed in the 64-bit grid key */
    int32_t key_bits;         /* bits per packed dimension */
    int32_t active_dims;      /* active dimensions used */
    int32_t kernel;           /* kernel used */
    uint64_t early_exits;     /* MFMA filter: tile pairs decided after the first K chunk */
    int32_t frames;           /* MFMA filter: 0 = every tile in the join's frame; 1 = some
                                 tiles in their own frames, used by the slow path; 2 = by
                                 the fast path too (hj_opts.filter_frames) */
    int32_t sampled_ranges;   /* MFMA filter, d > 16: 1 = the operand ranges of the dimensions
                                 past the 16th came from a sample of the rows that every point
                                 met; 2 = a point fell outside it and the join ran again from
                                 exact ranges; 0 = exact ranges from the start */
} hj_stats;

typedef struct hj_ctx hj_ctx;

/* Fills *o with the defaults (FP64 kernel, active dims from the data, row order). */
HJ_API void hj_opts_default(hj_opts *o);

/* Creates a context on HIP device `device` (one stream of default priority).
 * Replaces the reference's process-global state (EGO_epsilon, EGO_d,
 * min_size_qsort, egojoin.cpp:3-6) with per-context state. */
HJ_API int hj_ctx_create(hj_ctx **out, int device);
HJ_API int hj_ctx_destroy(hj_ctx *ctx);
HJ_API const char *hj_last_error(const hj_ctx *ctx);
/* number of compute units and device name (name may be NULL) */
HJ_API int hj_device_info(hj_ctx *ctx, int *num_cus, char *name, int name_len);

/* Device buffers on the context's device, for callers that keep point sets
 * resident in HBM and pass them to the *_device entry points.  No counterpart
 * in the reference (it has no device).  Copies are ordered on the context's
 * stream and complete before the call returns. */
HJ_API int hj_device_alloc(hj_ctx *ctx, uint64_t bytes, void **ptr);
HJ_API int hj_device_free(hj_ctx *ctx, void *ptr);
HJ_API int hj_memcpy_to_device(hj_ctx *ctx, void *dst_dev, const void *src_host, uint64_t bytes);
HJ_API int hj_memcpy_to_host(hj_ctx *ctx, void *dst_host, const void *src_dev, uint64_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte loads and
 * stores); bytes must be a multiple of 16.  Also the reference point for
 * calibrating the HBM byte counters of the join kernels. */
HJ_API int hj_device_copy(hj_ctx *ctx, void *dst_dev, const void *src_dev, uint64_t bytes);

/* Self-join count.  Replaces test_ego_loop3_macro (egojoin.cpp:777-838,
 * declared egojoin.h:56): same n, d, epsilon, array and count; the stripe
 * count becomes opts->active_dims, sortTime/indextime/loadpercent come back
 * in *stats.  `pts` is a host pointer to n x d row-major FP64, read only.
 * stats may be NULL. */
HJ_API int hj_self_join_count(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d,
                              double eps, const hj_opts *opts, uint64_t *count,
                              hj_stats *stats);

/* A x B join count.  Replaces test_ego_loop3_noself (egojoin.cpp:840-907,
 * declared egojoin.h:64); activedims is honoured (the reference hard-codes 2,
 * egojoin.cpp:847).  Host pointers, read only. */
HJ_API int hj_join_count(hj_ctx *ctx, const double *A, uint64_t nA, const double *B,
                         uint64_t nB, uint32_t d, double eps, const hj_opts *opts,
                         uint64_t *count, hj_stats *stats);

/* Same two joins on device-resident inputs (device pointers to n x d
 * row-major FP64), ordered on the context's stream.  No counterpart in the
 * reference (it has no device); these are the entry points for callers that
 * keep the points in HBM. */
HJ_API int hj_self_join_count_device(hj_ctx *ctx, const double *pts_dev, uint64_t n,
                                     uint32_t d, double eps, const hj_opts *opts,
                                     uint64_t *count, hj_stats *stats);
HJ_API int hj_join_count_device(hj_ctx *ctx, const double *A_dev, uint64_t nA,
                                const double *B_dev, uint64_t nB, uint32_t d, double eps,
                                const hj_opts *opts, uint64_t *count, hj_stats *stats);

/* The self-join count of hj_self_join_count_device without any allocation:
 * every internal device buffer comes from a fixed scratch array of
 * HJ_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends, every host staging buffer from a fixed host array;
 * no hipMalloc, hipHostMalloc or hipFree is called.  No counterpart in the
 * reference; this is the entry point for callers that cannot allocate (code
 * under graph capture, tight serving loops, harnesses that forbid it).  A join
 * that needs more scratch than that (about n * (24 + 8 * d) bytes for the FP64
 * kernel plus its work list) fails with HJ_ERR_CAPACITY and a message; opts->tune
 * must be 0.  All calls on one GPU share its scratch array: calls that overlap
 * on one GPU (other threads, other contexts) race and must be serialized by
 * the caller.  Same count and statistics as hj_self_join_count_device. */
#define HJ_NOALLOC_SCRATCH_BYTES ((uint64_t)256 << 20)
HJ_API int hj_self_join_count_device_noalloc(hj_ctx *ctx, const double *pts_dev, uint64_t n,
                                             uint32_t d, double eps, const hj_opts *opts,
                                             uint64_t *count, hj_stats *stats);

/* Batched joins in one pipeline pass and one persistent kernel launch.  Job
 * k self-joins rows [offsets[k], offsets[k+1]) of pts (n = offsets[njobs]
 * rows of d FP64 values) under radius eps[k]; counts[k] receives its count,
 * identical to hj_self_join_count on those rows alone.  Each job gets its own
 * grid (cells of width eps[k] over its own coordinate range) and its own
 * filter thresholds; the keys carry the job in their high bits, so one sort,
 * one work list and one kernel launch cover all jobs.  No counterpart in the
 * reference, which runs one join per process (main.cpp:102-124); this is the
 * entry point for callers with many small joins.  njobs <= 2^20; stats
 * aggregate over all jobs. */
HJ_API int hj_batch_self_join_count(hj_ctx *ctx, const double *pts, const uint64_t *offsets,
                                    uint32_t njobs, uint32_t d, const double *eps,
                                    const hj_opts *opts, uint64_t *counts, hj_stats *stats);
/* Batched A x B joins: job k joins rows [a_offsets[k], a_offsets[k+1]) of A
 * with rows [b_offsets[k], b_offsets[k+1]) of B. */
HJ_API int hj_batch_join_count(hj_ctx *ctx, const double *A, const uint64_t *a_offsets,
                               const double *B, const uint64_t *b_offsets, uint32_t njobs,
                               uint32_t d, const double *eps, const hj_opts *opts,
                               uint64_t *counts, hj_stats *stats);
/* The same on device-resident points (offsets, eps and counts on the host). */
HJ_API int hj_batch_self_join_count_device(hj_ctx *ctx, const double *pts_dev,
                                           const uint64_t *offsets, uint32_t njobs, uint32_t d,
                                           const double *eps, const hj_opts *opts,
                                           uint64_t *counts, hj_stats *stats);
HJ_API int hj_batch_join_count_device(hj_ctx *ctx, const double *A_dev,
                                      const uint64_t *a_offsets, const double *B_dev,
                                      const uint64_t *b_offsets, uint32_t njobs, uint32_t d,
                                      const double *eps, const hj_opts *opts, uint64_t *counts,
                                      hj_stats *stats);

/* Autotuner.  Times candidate settings of one join (or batch) on the given
 * device-resident input and returns the fastest (whole-call wall time, best of
 * two runs per candidate) by coordinate descent over: active dimensions
 * (single joins: the data-driven choice and its neighbours), grid size
 * (filter: the default, 2, 4, 8 workgroups per CU; FP64: the default, 4, 64,
 * 256 per CU), the kernel's tile shape (FP64:
 * tile_rows 8 / 16 x kblock 4 / 8 / 16; filter: mfma_chunk 2 / 4 / 8 and
 * filter_split auto / none) and the work-list order.  The result is also
 * cached in the context under the input's shape class (kernel, self / A x B,
 * single / batch, d, floor(log2 n), floor(4 log2 eps)), where calls with
 * opts->tune = 1 find it.  No counterpart in the reference, whose tile shape
 * and stripes are compile-time or command-line constants (KBLOCK,
 * NUM_THREADS, -s). */
typedef struct hj_tune_report {
    int32_t trials;   /* settings timed */
    int32_t reserved;
    double base_ms;   /* the starting settings */
    double best_ms;   /* the settings returned */
    hj_opts best;
} hj_tune_report;

HJ_API int hj_autotune_self(hj_ctx *ctx, const double *pts_dev, uint64_t n, uint32_t d,
                            double eps, const hj_opts *base, hj_tune_report *rep);
HJ_API int hj_autotune_join(hj_ctx *ctx, const double *A_dev, uint64_t nA, const double *B_dev,
                            uint64_t nB, uint32_t d, double eps, const hj_opts *base,
                            hj_tune_report *rep);
HJ_API int hj_autotune_batch_self(hj_ctx *ctx, const double *pts_dev, const uint64_t *offsets,
                                  uint32_t njobs, uint32_t d, const double *eps,
                                  const hj_opts *base, hj_tune_report *rep);

/* Self-join pair lists in CSR form over the ORIGINAL point indices: row i
 * lists every j != i with ||x_i - x_j|| <= eps, ascending (both directions of
 * each unordered pair).  rowptr has n+1 entries.  colidx has room for cap
 * entries; if 2*count > cap, HJ_ERR_CAPACITY is returned, rowptr is filled
 * and *total holds the required size.  The reference's pair output exists
 * only as dead code (join_pair egojoin.h:32-35, test_ego_loop3_queue.cpp). */
HJ_API int hj_self_join_csr(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d, double eps,
                            const hj_opts *opts, int64_t *rowptr, int64_t *colidx,
                            uint64_t cap, uint64_t *total, hj_stats *stats);
/* A x B pair lists: row a lists every b within eps, ascending. */
HJ_API int hj_join_csr(hj_ctx *ctx, const double *A, uint64_t nA, const double *B, uint64_t nB,
                       uint32_t d, double eps, const hj_opts *opts, int64_t *rowptr,
                       int64_t *colidx, uint64_t cap, uint64_t *total, hj_stats *stats);
/* The same pair lists with points, rowptr and colidx all in device memory of
 * the context's device (the lists are written in place and stay in HBM, no
 * copy over PCIe); same capacity rule.  No counterpart in the reference,
 * whose pair output is dead code. */
HJ_API int hj_self_join_csr_device(hj_ctx *ctx, const double *pts_dev, uint64_t n, uint32_t d,
                                   double eps, const hj_opts *opts, int64_t *rowptr_dev,
                                   int64_t *colidx_dev, uint64_t cap, uint64_t *total,
                                   hj_stats *stats);
HJ_API int hj_join_csr_device(hj_ctx *ctx, const double *A_dev, uint64_t nA, const double *B_dev,
                              uint64_t nB, uint32_t d, double eps, const hj_opts *opts,
                              int64_t *rowptr_dev, int64_t *colidx_dev, uint64_t cap,
                              uint64_t *total, hj_stats *stats);

/* EGO order of a point set (the reference's epsilonGridOrdering,
 * egojoin.cpp:159-168, comparator :9-17): perm[k] is the index of the k-th
 * point in ascending order of its grid cells floor(x * (1/eps)) compared
 * lexicographically over the first *key_dims dimensions (all d when they fit
 * the 64-bit key), ties in index order.  keys (may be NULL) receives the
 * sorted 64-bit keys. */
HJ_API int hj_ego_sort(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d, double eps,
                       uint32_t *perm, uint64_t *keys, int32_t *key_dims, int32_t *key_bits);

/* The tile-pair work list of a self-join (I, J in units of 64 sorted points,
 * J >= I) in drain order, as pairs[2k] = I, pairs[2k+1] = J; perm receives
 * the EGO permutation (may be NULL).  Replaces the stripe bounds and the
 * Hilbert traversal (prepareStripes egojoin.cpp:270-464, FGF_HILBERT_FOR
 * hilloop.h:9-82).  Returns HJ_ERR_CAPACITY with *npairs set if cap is too
 * small. */
HJ_API int hj_self_tile_pairs(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d,
                              double eps, const hj_opts *opts, uint32_t *pairs, uint64_t cap,
                              uint64_t *npairs, uint32_t *perm);
/* Same for an A x B join (I over A tiles, J over B tiles). */
HJ_API int hj_join_tile_pairs(hj_ctx *ctx, const double *A, uint64_t nA, const double *B,
                              uint64_t nB, uint32_t d, double eps, const hj_opts *opts,
                              uint32_t *pairs, uint64_t cap, uint64_t *npairs, uint32_t *permA,
                              uint32_t *permB);

/* Dimension statistics (outputStatistics, egojoin.cpp:523-591): cost[k] =
 * number of point pairs whose grid cells floor(x_k * (1/eps)) in dimension k
 * are equal or adjacent (exact over all points; the reference estimates it
 * from a random sample), and reorder_dim[k] = position of dimension k when
 * the dimensions are sorted by increasing cost (the most selective first;
 * ties keep the input order).  reorder_dim and cost (may be NULL) have d
 * entries.  Host points, read only. */
HJ_API int hj_dim_statistics(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d,
                             double eps, int32_t *reorder_dim, uint64_t *cost);
/* Moves column k of every row of the n x d host array to column
 * reorder_dim[k], in place (reorder_dimensions, egojoin.cpp:663-732, correct
 * for every d; the reference's mask leaves the last d % 8 columns unmoved).
 * The join's fixed-order distance then sums the dimensions in the new order. */
HJ_API int hj_reorder_dimensions(hj_ctx *ctx, double *pts, uint64_t n, uint32_t d,
                                 const int32_t *reorder_dim);

/* The reduced-precision filter's operands and error band for one join,
 * computed on the host (no device work).  A join with points within `radius`
 * of the operand centre c scales y = (x - c) * scale (a power of two that keeps
 * every |y| <= 2^14), rounds y to FP16 / BF16 and splits each squared norm
 * |y|^2 / 2^16 into `parts` FP16 / BF16 values; the MFMA chain then yields
 * acc = -|a - b|^2 * scale^2 / 2 up to rounding, and a pair is
 *   certainly outside eps  when acc <  t_out,
 *   certainly within eps   when acc >= t_in,
 * and rechecked with the fixed-order FP64 predicate otherwise.  K layout:
 * dimensions [0, split) and their norm slots fill k-steps [0, steps1) of 16
 * slots, the rest fill [steps1, steps).  No counterpart in the reference, whose
 * tile computes in FP64 only (egojoin.h:311-392); this is the entry point for
 * callers and tests that check the filter's error band.  kernel is
 * HJ_KERNEL_MFMA_F16 or HJ_KERNEL_MFMA_BF16; filter_split as in hj_opts. */
typedef struct hj_filter_info {
    double scale;
    float t_in, t_out;
    int32_t split, steps1, steps, parts;
} hj_filter_info;

HJ_API int hj_filter_band(int32_t kernel, uint32_t d, double eps, double radius,
                          int32_t filter_split, hj_filter_info *out);

/* The thresholds of one tile pair when tiles take their own operand frames
 * (hj_opts.filter_frames; no counterpart in the reference, for callers and
 * tests that check the filter's error band): kernel, d, eps, radius and
 * filter_split as in hj_filter_band (radius bounds |x - c| over the join, the
 * operands are scaled by that call's scale); rho_a, rho_b bound the norms of
 * the two tiles' rounded operand vectors, each in its tile's frame; dd_bound
 * bounds and dd_shift is the computed squared norm of the anchors' difference
 * over the K slots tested.  A pair whose accumulated value is below *t_out
 * is certainly farther than eps, one at or above *t_in certainly within. */
HJ_API int hj_filter_pair_band(int32_t kernel, uint32_t d, double eps, double radius,
                               int32_t filter_split, float rho_a, float rho_b, float dd_bound,
                               float dd_shift, float *t_in, float *t_out);

/* Hilbert index of (x[k], y[k]) on a 2^order x 2^order grid, computed on the
 * device (the order of the reference's HILLOOP traversal, hilloop.h:9-82). */
HJ_API int hj_hilbert_index(hj_ctx *ctx, int order, const uint32_t *x, const uint32_t *y,
                            uint64_t n, uint64_t *out);

/* ---- resident index and serving -------------------------------------------
 * No counterpart in the reference, which answers one join per process
 * (main.cpp:102-124) and keeps nothing between calls.  An index holds a
 * point set EGO-sorted once in HBM under a grid of cell width eps_max, with
 * its FP64 tile layout and (for a filter kernel) its MFMA operands; query
 * batches are joined against it without touching it again.  Request k of a
 * batch is the query point set rows [q_off[k], q_off[k+1]) of Q with radius
 * eps[k] <= eps_max; its count is the number of (query, index point) pairs
 * within eps[k] (the A x B count of hj_join_count with A = the request,
 * B = the index).  A batch holds at most 65536 requests. */
typedef struct hj_index hj_index;

/* Builds an index from n x d host (or, _device, device) points on the
 * context's device.  opts->kernel selects the filter operands kept (FP64 keeps
 * none); opts->active_dims and filter_split as for the joins.  The index is
 * independent of the context afterwards and may be shared by any contexts on
 * the same device (read only). */
HJ_API int hj_index_create(hj_ctx *ctx, const double *pts, uint64_t n, uint32_t d,
                           double eps_max, const hj_opts *opts, hj_index **out);
HJ_API int hj_index_create_device(hj_ctx *ctx, const double *pts_dev, uint64_t n, uint32_t d,
                                  double eps_max, const hj_opts *opts, hj_index **out);
HJ_API int hj_index_destroy(hj_index *idx);
/* n, d, eps_max, key dimensions / bits and active dimensions of the index */
HJ_API int hj_index_info(const hj_index *idx, uint64_t *n, uint32_t *d, double *eps_max,
                         int32_t *key_dims, int32_t *key_bits, int32_t *active_dims);

/* One batch of count requests against the index (host or device query
 * points; offsets, eps and counts on the host).  opts->kernel = FP64 forces
 * the FP64 kernel; otherwise the index's kernel is used, except that a batch
 * with a query point beyond the filter operands' range (4x the index's own
 * radius around its centre) falls back to FP64 -- counts are the same. */
HJ_API int hj_index_query_count(hj_ctx *ctx, const hj_index *idx, const double *Q,
                                const uint64_t *q_off, uint32_t nreq, const double *eps,
                                const hj_opts *opts, uint64_t *counts, hj_stats *stats);
HJ_API int hj_index_query_count_device(hj_ctx *ctx, const hj_index *idx, const double *Q_dev,
                                       const uint64_t *q_off, uint32_t nreq, const double *eps,
                                       const hj_opts *opts, uint64_t *counts, hj_stats *stats);
/* Pair lists of a batch: rowptr has q_off[nreq] + 1 entries over all query
 * rows of the batch (row i of request k is rowptr index q_off[k] + i);
 * colidx lists index point ids, ascending per row.  HJ_ERR_CAPACITY with
 * *total set when colidx (cap entries) is too small; rowptr is filled. */
HJ_API int hj_index_query_csr(hj_ctx *ctx, const hj_index *idx, const double *Q,
                              const uint64_t *q_off, uint32_t nreq, const double *eps,
                              const hj_opts *opts, int64_t *rowptr, int64_t *colidx,
                              uint64_t cap, uint64_t *total, hj_stats *stats);

/* Serving loop over a resident index: requests are submitted one by one from
 * any thread; worker threads (each with its own stream) form batches from the
 * waiting requests in arrival order, so requests that arrive while other
 * batches run on the device are admitted as soon as a worker and result
 * memory are free, and the workers' batches overlap on the device.
 * Admission is by result-buffer memory: a count request reserves 8 bytes, a
 * pair-list request (nq + 1) * 8 bytes for its rowptr, then count * 8 bytes
 * for its colidx once its count is known (a request whose pairs do not fit
 * waits and runs again); results hold their memory until released. */
typedef struct hj_server hj_server;

typedef struct hj_server_opts {
    int32_t workers;             /* concurrent batches (default 2) */
    int32_t kernel;              /* -1 (default): the index's kernel; HJ_KERNEL_FP64: FP64 */
    uint32_t max_batch_requests; /* default 4096, at most 65536 */
    uint32_t reserved0;
    uint64_t max_batch_points;   /* query points per batch (default 1 << 18); a larger
                                    request runs in a batch of its own */
    uint64_t result_bytes;       /* result-buffer budget (default 1 GiB) */
    int32_t order;               /* HJ_ORDER_* of the batches' work lists */
    int32_t reserved1[3];
} hj_server_opts;

typedef struct hj_result {
    uint64_t ticket;
    int32_t status;          /* HJ_OK or the HJ_ERR_* code of the failure */
    int32_t pairs;           /* 1: rowptr / colidx are filled */
    uint64_t nq;             /* query points */
    uint64_t count;          /* pairs within eps */
    const int64_t *rowptr;   /* nq + 1 entries (pairs requests), owned by the server */
    const int64_t *colidx;   /* count entries, index point ids, ascending per row */
    double queue_ms;         /* submit -> batch start */
    double run_ms;           /* batch start -> result ready */
    double latency_ms;       /* submit -> result ready */
    uint64_t batch;          /* batch id */
    uint32_t batch_requests; /* requests in that batch */
    uint32_t reserved;
} hj_result;

typedef struct hj_server_stats {
    uint64_t submitted, completed, failed, batches, pending;
    double p50_ms, p99_ms, mean_ms, max_ms; /* latency of completed requests */
    double requests_per_s;                  /* completed / (last completion - first submit) */
    double mean_batch_requests;
    uint64_t result_bytes_used;
} hj_server_stats;

HJ_API void hj_server_opts_default(hj_server_opts *o);
/* The index must outlive the server.  Returns once every worker has set up
 * its stream; HJ_ERR_HIP when none could. */
HJ_API int hj_server_create(const hj_index *idx, int device, const hj_server_opts *opts,
                            hj_server **out);
/* Copies the nq x d query points; returns at once with *ticket. */
HJ_API int hj_server_submit(hj_server *srv, const double *q, uint64_t nq, double eps,
                            int want_pairs, uint64_t *ticket);
/* Blocks until the request is done (timeout_ms < 0: no limit; HJ_ERR_TIMEOUT
 * when it is not done in time) and fills *res. */
HJ_API int hj_server_wait(hj_server *srv, uint64_t ticket, double timeout_ms, hj_result *res);
/* Frees the result's buffers and returns its memory to the admission budget. */
HJ_API int hj_server_release(hj_server *srv, uint64_t ticket);
HJ_API int hj_server_stats_get(hj_server *srv, hj_server_stats *st);
/* Clears the latency statistics. */
HJ_API int hj_server_stats_reset(hj_server *srv);
/* Finishes the submitted requests, stops the workers and frees everything. */
HJ_API int hj_server_destroy(hj_server *srv);
/* A copy of the server's last error message, valid until the calling
 * thread's next call of this function. */
HJ_API const char *hj_server_last_error(const hj_server *srv);

/* Reads n x d FP64 points: `.bin` = raw row-major doubles, anything else =
 * text with ',' separators and '%' comment lines.  Replaces read_file
 * (dataIo.cpp:113-165); returns HJ_ERR_INVALID (instead of exit(1)) on a
 * missing or short file. */
HJ_API int hj_read_points(const char *path, uint64_t n, uint32_t d, double *out);

/* Seeded synthetic data.  Uniform [0,1): element k of seed s is
 * (mix(mix(s) + (k+1) * 0x9E3779B97F4A7C15) >> 11) * 2^-53, mix = SplitMix64
 * finaliser, k = i*d + j.  Replaces the MKL generator random_init_unif
 * (dataIo.cpp:4-33), whose stream cannot be reproduced without MKL. */
HJ_API void hj_gen_uniform(uint64_t seed, uint64_t n, uint32_t d, double *out);
/* Gaussian mixture: ncomp centres uniform in [0,1)^d, isotropic std sigma,
 * weights Dirichlet(alpha), coordinates clipped to [0, nextafter(1,0)]. */
HJ_API void hj_gen_mixture(uint64_t seed, uint64_t n, uint32_t d, uint32_t ncomp, double sigma,
                           double alpha, double *out);

#ifdef __cplusplus
}
#endif
#endif
