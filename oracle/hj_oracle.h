/*
 * hj_oracle.h -- CPU oracle for the epsilon similarity join.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for the cpu_baseline leg, as the timed CPU restatement).  The product
 * library (hilbertjoin_amd) never links, loads or calls it.
 *
 * What it restates (reference = perdacherMartin/hilbertJoin):
 *   - grid cell of a coordinate, floor(x * (1/eps)) -- egojoin.cpp:9-17
 *     (epsilonGridCompare; under -ffast-math the division is a multiply by
 *     the reciprocal, so the cell rule is floor(x * (1.0/eps))).
 *   - lexicographic cell comparison -- egojoin.cpp:9-17.
 *   - the join predicate ||p-q||_2 <= eps, counted over unordered pairs i<j
 *     (self-join, egojoin.cpp:777-838 with the diagonal rule egojoin.h:383-392)
 *     or over all (a,b) pairs (A x B join, egojoin.cpp:840-907).
 *   - the reference's own tile arithmetic (expanded form
 *     eps^2/2 - |a|^2/2 - |b|^2/2 + a.b >= 0 with per-KBLOCK partial norms,
 *     egojoin.h:182-195 and :329-381, sign rule egojoin.cpp:803-811), used
 *     only to list the pairs on which the reference's arithmetic and the
 *     fixed-order oracle can disagree (pairs within rounding of eps).
 *   - an exact CPU EGO join (sort by cells, per-point candidate ranges, 8x8
 *     blocks with an early exit every KBLOCK dimensions) restating the
 *     reference's algorithm (egojoin.cpp:159-168, :270-464, :1126-1240,
 *     egojoin.h:164-222, :311-392) -- the CPU baseline.
 *   - the textbook Hilbert curve index (xy2d / d2xy), the order in which the
 *     reference's traversal (hilloop.h:9-82) visits block pairs.
 *
 * The FIXED-ORDER FP64 PREDICATE (the count every product path must match
 * bit-for-bit):
 *     s = 0;  for k = 0..d-1 { t = a[k] - b[k];  s = fma(t, t, s); }
 *     pair counts  <=>  s <= eps*eps        (eps*eps rounded once, in FP64)
 * Partial sums are non-decreasing under round-to-nearest, so an early exit
 * once s > eps*eps never changes the result.
 *
 * Parity pins (tests/test_oracle.py): closed forms (n identical points,
 * integer lattices, separated clusters), invariants (permutations of points
 * and of dimensions, self = (join(X,X) - n)/2, join(A,B) = join(B,A),
 * monotonicity in eps), an independent numpy brute force on data without
 * boundary pairs, the EGO port against brute force, and the Hilbert round
 * trip / adjacency properties.  No function here is "parity unpinned".
 */
#ifndef HJ_ORACLE_H
#define HJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* grid cell and lexicographic comparison (egojoin.cpp:9-17) */
double hjo_cell(double x, double eps);
int hjo_ego_compare(const double *a, const double *b, int d, double eps);

/* fixed-order FP64 squared distance and predicate */
double hjo_dist2(const double *a, const double *b, int d);
int hjo_pair_within(const double *a, const double *b, int d, double eps);

/* brute force counts; early_exit != 0 stops a pair once s > eps^2 */
uint64_t hjo_self_count(const double *X, int64_t n, int d, double eps, int early_exit);
uint64_t hjo_join_count(const double *A, int64_t nA, const double *B, int64_t nB,
                        int d, double eps, int early_exit);

/* the reference's expanded-form tile arithmetic, pair by pair */
int hjo_pair_within_expanded(const double *a, const double *b, int d, double eps, int kblock);
uint64_t hjo_self_count_expanded(const double *X, int64_t n, int d, double eps, int kblock);
/* A x B count of the reference's NOSELF expanded form (egojoin.h:482-553) */
uint64_t hjo_join_count_expanded(const double *A, int64_t nA, const double *B, int64_t nB,
                                 int d, double eps, int kblock);
/* pairs (i<j) where the expanded form (evaluated to the end, or stopped at
 * the first check point where its own sum is negative -- the two results a
 * pair can get in the reference's 8x8 tile) and the fixed-order predicate
 * disagree; writes up to cap pairs as (i,j) into out[2*k], returns the total
 * number */
int64_t hjo_self_disagreements(const double *X, int64_t n, int d, double eps, int kblock,
                               int64_t *out, int64_t cap);
/* the same for the pairs (a, b) of an A x B join (egojoin.h:482-553,
 * egojoin.cpp:868-884) */
int64_t hjo_join_disagreements(const double *A, int64_t nA, const double *B, int64_t nB, int d,
                               double eps, int kblock, int64_t *out, int64_t cap);

/* pairs with |s - eps^2| <= tol (fixed-order s); same output convention */
int64_t hjo_self_band(const double *X, int64_t n, int d, double eps, double tol,
                      int64_t *out, int64_t cap);

/* CSR neighbour lists.  Self: row i lists every j != i within eps (both
 * directions), ascending.  Join: row a lists every b within eps, ascending.
 * rowptr has n+1 entries; colidx may be NULL to only fill rowptr. */
uint64_t hjo_self_csr(const double *X, int64_t n, int d, double eps,
                      int64_t *rowptr, int64_t *colidx);
uint64_t hjo_join_csr(const double *A, int64_t nA, const double *B, int64_t nB, int d,
                      double eps, int64_t *rowptr, int64_t *colidx);

/* EGO order: a stable permutation sorting rows by their cell vectors */
void hjo_ego_sort(const double *X, int64_t n, int d, double eps, int64_t *perm);

/* CPU EGO join (the CPU baseline).  threads <= 0 uses all cores.
 * times[0] = sort s, times[1] = index s, times[2] = join loop s (may be NULL).
 * evaluated (may be NULL) receives the number of pair evaluations executed. */
uint64_t hjo_ego_self_count(const double *X, int64_t n, int d, double eps, int active_dims,
                            int kblock, int threads, double *times, uint64_t *evaluated);
uint64_t hjo_ego_join_count(const double *A, int64_t nA, const double *B, int64_t nB, int d,
                            double eps, int active_dims, int kblock, int threads,
                            double *times, uint64_t *evaluated);

/* CPU baseline sample: sort and index in full, join loop on every
 * iblock_stride-th 8-block of the query set only */
uint64_t hjo_ego_self_count_sampled(const double *X, int64_t n, int d, double eps,
                                    int active_dims, int kblock, int threads,
                                    int64_t iblock_stride, double *times, uint64_t *evaluated);

/* Per-dimension selectivity cost of outputStatistics (egojoin.cpp:523-591),
 * computed exactly over all points: cost[k] = pairs of points whose cells
 * floor(x_k * (1/eps)) are equal or adjacent.  Returns 0, or -1 when a
 * dimension spans more than 2^26 cells. */
int hjo_dim_costs(const double *X, int64_t n, int d, double eps, uint64_t *cost);

/* textbook Hilbert curve on a 2^order x 2^order grid */
uint64_t hjo_hilbert_xy2d(int order, uint32_t x, uint32_t y);
void hjo_hilbert_d2xy(int order, uint64_t h, uint32_t *x, uint32_t *y);

int hjo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
