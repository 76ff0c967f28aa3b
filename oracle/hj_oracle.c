/*
 * hj_oracle.c -- CPU oracle for the epsilon similarity join.
 *
 * TEST INFRASTRUCTURE ONLY (see hj_oracle.h for who may use it and for the
 * reference file:line each function restates).  Plain C99 + OpenMP, built
 * with -ffp-contract=off so every floating-point operation is the one
 * written here; fused multiply-adds are explicit fma() calls.
 */
#include "hj_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

int hjo_max_threads(void) { return omp_get_max_threads(); }

static double now_s(void) { return omp_get_wtime(); }

/* ------------------------------------------------------------------------ */
/* grid cells (egojoin.cpp:9-17)                                             */
/* ------------------------------------------------------------------------ */

/* egojoin.cpp:12 computes floor(A[i] / EGO_epsilon); the reference is built
 * with -ffast-math (CMakeLists.txt:24), under which g++ hoists 1/eps and
 * multiplies, so the cell of x is floor(x * (1.0/eps)). */
double hjo_cell(double x, double eps) {
    double r = 1.0 / eps;
    return floor(x * r);
}

/* egojoin.cpp:9-17: first non-zero cell difference decides. */
int hjo_ego_compare(const double *a, const double *b, int d, double eps) {
    double r = 1.0 / eps;
    for (int k = 0; k < d; k++) {
        double ca = floor(a[k] * r), cb = floor(b[k] * r);
        if (ca < cb) return -1;
        if (ca > cb) return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* fixed-order FP64 predicate                                                */
/* ------------------------------------------------------------------------ */

double hjo_dist2(const double *a, const double *b, int d) {
    double s = 0.0;
    for (int k = 0; k < d; k++) {
        double t = a[k] - b[k];
        s = fma(t, t, s);
    }
    return s;
}

static inline int within_ee(const double *a, const double *b, int d, double eps2) {
    double s = 0.0;
    for (int k = 0; k < d; k++) {
        double t = a[k] - b[k];
        s = fma(t, t, s);
        if (s > eps2) return 0; /* partial sums never decrease */
    }
    return 1;
}

int hjo_pair_within(const double *a, const double *b, int d, double eps) {
    double eps2 = eps * eps;
    return hjo_dist2(a, b, d) <= eps2;
}

uint64_t hjo_self_count(const double *X, int64_t n, int d, double eps, int early_exit) {
    double eps2 = eps * eps;
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t i = 0; i < n; i++) {
        const double *a = X + i * d;
        uint64_t c = 0;
        for (int64_t j = i + 1; j < n; j++) {
            const double *b = X + j * d;
            if (early_exit)
                c += (uint64_t)within_ee(a, b, d, eps2);
            else
                c += (uint64_t)(hjo_dist2(a, b, d) <= eps2);
        }
        total += c;
    }
    return total;
}

uint64_t hjo_join_count(const double *A, int64_t nA, const double *B, int64_t nB,
                        int d, double eps, int early_exit) {
    double eps2 = eps * eps;
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t i = 0; i < nA; i++) {
        const double *a = A + i * d;
        uint64_t c = 0;
        for (int64_t j = 0; j < nB; j++) {
            const double *b = B + j * d;
            if (early_exit)
                c += (uint64_t)within_ee(a, b, d, eps2);
            else
                c += (uint64_t)(hjo_dist2(a, b, d) <= eps2);
        }
        total += c;
    }
    return total;
}

/* ------------------------------------------------------------------------ */
/* the reference's expanded-form arithmetic                                  */
/* ------------------------------------------------------------------------ */

/* Block norm terms of one point (egojoin.h:182-195):
 *   blk[0] = (eps^2/2 - sum_{k < 2*KBLOCK} x_k^2) / 2
 *   blk[1] = 0                        (calloc'd, never written)
 *   blk[b] = (0 - sum_{k in [b*KBLOCK, (b+1)*KBLOCK)} x_k^2) / 2,  b >= 2 */
static void expanded_blocks(const double *x, int d, double eps, int kblock, double *blk) {
    int nb = (d + kblock - 1) / kblock;
    double e = eps * eps / 2;
    double h = e;
    for (int j = 0; j < 2 * kblock && j < d; j++) h -= x[j] * x[j];
    blk[0] = h / 2;
    if (nb > 1) blk[1] = 0.0;
    for (int b = 2; b < nb; b++) {
        h = 0;
        for (int j = b * kblock; j < (b + 1) * kblock && j < d; j++) h -= x[j] * x[j];
        blk[b] = h / 2;
    }
}

/* egojoin.h:329-381 for one (lane, R) pair, evaluated to the end: the sum
 * starts at selfA[0] + selfB[0], takes fma(a_k, b_k, sum) for every k and,
 * at k % KBLOCK == 0 && k > KBLOCK, adds selfA[k/KBLOCK] + selfB[k/KBLOCK].
 * The pair counts iff the sign bit of the final sum is clear
 * (egojoin.cpp:803-811). */
static int within_expanded_blk(const double *a, const double *b, int d, int kblock,
                               const double *ba, const double *bb) {
    double sum = ba[0] + bb[0];
    sum = fma(a[0], b[0], sum);
    for (int k = 1; k < d; k++) {
        if (k % kblock == 0 && k > kblock) sum += ba[k / kblock] + bb[k / kblock];
        sum = fma(a[k], b[k], sum);
    }
    return !signbit(sum);
}

/* The same sum, stopped at the first check point (k % KBLOCK == 0 && k >
 * KBLOCK, before the block norms are added) where it is negative.  The
 * reference stops a whole 8x8 tile there only when all 64 of its sums are
 * negative (egojoin.h:350-359, :521-530), so a pair's result in any tile is
 * either this one or the one evaluated to the end. */
static int within_expanded_stop(const double *a, const double *b, int d, int kblock,
                                const double *ba, const double *bb) {
    double sum = ba[0] + bb[0];
    sum = fma(a[0], b[0], sum);
    for (int k = 1; k < d; k++) {
        if (k % kblock == 0 && k > kblock) {
            if (signbit(sum)) return 0;
            sum += ba[k / kblock] + bb[k / kblock];
        }
        sum = fma(a[k], b[k], sum);
    }
    return !signbit(sum);
}

/* 1 when the reference's arithmetic can give this pair a result other than
 * the fixed-order predicate's (in some tile context) */
static int expanded_may_disagree(const double *a, const double *b, int d, int kblock,
                                 const double *ba, const double *bb, double eps2) {
    const int f = hjo_dist2(a, b, d) <= eps2;
    return within_expanded_blk(a, b, d, kblock, ba, bb) != f ||
           within_expanded_stop(a, b, d, kblock, ba, bb) != f;
}

int hjo_pair_within_expanded(const double *a, const double *b, int d, double eps, int kblock) {
    int nb = (d + kblock - 1) / kblock;
    double *ba = (double *)malloc(sizeof(double) * (size_t)(2 * nb + 2));
    double *bb = ba + nb + 1;
    expanded_blocks(a, d, eps, kblock, ba);
    expanded_blocks(b, d, eps, kblock, bb);
    int r = within_expanded_blk(a, b, d, kblock, ba, bb);
    free(ba);
    return r;
}

static double *all_blocks(const double *X, int64_t n, int d, double eps, int kblock, int *nb_out) {
    int nb = (d + kblock - 1) / kblock;
    double *blk = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * (size_t)nb);
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++) expanded_blocks(X + i * d, d, eps, kblock, blk + i * nb);
    *nb_out = nb;
    return blk;
}

uint64_t hjo_self_count_expanded(const double *X, int64_t n, int d, double eps, int kblock) {
    int nb;
    double *blk = all_blocks(X, n, d, eps, kblock, &nb);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++)
            total += (uint64_t)within_expanded_blk(X + i * d, X + j * d, d, kblock,
                                                   blk + i * nb, blk + j * nb);
    free(blk);
    return total;
}

static int64_t append_pair(int64_t *out, int64_t cap, int64_t *cnt, int64_t i, int64_t j) {
    int64_t k;
#pragma omp atomic capture
    k = (*cnt)++;
    if (out && k < cap) {
        out[2 * k] = i;
        out[2 * k + 1] = j;
    }
    return k;
}

/* A x B count of the expanded form evaluated to the end (egojoin.h:482-553,
 * sign rule egojoin.cpp:868-884): the NOSELF tile takes its block norms from
 * A and B separately (egojoin.h:425-450) but adds them exactly as the self
 * tile does. */
uint64_t hjo_join_count_expanded(const double *A, int64_t nA, const double *B, int64_t nB,
                                 int d, double eps, int kblock) {
    int nba, nbb;
    double *ba = all_blocks(A, nA, d, eps, kblock, &nba);
    double *bb = all_blocks(B, nB, d, eps, kblock, &nbb);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t i = 0; i < nA; i++)
        for (int64_t j = 0; j < nB; j++)
            total += (uint64_t)within_expanded_blk(A + i * d, B + j * d, d, kblock, ba + i * nba,
                                                   bb + j * nbb);
    free(ba);
    free(bb);
    return total;
}

int64_t hjo_self_disagreements(const double *X, int64_t n, int d, double eps, int kblock,
                               int64_t *out, int64_t cap) {
    int nb;
    double *blk = all_blocks(X, n, d, eps, kblock, &nb);
    double eps2 = eps * eps;
    int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++)
            if (expanded_may_disagree(X + i * d, X + j * d, d, kblock, blk + i * nb,
                                      blk + j * nb, eps2))
                append_pair(out, cap, &cnt, i, j);
    free(blk);
    return cnt;
}

/* Pairs (a, b) of A x B on which the reference's NOSELF tile arithmetic
 * (egojoin.h:482-553 with the block norms of :425-450, counted by the sign
 * rule of egojoin.cpp:868-884) can disagree with the fixed-order predicate. */
int64_t hjo_join_disagreements(const double *A, int64_t nA, const double *B, int64_t nB, int d,
                               double eps, int kblock, int64_t *out, int64_t cap) {
    int nba, nbb;
    double *ba = all_blocks(A, nA, d, eps, kblock, &nba);
    double *bb = all_blocks(B, nB, d, eps, kblock, &nbb);
    double eps2 = eps * eps;
    int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nA; i++)
        for (int64_t j = 0; j < nB; j++)
            if (expanded_may_disagree(A + i * d, B + j * d, d, kblock, ba + i * nba,
                                      bb + j * nbb, eps2))
                append_pair(out, cap, &cnt, i, j);
    free(ba);
    free(bb);
    return cnt;
}

int64_t hjo_self_band(const double *X, int64_t n, int d, double eps, double tol,
                      int64_t *out, int64_t cap) {
    double eps2 = eps * eps;
    int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++)
            if (fabs(hjo_dist2(X + i * d, X + j * d, d) - eps2) <= tol)
                append_pair(out, cap, &cnt, i, j);
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* CSR neighbour lists                                                        */
/* ------------------------------------------------------------------------ */

static uint64_t csr_fill(const double *A, int64_t nA, const double *B, int64_t nB, int d,
                         double eps, int self, int64_t *rowptr, int64_t *colidx) {
    double eps2 = eps * eps;
    rowptr[0] = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nA; i++) {
        int64_t c = 0;
        for (int64_t j = 0; j < nB; j++) {
            if (self && j == i) continue;
            c += within_ee(A + i * d, B + j * d, d, eps2);
        }
        rowptr[i + 1] = c;
    }
    for (int64_t i = 0; i < nA; i++) rowptr[i + 1] += rowptr[i];
    if (colidx) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < nA; i++) {
            int64_t p = rowptr[i];
            for (int64_t j = 0; j < nB; j++) {
                if (self && j == i) continue;
                if (within_ee(A + i * d, B + j * d, d, eps2)) colidx[p++] = j;
            }
        }
    }
    return (uint64_t)rowptr[nA];
}

uint64_t hjo_self_csr(const double *X, int64_t n, int d, double eps,
                      int64_t *rowptr, int64_t *colidx) {
    return csr_fill(X, n, X, n, d, eps, 1, rowptr, colidx);
}

uint64_t hjo_join_csr(const double *A, int64_t nA, const double *B, int64_t nB, int d,
                      double eps, int64_t *rowptr, int64_t *colidx) {
    return csr_fill(A, nA, B, nB, d, eps, 0, rowptr, colidx);
}

/* ------------------------------------------------------------------------ */
/* EGO sort (egojoin.cpp:159-168 with the comparator of :9-17)                */
/* ------------------------------------------------------------------------ */

typedef struct {
    const int64_t *cells; /* n x d */
    int d;
} cell_ctx;

static int cmp_cells(const int64_t *a, const int64_t *b, int d) {
    for (int k = 0; k < d; k++) {
        if (a[k] < b[k]) return -1;
        if (a[k] > b[k]) return 1;
    }
    return 0;
}

/* stable merge sort of indices by cell vectors; parallel over chunks */
static void merge_runs(const cell_ctx *c, const int64_t *src, int64_t *dst, int64_t lo,
                       int64_t mid, int64_t hi) {
    int64_t i = lo, j = mid, k = lo;
    while (i < mid && j < hi) {
        if (cmp_cells(c->cells + src[j] * c->d, c->cells + src[i] * c->d, c->d) < 0)
            dst[k++] = src[j++];
        else
            dst[k++] = src[i++];
    }
    while (i < mid) dst[k++] = src[i++];
    while (j < hi) dst[k++] = src[j++];
}

static void ego_sort_cells(const int64_t *cells, int64_t n, int d, int64_t *perm) {
    cell_ctx c = {cells, d};
    int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) perm[i] = i;
    /* bottom-up merge sort: run width doubles each round */
    int64_t *src = perm, *dst = tmp;
    for (int64_t w = 1; w < n; w *= 2) {
#pragma omp parallel for schedule(static)
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n;
            int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            merge_runs(&c, src, dst, lo, mid, hi);
        }
        int64_t *t = src;
        src = dst;
        dst = t;
    }
    if (src != perm) memcpy(perm, src, sizeof(int64_t) * (size_t)n);
    free(tmp);
}

static int64_t *compute_cells(const double *X, int64_t n, int d, double eps) {
    int64_t *cells = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1) * (size_t)d);
    double r = 1.0 / eps;
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < d; k++) cells[i * d + k] = (int64_t)floor(X[i * d + k] * r);
    return cells;
}

void hjo_ego_sort(const double *X, int64_t n, int d, double eps, int64_t *perm) {
    int64_t *cells = compute_cells(X, n, d, eps);
    ego_sort_cells(cells, n, d, perm);
    free(cells);
}

/* ------------------------------------------------------------------------ */
/* CPU EGO join (the CPU baseline)                                            */
/* ------------------------------------------------------------------------ */

/* Per-dimension partner cell window of a point.  A partner q of p that the
 * fixed-order predicate accepts has |x_q - x_p| <= eps (1 + 3u), so
 * |fl(x_q r) - fl(x_p r)| <= 1 + eta with eta << 2^-48 (1 + |x_p r|).  The
 * window is [c-1, c+1], widened by one cell on a side where fl(x_p r) lies
 * within eta of a cell boundary.  (The reference displaces x by +-eps and
 * re-floors, egojoin.cpp:281-315 and :1133-1144, which can miss such
 * boundary pairs.) */
static void cell_window(double x, double r, int64_t *lo, int64_t *hi) {
    double X = x * r;
    double c = floor(X);
    double f = X - c;
    double eta = ldexp(1.0 + fabs(X), -48);
    *lo = (int64_t)c - 1 - (f < eta ? 1 : 0);
    *hi = (int64_t)c + 1 + (f >= 1.0 - eta ? 1 : 0);
}

/* first index in [0, n) whose cell vector is >= q (lex), or n */
static int64_t lower_bound_cells(const int64_t *scells, int64_t n, int d, const int64_t *q) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t m = lo + (hi - lo) / 2;
        if (cmp_cells(scells + m * d, q, d) < 0)
            lo = m + 1;
        else
            hi = m;
    }
    return lo;
}

/* first index in [0, n) whose cell vector is > q (lex), or n */
static int64_t upper_bound_cells(const int64_t *scells, int64_t n, int d, const int64_t *q) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t m = lo + (hi - lo) / 2;
        if (cmp_cells(scells + m * d, q, d) <= 0)
            lo = m + 1;
        else
            hi = m;
    }
    return lo;
}

typedef struct {
    int64_t s, e;
} ival;

static int cmp_ival(const void *a, const void *b) {
    const ival *x = (const ival *)a, *y = (const ival *)b;
    return x->s < y->s ? -1 : (x->s > y->s ? 1 : 0);
}

/* Candidate 8-block ranges of every 8-block of the (sorted) query set against
 * the sorted target set: for each point and each combination w of cells in
 * the first A dimensions within the point's window, the targets whose cell
 * vectors lie between (w, lo_A..lo_{d-1}) and (w, hi_A..hi_{d-1}) -- the
 * stripes of egojoin.cpp:270-464 / :1126-1240 -- are unioned per 8-block
 * (block min/max as egojoin.cpp:437-444) and merged.  self: only targets in
 * blocks >= the query block are kept (pairs i<j).  Returns per-block
 * interval lists in *iv (blk_off has nblk+1 entries). */
static void ego_block_ranges(const double *Q, const int64_t *qcells, int64_t nq,
                             const int64_t *tcells, int64_t nt, int d, double eps, int A,
                             int self, ival **iv_out, int64_t **off_out) {
    double r = 1.0 / eps;
    int64_t nblk = (nq + 7) / 8;
    int64_t ntblk = (nt + 7) / 8;
    int64_t maxw = 1;
    for (int k = 0; k < A; k++) maxw *= 4;
    int64_t cap_per_blk = 8 * maxw;
    ival *tmp = (ival *)malloc(sizeof(ival) * (size_t)(nblk > 0 ? nblk : 1) * (size_t)cap_per_blk);
    int64_t *cnt = (int64_t *)calloc((size_t)nblk + 1, sizeof(int64_t));
#pragma omp parallel
    {
        int64_t *lo = (int64_t *)malloc(sizeof(int64_t) * (size_t)d * 4);
        int64_t *hi = lo + d, *ql = lo + 2 * d, *qh = lo + 3 * d;
#pragma omp for schedule(dynamic, 64)
        for (int64_t b = 0; b < nblk; b++) {
            ival *mine = tmp + b * cap_per_blk;
            int64_t m = 0;
            for (int64_t p = b * 8; p < b * 8 + 8 && p < nq; p++) {
                for (int k = 0; k < d; k++) cell_window(Q[p * d + k], r, lo + k, hi + k);
                /* enumerate w over the active-dimension windows */
                int64_t w[8];
                for (int k = 0; k < A; k++) w[k] = lo[k];
                for (;;) {
                    int keep = 1;
                    if (self) {
                        /* w - c >=lex 0 : targets after p in sorted order */
                        for (int k = 0; k < A; k++) {
                            int64_t c = qcells[p * d + k];
                            if (w[k] > c) break;
                            if (w[k] < c) { keep = 0; break; }
                        }
                    }
                    if (keep) {
                        for (int k = 0; k < d; k++) {
                            ql[k] = k < A ? w[k] : lo[k];
                            qh[k] = k < A ? w[k] : hi[k];
                        }
                        int64_t s = lower_bound_cells(tcells, nt, d, ql);
                        int64_t e = upper_bound_cells(tcells, nt, d, qh);
                        if (s < e) {
                            mine[m].s = s / 8;
                            mine[m].e = (e + 7) / 8;
                            m++;
                        }
                    }
                    int k = A - 1;
                    while (k >= 0 && w[k] == hi[k]) { w[k] = lo[k]; k--; }
                    if (k < 0) break;
                    w[k]++;
                }
            }
            /* merge */
            qsort(mine, (size_t)m, sizeof(ival), cmp_ival);
            int64_t o = 0;
            for (int64_t t = 0; t < m; t++) {
                int64_t s = mine[t].s, e = mine[t].e;
                if (self && s < b) s = b;
                if (e > ntblk) e = ntblk;
                if (s >= e) continue;
                if (o > 0 && s <= mine[o - 1].e) {
                    if (e > mine[o - 1].e) mine[o - 1].e = e;
                } else {
                    mine[o].s = s;
                    mine[o].e = e;
                    o++;
                }
            }
            cnt[b + 1] = o;
        }
        free(lo);
    }
    for (int64_t b = 0; b < nblk; b++) cnt[b + 1] += cnt[b];
    ival *iv = (ival *)malloc(sizeof(ival) * (size_t)(cnt[nblk] > 0 ? cnt[nblk] : 1));
    for (int64_t b = 0; b < nblk; b++)
        memcpy(iv + cnt[b], tmp + b * cap_per_blk, sizeof(ival) * (size_t)(cnt[b + 1] - cnt[b]));
    free(tmp);
    *iv_out = iv;
    *off_out = cnt;
}

/* 8x8 block of pairs with the fixed-order predicate; the early exit is taken
 * every kblock dimensions once every pair of the block exceeds eps^2, as the
 * reference's tile loop does (egojoin.h:349-359).  The target block comes
 * transposed, [d][8] (the reference's AoSoA-8 layout, egojoin.cpp:234-267),
 * so the 8 partners of one row advance together; lanes past the set's end
 * hold NaN, which never counts.  Each pair still sums its dimensions in
 * ascending order with one fma per dimension. */
static uint64_t block8(const double *Ys, int64_t n_a, const double *ZT, int d, double eps2,
                       int64_t ib, int64_t jb, int diag, int kblock) {
    double s[8][8];
    int ia = (int)(n_a - ib * 8 < 8 ? n_a - ib * 8 : 8);
    const double *a = Ys + ib * 8 * d;
    const double *bt = ZT + jb * 8 * d;
    memset(s, 0, sizeof(s));
    for (int k0 = 0; k0 < d; k0 += kblock) {
        int k1 = k0 + kblock < d ? k0 + kblock : d;
        int alive = 0;
        for (int x = 0; x < ia; x++) {
            double *sx = s[x];
            for (int k = k0; k < k1; k++) {
                const double av = a[x * d + k];
                const double *bk = bt + k * 8;
#pragma omp simd
                for (int y = 0; y < 8; y++) {
                    double t = av - bk[y];
                    sx[y] = fma(t, t, sx[y]);
                }
            }
            for (int y = 0; y < 8; y++) alive |= (sx[y] <= eps2) & (!diag | (x < y));
        }
        if (!alive) return 0;
    }
    uint64_t c = 0;
    for (int x = 0; x < ia; x++)
        for (int y = 0; y < 8; y++)
            c += (uint64_t)((s[x][y] <= eps2) & (!diag | (x < y)));
    return c;
}

/* the sorted target set as 8-point blocks [blk][d][8], NaN past its end */
static double *transpose8(const double *Zs, int64_t n, int d) {
    int64_t nblk = (n + 7) / 8;
    double *T = (double *)malloc(sizeof(double) * (size_t)(nblk > 0 ? nblk : 1) * 8 * (size_t)d);
#pragma omp parallel for
    for (int64_t b = 0; b < nblk; b++)
        for (int k = 0; k < d; k++)
            for (int y = 0; y < 8; y++) {
                int64_t i = b * 8 + y;
                T[(b * d + k) * 8 + y] = i < n ? Zs[i * d + k] : NAN;
            }
    return T;
}

static uint64_t ego_join_impl(const double *A, int64_t nA, const double *B, int64_t nB, int d,
                              double eps, int active_dims, int kblock, int threads, int self,
                              int64_t iblock_stride, double *times, uint64_t *evaluated) {
    if (threads > 0) omp_set_num_threads(threads);
    if (kblock < 1) kblock = 4;
    if (iblock_stride < 1) iblock_stride = 1;
    int A_ = active_dims < 0 ? 0 : (active_dims > d ? d : active_dims);
    if (A_ > 6) A_ = 6;
    double t0 = now_s();
    int64_t *ca = compute_cells(A, nA, d, eps);
    int64_t *pa = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nA > 0 ? nA : 1));
    ego_sort_cells(ca, nA, d, pa);
    int64_t *cb = ca, *pb = pa;
    if (!self) {
        cb = compute_cells(B, nB, d, eps);
        pb = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nB > 0 ? nB : 1));
        ego_sort_cells(cb, nB, d, pb);
    }
    /* sorted copies of points and cells */
    double *Ys = (double *)malloc(sizeof(double) * (size_t)(nA > 0 ? nA : 1) * (size_t)d);
    int64_t *Yc = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nA > 0 ? nA : 1) * (size_t)d);
#pragma omp parallel for
    for (int64_t i = 0; i < nA; i++) {
        memcpy(Ys + i * d, A + pa[i] * d, sizeof(double) * (size_t)d);
        memcpy(Yc + i * d, ca + pa[i] * d, sizeof(int64_t) * (size_t)d);
    }
    double *Zs = Ys;
    int64_t *Zc = Yc;
    if (!self) {
        Zs = (double *)malloc(sizeof(double) * (size_t)(nB > 0 ? nB : 1) * (size_t)d);
        Zc = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nB > 0 ? nB : 1) * (size_t)d);
#pragma omp parallel for
        for (int64_t i = 0; i < nB; i++) {
            memcpy(Zs + i * d, B + pb[i] * d, sizeof(double) * (size_t)d);
            memcpy(Zc + i * d, cb + pb[i] * d, sizeof(int64_t) * (size_t)d);
        }
    }
    double t1 = now_s();
    ival *iv;
    int64_t *off;
    ego_block_ranges(Ys, Yc, nA, Zc, self ? nA : nB, d, eps, A_, self, &iv, &off);
    double *ZT = transpose8(Zs, self ? nA : nB, d);
    double t2 = now_s();
    double eps2 = eps * eps;
    int64_t nblk = (nA + 7) / 8;
    uint64_t total = 0, evals = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total, evals)
    for (int64_t ib = 0; ib < nblk; ib += iblock_stride) {
        for (int64_t t = off[ib]; t < off[ib + 1]; t++)
            for (int64_t jb = iv[t].s; jb < iv[t].e; jb++) {
                total += block8(Ys, nA, ZT, d, eps2, ib, jb, self && jb == ib, kblock);
                evals += 64;
            }
    }
    double t3 = now_s();
    free(ZT);
    if (times) {
        times[0] = t1 - t0;
        times[1] = t2 - t1;
        times[2] = t3 - t2;
    }
    if (evaluated) *evaluated = evals;
    free(iv);
    free(off);
    free(Ys);
    free(Yc);
    free(ca);
    free(pa);
    if (!self) {
        free(Zs);
        free(Zc);
        free(cb);
        free(pb);
    }
    return total;
}

uint64_t hjo_ego_self_count(const double *X, int64_t n, int d, double eps, int active_dims,
                            int kblock, int threads, double *times, uint64_t *evaluated) {
    return ego_join_impl(X, n, X, n, d, eps, active_dims, kblock, threads, 1, 1, times, evaluated);
}

uint64_t hjo_ego_join_count(const double *A, int64_t nA, const double *B, int64_t nB, int d,
                            double eps, int active_dims, int kblock, int threads,
                            double *times, uint64_t *evaluated) {
    return ego_join_impl(A, nA, B, nB, d, eps, active_dims, kblock, threads, 0, 1, times,
                         evaluated);
}

/* sampled variant for the CPU baseline: only every iblock_stride-th 8-block
 * of the query set runs the join loop (sort and index run in full) */
uint64_t hjo_ego_self_count_sampled(const double *X, int64_t n, int d, double eps,
                                    int active_dims, int kblock, int threads,
                                    int64_t iblock_stride, double *times, uint64_t *evaluated) {
    return ego_join_impl(X, n, X, n, d, eps, active_dims, kblock, threads, 1, iblock_stride,
                         times, evaluated);
}

/* ------------------------------------------------------------------------ */
/* textbook Hilbert curve (the traversal order of hilloop.h:9-82)            */
/* ------------------------------------------------------------------------ */

/* egojoin.cpp:538-570: histogram of the cells of dimension k, then
 * sum h (h - 1) / 2 over the cells plus the products of neighbouring cells. */
int hjo_dim_costs(const double *X, int64_t n, int d, double eps, uint64_t *cost) {
    const double r = 1.0 / eps;
    for (int k = 0; k < d; k++) {
        cost[k] = 0;
        if (n == 0) continue;
        double lo = floor(X[k] * r), hi = lo;
        for (int64_t i = 1; i < n; i++) {
            const double c = floor(X[i * d + k] * r);
            if (c < lo) lo = c;
            if (c > hi) hi = c;
        }
        if (hi - lo >= (double)(1 << 26)) return -1;
        const int64_t size = (int64_t)(hi - lo) + 1;
        uint64_t *h = (uint64_t *)calloc((size_t)size, sizeof(uint64_t));
        if (!h) return -1;
        for (int64_t i = 0; i < n; i++) h[(int64_t)(floor(X[i * d + k] * r) - lo)]++;
        uint64_t v = 0;
        for (int64_t b = 0; b < size; b++) {
            if (h[b]) v += h[b] * (h[b] - 1) / 2;
            if (b > 0) v += h[b - 1] * h[b];
        }
        cost[k] = v;
        free(h);
    }
    return 0;
}

uint64_t hjo_hilbert_xy2d(int order, uint32_t x, uint32_t y) {
    uint64_t h = 0;
    uint64_t nside = (uint64_t)1 << order;
    for (uint64_t s = nside >> 1; s > 0; s >>= 1) {
        uint32_t rx = (x & s) ? 1u : 0u;
        uint32_t ry = (y & s) ? 1u : 0u;
        h += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = (uint32_t)(nside - 1 - x);
                y = (uint32_t)(nside - 1 - y);
            }
            uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return h;
}

void hjo_hilbert_d2xy(int order, uint64_t h, uint32_t *xo, uint32_t *yo) {
    uint64_t nside = (uint64_t)1 << order;
    uint64_t x = 0, y = 0, t = h;
    for (uint64_t s = 1; s < nside; s <<= 1) {
        uint64_t rx = 1 & (t / 2);
        uint64_t ry = 1 & (t ^ rx);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            uint64_t tmp = x;
            x = y;
            y = tmp;
        }
        x += s * rx;
        y += s * ry;
        t /= 4;
    }
    *xo = (uint32_t)x;
    *yo = (uint32_t)y;
}
