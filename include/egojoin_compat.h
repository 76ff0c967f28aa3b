/*
 * egojoin_compat.h -- the reference's two join entry points, with their exact
 * C++ signatures, implemented on the MI355X library (libhilbertjoin_amd.so).
 *
 * A program written against hilbertjoin/egojoin.h links these instead of the
 * AVX-512 code.  Behaviour differences (deliberate):
 *   - the input arrays are NOT modified (the reference EGO-sorts them in place
 *     and restores them with a faulty transpose, egojoin.cpp:476-511);
 *   - every point takes part (the reference drops the n % 8 tail,
 *     egojoin.h:325,496), and n < 8 is handled;
 *   - the count uses the fixed-order FP64 predicate (see hilbertjoin_amd.h);
 *   - test_ego_loop3_noself honours activedims (the reference hard-codes 2,
 *     egojoin.cpp:847).
 * The device is HJ_DEVICE (default 0); the kernel is HJ_KERNEL = fp64
 * (default), f16 or bf16.  Errors print a message and exit(1), as the
 * reference's helpers do (allocation.cpp:13-16).
 */
#ifndef EGOJOIN_COMPAT_H
#define EGOJOIN_COMPAT_H

#include <stddef.h>

/* egojoin.h:56 -- self-join count; sortTime / indextime in seconds;
 * loadpercent = evaluated pairs / (n (n-1) / 2), the quantity of
 * egojoin.cpp:797. */
void test_ego_loop3_macro(size_t n, size_t d, double epsilon, double *array,
                          size_t *countresult, int stripes, double *sortTime, double *indextime,
                          double *loadpercent);

/* egojoin.h:64 -- A x B join count; loadpercent in the units of
 * egojoin.cpp:865: (evaluated pairs / 64) / (nA/8) / (nA/8 - 1) * 200. */
void test_ego_loop3_noself(const size_t nA, const size_t nB, const int d, const double epsilon,
                           double *arrayA, double *arrayB, size_t *countresult,
                           const int activedims, double *sortTime, double *indextime,
                           double *loadpercent);

/* egojoin.h:59 -- dimension statistics: reorder_dim (d + 8 entries, the last
 * 8 mapped to themselves) receives the position of each dimension when the
 * dimensions are sorted by increasing selectivity cost, computed exactly on
 * the GPU (the reference samples). */
void outputStatistics(int n, int d, double epsilon, double *array, int *reorder_dim);

/* egojoin.h:60 -- moves column j of every row to column reorder_dim[j], in
 * place, for any d (the reference's AVX-512 mask skips the last d % 8
 * columns, egojoin.cpp:674). */
void reorder_dimensions(int n, int d, double *array, int *reorder_dim);

#endif
