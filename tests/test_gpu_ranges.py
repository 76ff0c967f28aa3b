"""Coordinate ranges in two passes.

The grid's key fields use only the leading 16 dimensions, so their ranges
are taken first; the ranges of the other dimensions are taken in a second
pass on the second stream and complete the filter operands' centre and scale
and the thresholds.  These cases cross the split: d below, at and just past
16, a second pass wider than one 64-dimension block, A x B joins (both sides'
passes), batches (a chunk table of several segments, jobs of one point), and
non-finite coordinates on either side of the split, which must raise before
any result is returned and leave the context usable.
"""
import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle
import planted

pytestmark = pytest.mark.gpu

DIMS = [1, 15, 16, 17, 64, 100]


def _eps_for(d):
    # a radius that leaves a few natural pairs besides the planted ones
    return 0.05 * np.sqrt(d) if d > 1 else 0.001


@pytest.mark.parametrize("d", DIMS)
@pytest.mark.parametrize("kernel", ["f16", "fp64"])
def test_self_across_the_split(ctx, d, kernel):
    eps = _eps_for(d)
    X = hj.gen_uniform(80 + d, 3000, d)
    Xp, _, inside = planted.plant_self(X, eps, 40, seed=8000 + d)
    want = oracle.self_count(Xp, eps)
    assert want >= inside.sum() > 0
    got, _ = ctx.self_join_count(Xp, eps, kernel=kernel)
    assert got == want


@pytest.mark.parametrize("d", [16, 17, 100])
def test_join_across_the_split(ctx, d):
    # the two sides' ranges differ (B shifted), so each side's passes matter
    eps = _eps_for(d)
    A = hj.gen_uniform(90 + d, 2500, d)
    B = hj.gen_uniform(91 + d, 1800, d) * 0.8 + 0.15
    Ap, Bp, _, inside = planted.plant_join(A, B, eps, 30, seed=8100 + d)
    want = oracle.join_count(Ap, Bp, eps)
    assert want >= inside.sum() > 0
    for kernel in ("f16", "fp64"):
        got, _ = ctx.join_count(Ap, Bp, eps, kernel=kernel)
        assert got == want, kernel


@pytest.mark.parametrize("d", [17, 100])
def test_batch_across_the_split(ctx, d):
    eps = _eps_for(d)
    sizes = [1, 700, 2500, 64, 1, 1300]
    jobs = []
    for k, n in enumerate(sizes):
        J = hj.gen_uniform(9000 + 10 * d + k, n, d) * (1.0 + 0.5 * k) - 0.3 * k
        if n > 10:
            J, _, _ = planted.plant_self(J, eps, 10, seed=9100 + k, lo=-0.3 * k,
                                         hi=1.0 + 0.2 * k)
        jobs.append(J)
    off = np.zeros(len(sizes) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    want = np.array([oracle.self_count(J, eps) for J in jobs])
    assert want.sum() > 0
    dev = ctx.to_device(np.concatenate(jobs))
    for kernel in ("f16", "fp64"):
        got, _ = ctx.batch_self_join_count(dev, off, eps, kernel=kernel)
        assert np.array_equal(got, want), kernel
    dev.free()


@pytest.mark.parametrize("dim", [3, 30])
@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_nonfinite_on_either_side_of_the_split(ctx, dim, bad):
    d, eps = 40, 0.4
    X = hj.gen_uniform(95, 2000, d)
    Y = X.copy()
    Y[1234, dim] = bad
    with pytest.raises(hj.HJError):
        ctx.self_join_count(Y, eps)
    with pytest.raises(hj.HJError):
        ctx.join_count(X, Y, eps)
    off = np.array([0, 1000, 1500, 2000], dtype=np.uint64)
    dev = ctx.to_device(Y)
    with pytest.raises(hj.HJError):
        ctx.batch_self_join_count(dev, off, eps)
    dev.free()
    # the context stays usable
    Xp, _, _ = planted.plant_self(X, eps, 20, seed=9500 + dim)
    got, _ = ctx.self_join_count(Xp, eps)
    assert got == oracle.self_count(Xp, eps)


# A single FP16 filter join whose sides have at least 8192 rows takes the
# ranges of the dimensions past the 16th from every 16th row (widened by a
# sixteenth of their width on each side) when the sample's ranges of the key
# dimensions come close to their exact ones, and checks every point against
# them where the operands are made; a point outside (here in a row that is not
# sampled: rows 16 j from the side's start are) makes the join run again from
# exact ranges (stats.sampled_ranges 2 instead of 1).  Data with long tails,
# batches, BF16 (whose band is the more sensitive to a wider range) and the
# FP64 kernel take exact ranges (0).
N_BIG = 20_000


def _with_outlier(seed, n, d, eps, row, dim, value):
    # pairs planted at 0.02 (they reach at most about 1% past the unit cube,
    # so the key dimensions' tails stay short), one coordinate far out in a
    # row that is not sampled, and a pair at 0.3 eps through it
    X = hj.gen_uniform(seed, n, d)
    Xp, _, _ = planted.plant_self(X, 0.02, 25, seed=seed + 1)
    Xp = Xp.copy()
    Xp[row, dim] = value
    Xp[row + 1000] = Xp[row]
    k = 17  # a dimension past the key candidates, moved towards the middle
    Xp[row + 1000, k] += -0.3 * eps if Xp[row, k] > 0.5 else 0.3 * eps
    return Xp


@pytest.mark.parametrize("kernel", ["f16", "bf16"])
def test_sampled_ranges_hold(ctx, kernel):
    # pairs planted at eps 0.02 reach at most about 1% past the unit cube
    d, eps = 40, 0.02
    X = hj.gen_uniform(97, N_BIG, d)
    Xp, _, _ = planted.plant_self(X, eps, 25, seed=98)
    got, st = ctx.self_join_count(Xp, eps, kernel=kernel)
    assert got == oracle.self_count(Xp, eps)
    assert st.sampled_ranges == (1 if kernel == "f16" else 0)
    _, st64 = ctx.self_join_count(Xp, eps, kernel="fp64")
    assert st64.sampled_ranges == 0
    _, st16 = ctx.self_join_count(np.ascontiguousarray(Xp[:, :16]), eps, kernel=kernel)
    assert st16.sampled_ranges == 0


@pytest.mark.parametrize("value", [3.0, -2.5])
@pytest.mark.parametrize("kernel", ["f16", "bf16"])
def test_sampled_ranges_missed_self(ctx, value, kernel):
    d, eps = 40, 0.35
    Xp = _with_outlier(99, N_BIG, d, eps, 1001, 25, value)
    want = oracle.self_count(Xp, eps)
    got, st = ctx.self_join_count(Xp, eps, kernel=kernel)
    assert got == want
    assert st.sampled_ranges == (2 if kernel == "f16" else 0)
    # the context goes on from sampled ranges
    X = hj.gen_uniform(100, N_BIG, d)
    got, st = ctx.self_join_count(X, eps, kernel=kernel)
    assert got == oracle.self_count(X, eps)
    assert st.sampled_ranges == (1 if kernel == "f16" else 0)


def test_sampled_ranges_missed_csr(ctx):
    d, eps = 40, 0.35
    Xp = _with_outlier(101, N_BIG, d, eps, 2003, 33, 4.0)
    rp, cols, st = ctx.self_join_csr(Xp, eps, kernel="f16")
    rowptr, want = oracle.self_csr(Xp, eps)
    rp, cols = np.asarray(rp).ravel(), np.asarray(cols).ravel()  # host input: host arrays
    assert np.array_equal(rp.astype(np.int64), np.asarray(rowptr, dtype=np.int64))
    got = planted.csr_pairs(rp, cols)
    assert got == planted.csr_pairs(np.asarray(rowptr), np.asarray(want))
    assert st.sampled_ranges == 2


def test_sampled_ranges_missed_join(ctx):
    d, eps = 40, 0.35
    A = hj.gen_uniform(102, N_BIG, d)
    B = _with_outlier(103, 12_000, d, eps, 501, 20, 2.0)
    # an A point on B's outlier (row 8 of A is not sampled either)
    A[8] = B[501]
    want = oracle.join_count(A, B, eps)
    got, st = ctx.join_count(A, B, eps, kernel="f16")
    assert got == want
    assert st.sampled_ranges == 2


def test_sampled_ranges_batch_takes_exact_ranges(ctx):
    d, eps = 40, 0.35
    sizes = [300, N_BIG, 5000, 9000]
    jobs = [hj.gen_uniform(104 + k, n, d) for k, n in enumerate(sizes)]
    jobs[1] = _with_outlier(110, N_BIG, d, eps, 3001, 39, -1.5)
    off = np.zeros(len(sizes) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    want = np.array([oracle.self_count(J, eps) for J in jobs])
    dev = ctx.to_device(np.concatenate(jobs))
    got, st = ctx.batch_self_join_count(dev, off, eps, kernel="f16")
    assert np.array_equal(got, want)
    assert st.sampled_ranges == 0  # batches take exact ranges
    dev.free()


def test_sampled_ranges_long_tails_take_exact_ranges(ctx):
    # Gaussian coordinates: a sample's extremes fall well short of the data's
    rng = np.random.default_rng(112)
    X = rng.normal(size=(N_BIG, 40)) * 0.2
    eps = 0.35
    Xp, _, _ = planted.plant_self(X, eps, 25, seed=113, lo=-0.8, hi=0.8)
    got, st = ctx.self_join_count(Xp, eps, kernel="f16")
    assert got == oracle.self_count(Xp, eps)
    assert st.sampled_ranges == 0


def test_sampled_ranges_nonfinite_outside_the_sample(ctx):
    d, eps = 40, 0.35
    X = hj.gen_uniform(111, N_BIG, d)
    Y = X.copy()
    Y[1001, 30] = np.nan  # not a sampled row: found by the operand check
    with pytest.raises(hj.HJError):
        ctx.self_join_count(Y, eps, kernel="f16")
    got, _ = ctx.self_join_count(X, eps, kernel="f16")
    assert got == oracle.self_count(X, eps)
