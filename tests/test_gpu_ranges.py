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
