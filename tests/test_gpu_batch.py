"""GPU parity of the batched joins: every job's count equals the oracle's count
of that job alone (bit-exact), whatever the other jobs in the batch are."""
import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle

pytestmark = pytest.mark.gpu

KERNELS = ["fp64", "f16", "bf16"]


def _jobs(seed, sizes, d, scales=None):
    """Concatenated points of len(sizes) jobs; job k is a Gaussian mixture
    scaled by scales[k] and shifted by k (so the jobs' coordinate ranges and
    grids differ)."""
    parts = []
    for k, n in enumerate(sizes):
        X = hj.gen_mixture(seed * 1000 + k, n, d, 6, 0.06) if n else np.zeros((0, d))
        s = 1.0 if scales is None else scales[k]
        parts.append(X * s + 0.37 * k)
    off = np.zeros(len(sizes) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    return np.concatenate(parts, axis=0), off


@pytest.mark.parametrize("kernel", KERNELS)
def test_batch_self_matches_oracle_per_job(ctx, kernel):
    sizes = [0, 1, 2, 63, 64, 65, 500, 1000, 2000, 129]
    rng = np.random.default_rng(5)
    eps = rng.uniform(0.03, 0.3, size=len(sizes))
    scales = rng.uniform(0.5, 4.0, size=len(sizes))
    X, off = _jobs(1, sizes, 16, scales)
    print(f"batch self kernel={kernel} jobs={len(sizes)}", flush=True)
    counts, st = ctx.batch_self_join_count(X, off, eps, kernel=kernel)
    for k in range(len(sizes)):
        want = oracle.self_count(X[off[k]:off[k + 1]], eps[k]) if sizes[k] > 1 else 0
        assert counts[k] == want, (k, sizes[k], eps[k])
    assert counts.sum() > 0


@pytest.mark.parametrize("kernel", KERNELS)
def test_batch_join_matches_oracle_per_job(ctx, kernel):
    sa = [300, 0, 64, 1000, 5, 700]
    sb = [200, 50, 0, 1500, 65, 1]
    rng = np.random.default_rng(6)
    eps = rng.uniform(0.05, 0.3, size=len(sa))
    A, aoff = _jobs(2, sa, 12)
    B, boff = _jobs(2, sb, 12)  # same mixture centres as A, perturbed
    B = B + rng.normal(0.0, 0.02, size=B.shape)
    print(f"batch join kernel={kernel} jobs={len(sa)}", flush=True)
    counts, _ = ctx.batch_join_count(A, aoff, B, boff, eps, kernel=kernel)
    for k in range(len(sa)):
        a, b = A[aoff[k]:aoff[k + 1]], B[boff[k]:boff[k + 1]]
        want = oracle.join_count(a, b, eps[k]) if len(a) and len(b) else 0
        assert counts[k] == want, (k, eps[k])
    assert counts.sum() > 0


@pytest.mark.parametrize("kernel", ["fp64", "f16"])
def test_batch_equals_single_joins(ctx, kernel):
    # a job's count does not depend on its neighbours in the batch
    sizes = [3000, 40, 1200]
    X, off = _jobs(4, sizes, 24)
    eps = [0.15, 0.5, 0.2]
    counts, _ = ctx.batch_self_join_count(X, off, eps, kernel=kernel)
    for k in range(3):
        single, _ = ctx.self_join_count(X[off[k]:off[k + 1]], eps[k], kernel=kernel)
        assert counts[k] == single


def test_batch_many_jobs(ctx):
    # 300 jobs: 9 prefix bits above the cell keys
    rng = np.random.default_rng(7)
    sizes = rng.integers(0, 120, size=300)
    X, off = _jobs(5, sizes, 6)
    eps = rng.uniform(0.05, 0.2, size=300)
    counts, _ = ctx.batch_self_join_count(X, off, eps, kernel="f16")
    want = [oracle.self_count(X[off[k]:off[k + 1]], eps[k]) if sizes[k] > 1 else 0
            for k in range(300)]
    assert np.array_equal(counts, np.array(want, dtype=np.uint64))


@pytest.mark.parametrize("kernel", ["fp64", "f16"])
def test_batch_more_jobs_than_grid_rows(ctx, kernel):
    # 70000 jobs of 0-4 points: more segments than 65536 (the per-segment
    # kernels' grids must not put them on a 16-bit axis)
    rng = np.random.default_rng(9)
    sizes = rng.integers(0, 5, size=70000)
    off = np.zeros(sizes.size + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    X = rng.uniform(0.0, 1.0, size=(int(off[-1]), 4)) + np.repeat(
        np.arange(sizes.size) % 7, sizes)[:, None]
    eps = rng.uniform(0.2, 0.8, size=sizes.size)
    print(f"batch of {sizes.size} jobs kernel={kernel}", flush=True)
    counts, _ = ctx.batch_self_join_count(X, off, eps, kernel=kernel)
    want = np.array([oracle.self_count(X[off[k]:off[k + 1]], eps[k]) if sizes[k] > 1 else 0
                     for k in range(sizes.size)], dtype=np.uint64)
    assert np.array_equal(counts, want)
    assert want.sum() > 1000


def test_batch_device_inputs(ctx):
    sizes = [700, 900]
    X, off = _jobs(6, sizes, 10)
    A, aoff = _jobs(6, [400, 300], 10)
    A = A + np.random.default_rng(8).normal(0.0, 0.03, size=A.shape)
    dX, dA = ctx.to_device(X), ctx.to_device(A)
    got, _ = ctx.batch_self_join_count(dX, off, 0.2, kernel="f16")
    want, _ = ctx.batch_self_join_count(X, off, 0.2, kernel="f16")
    assert np.array_equal(got, want)
    got, _ = ctx.batch_join_count(dA, aoff, dX, off, 0.2, kernel="bf16")
    want = [oracle.join_count(A[aoff[k]:aoff[k + 1]], X[off[k]:off[k + 1]], 0.2)
            for k in range(2)]
    assert list(got) == want and sum(want) > 0


def test_batch_invalid_arguments(ctx):
    X, off = _jobs(8, [10, 10], 3)
    with pytest.raises(ValueError):
        ctx.batch_self_join_count(X, off[:-1], 0.1)
    bad = off.copy()
    bad[1] = 15
    bad[2] = 12
    with pytest.raises(hj.HJError):
        ctx.batch_self_join_count(X[:12], bad, 0.1)
    with pytest.raises(hj.HJError):
        ctx.batch_self_join_count(X, off, [0.1, -1.0])
    counts, _ = ctx.batch_self_join_count(np.zeros((0, 3)), np.zeros(1, dtype=np.uint64), [])
    assert counts.size == 0
