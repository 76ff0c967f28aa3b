"""The MFMA filter's one-row chunks (row-order lists, FP16).

In a row-order work list the tile pairs of a row tile are consecutive, so a
chunk whose first and last pair share the row tile is tested with one vote for
the whole chunk, and pair by pair only when some pair may be near (a diagonal
tile pair always is).  Chunks across a row-tile boundary, the last partial
chunk and Hilbert-ordered lists take the per-pair path.  Counts and pair lists
must not depend on the path a chunk took: row order (both paths) against
Hilbert order (per-pair path only) and the oracle, for every chunk size, for
self and A x B joins, for batches of jobs of a few tiles each (chunks across
segments) and with planted pairs at eps (1 +- 2^-k).
"""
import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle
import planted

pytestmark = pytest.mark.gpu

CHUNKS = [2, 4, 8]


@pytest.mark.parametrize("chunk", CHUNKS)
def test_one_row_chunks_self_count(ctx, chunk):
    eps = 0.12
    X = hj.gen_mixture(71, 20_000, 24, 16, 0.05)
    Xp, _, inside = planted.plant_self(X, eps, 200, seed=7100 + chunk)
    want = oracle.self_count(Xp, eps)
    assert want >= inside.sum() > 50
    dev = ctx.to_device(Xp)
    for order in ("row", "hilbert"):
        got, st = ctx.self_join_count(dev, eps, kernel="f16", mfma_chunk=chunk, order=order)
        print(f"  chunk={chunk} order={order}: {got} band={st.band_pairs}", flush=True)
        assert got == want, order
    dev.free()


@pytest.mark.parametrize("chunk", CHUNKS)
def test_one_row_chunks_join_csr(ctx, chunk):
    eps = 0.15
    A = hj.gen_mixture(72, 12_000, 20, 12, 0.06)
    B = hj.gen_mixture(73, 9_000, 20, 12, 0.06)
    Ap, Bp, pairs, inside = planted.plant_join(A, B, eps, 150, seed=7200 + chunk)
    want = oracle.join_count(Ap, Bp, eps)
    dA, dB = ctx.to_device(Ap), ctx.to_device(Bp)
    lists = {}
    for order in ("row", "hilbert"):
        got, _ = ctx.join_count(dA, dB, eps, kernel="f16", mfma_chunk=chunk, order=order)
        assert got == want, order
        rp, cols, _ = ctx.join_csr(dA, dB, eps, kernel="f16", mfma_chunk=chunk, order=order)
        lists[order] = planted.csr_pairs(rp.numpy().ravel(), cols.numpy().ravel())
        assert len(lists[order]) == want
    assert lists["row"] == lists["hilbert"]
    planted.verify_listed(lists["row"], Ap, Bp, eps)
    assert {tuple(p) for p, ok in zip(pairs.tolist(), inside) if ok} <= lists["row"]
    dA.free()
    dB.free()


@pytest.mark.parametrize("chunk", CHUNKS)
def test_one_row_chunks_batch_of_small_jobs(ctx, chunk):
    # jobs of 1..320 points: one to five tiles each, so chunks span segments
    rng = np.random.default_rng(74)
    sizes = rng.integers(1, 321, size=150)
    eps = 0.3
    jobs = [hj.gen_mixture(7400 + k, int(n), 12, 3, 0.08) for k, n in enumerate(sizes)]
    off = np.zeros(len(sizes) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    dev = ctx.to_device(np.concatenate(jobs))
    want = np.array([oracle.self_count(J, eps) for J in jobs])
    assert want.sum() > 1000
    for order in ("row", "hilbert"):
        got, _ = ctx.batch_self_join_count(dev, off, eps, kernel="f16", mfma_chunk=chunk,
                                           order=order)
        assert np.array_equal(got, want), order
    dev.free()
