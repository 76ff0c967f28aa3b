"""Tile frames of the MFMA filter on the HIP path (hj_opts.filter_frames).

Tiles whose points lie close together take their own operand frame (an anchor
near the tile's centre) and a tile pair is tested against its own error band;
the band then follows the tiles' extent, not the join's.  Counts and pair
lists must stay identical to the FP64 kernel's (and the oracle's, where it
finishes in seconds) with frames on and off, on both filters, in the slow path
only (FP16: narrow grid-frame band) and in the fast path too (BF16: wide band),
for self, A x B, batched and pair-list joins, with planted pairs at
eps (1 +- 2^-k) and with framed and grid-frame tiles in one join.
"""
import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle
import planted

pytestmark = pytest.mark.gpu

FILTERS = ["f16", "bf16"]


def _counts(ctx, dev, eps, **kw):
    c64, _ = ctx.self_join_count(dev, eps, kernel="fp64", **kw)
    out = {}
    for kernel in FILTERS:
        for frames in (False, True):
            c, st = ctx.self_join_count(dev, eps, kernel=kernel, filter_frames=frames, **kw)
            out[(kernel, frames)] = (c, st)
    return c64, out


@pytest.mark.parametrize("d,sigma,eps,ncomp", [(8, 0.01, 0.03, 32), (16, 0.02, 0.06, 64),
                                               (28, 0.01, 0.05, 64), (64, 0.004, 0.04, 16)])
def test_frames_counts_match_fp64_on_clusters(ctx, d, sigma, eps, ncomp):
    X = hj.gen_mixture(11 + d, 200_000, d, ncomp, sigma)
    dev = ctx.to_device(X)
    c64, out = _counts(ctx, dev, eps)
    assert c64 > 1000
    for (kernel, frames), (c, st) in out.items():
        print(f"  {kernel} frames={frames}: {c} band={st.band_pairs} mode={st.frames}", flush=True)
        assert c == c64, (kernel, frames)
        if not frames:
            assert st.frames == 0
    for kernel in FILTERS:
        on, off = out[(kernel, True)][1], out[(kernel, False)][1]
        assert on.frames >= 1  # compact clusters: tiles took their own frames
        assert on.band_pairs < off.band_pairs / 4  # the band follows the tiles
    if d == 28:
        # FP16's band (0.09 of eps here) leaves the fast path in the grid frame,
        # BF16's (0.4) moves it to the tiles' frames too
        assert out[("f16", True)][1].frames == 1 and out[("bf16", True)][1].frames == 2
    dev.free()


def test_frames_exact_against_brute_force_with_planted_pairs(ctx):
    eps = 0.05
    X = hj.gen_mixture(21, 12_000, 28, 8, 0.012)
    Xp, pairs, inside = planted.plant_self(X, eps, 200, seed=2101)
    want = oracle.self_count(Xp, eps)
    dev = ctx.to_device(Xp)
    for kernel in FILTERS:
        c, st = ctx.self_join_count(dev, eps, kernel=kernel)
        assert st.frames >= 1 and c == want, kernel
    # the pair lists: every planted pair within eps, nothing outside it
    rp, cols, _ = ctx.self_join_csr(dev, eps, kernel="f16")
    listed = planted.csr_pairs(rp.numpy().ravel(), cols.numpy().ravel())
    assert len(listed) == 2 * want
    und = {(i, j) for i, j in listed if i < j}
    planted.verify_listed(und, Xp, Xp, eps)
    assert {tuple(sorted(p)) for p, ok in zip(pairs.tolist(), inside) if ok} <= und
    dev.free()


def test_frames_with_grid_frame_tiles_in_the_same_join(ctx):
    # tight clusters (framed tiles) in a uniform background (tiles that keep
    # the grid frame), with pairs across both kinds
    rng = np.random.default_rng(31)
    C = hj.gen_mixture(31, 60_000, 16, 24, 0.01)
    U = rng.random((60_000, 16))
    X = np.concatenate([C, U])[rng.permutation(120_000)]
    eps = 0.04
    dev = ctx.to_device(X)
    c64, out = _counts(ctx, dev, eps)
    for (kernel, frames), (c, st) in out.items():
        assert c == c64, (kernel, frames)
    assert out[("f16", True)][1].frames >= 1
    assert c64 == oracle.ego_self_count(X, eps, active_dims=3)[0]
    dev.free()


def test_frames_join_a_x_b_and_symmetry(ctx):
    A = hj.gen_mixture(41, 150_000, 28, 32, 0.01)
    B = hj.gen_mixture(42, 120_000, 28, 32, 0.01)  # other clusters, some overlapping
    B[:60_000] = A[:60_000] + np.random.default_rng(43).normal(scale=0.01, size=(60_000, 28))
    eps = 0.05
    dA, dB = ctx.to_device(A), ctx.to_device(B)
    c64, _ = ctx.join_count(dA, dB, eps, kernel="fp64")
    assert c64 > 1000
    for kernel in FILTERS:
        c, st = ctx.join_count(dA, dB, eps, kernel=kernel)
        c2, _ = ctx.join_count(dB, dA, eps, kernel=kernel)
        assert st.frames >= 1 and c == c2 == c64, kernel
    rp, cols, _ = ctx.join_csr(dA, dB, eps, kernel="bf16")
    listed = planted.csr_pairs(rp.numpy().ravel(), cols.numpy().ravel())
    rp64, cols64, _ = ctx.join_csr(dA, dB, eps, kernel="fp64")
    assert listed == planted.csr_pairs(rp64.numpy().ravel(), cols64.numpy().ravel())
    dA.free()
    dB.free()


def test_frames_batch_of_clustered_joins(ctx):
    sizes = [3000, 17000, 40000, 800, 25000, 9000]
    jobs = [hj.gen_mixture(50 + k, n, 16, 6, 0.012) for k, n in enumerate(sizes)]
    off = np.zeros(len(sizes) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    dev = ctx.to_device(np.concatenate(jobs))
    eps = 0.045
    c64, _ = ctx.batch_self_join_count(dev, off, eps, kernel="fp64")
    for kernel in FILTERS:
        c, st = ctx.batch_self_join_count(dev, off, eps, kernel=kernel)
        assert np.array_equal(c, c64), kernel
        assert st.frames >= 1
    for k in (0, 3):
        assert c64[k] == oracle.self_count(jobs[k], eps)
    dev.free()


def test_frames_off_where_they_cannot_apply(ctx):
    # d > 64 (tiles not staged in FP64) and a first K chunk of two k-steps keep
    # the grid frame; counts unchanged
    X = hj.gen_mixture(61, 40_000, 72, 8, 0.005)
    dev = ctx.to_device(X)
    c64, _ = ctx.self_join_count(dev, 0.05, kernel="fp64")
    c, st = ctx.self_join_count(dev, 0.05, kernel="f16")
    assert c == c64 and st.frames == 0
    dev.free()
    Y = hj.gen_mixture(62, 40_000, 40, 8, 0.006)
    dev = ctx.to_device(Y)
    c64, _ = ctx.self_join_count(dev, 0.05, kernel="fp64")
    c, st = ctx.self_join_count(dev, 0.05, kernel="f16", filter_split=28)
    assert c == c64 and st.frames == 0
    c, st = ctx.self_join_count(dev, 0.05, kernel="f16")
    assert c == c64 and st.frames >= 1
    dev.free()
