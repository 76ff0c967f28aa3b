"""GPU tests of the autotuner: tuned settings are valid, never change a count,
and calls with tune=True reuse the context's tuned settings."""
import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", ["fp64", "f16"])
def test_autotune_self_keeps_the_count(ctx, kernel):
    X = hj.gen_mixture(41, 6000, 24, 12, 0.05)
    eps = 0.15
    want = oracle.self_count(X, eps)
    assert want > 0
    dev = ctx.to_device(X)
    rep = ctx.autotune_self(dev, eps, kernel=kernel)
    print(f"tune kernel={kernel}: {rep}", flush=True)
    assert rep["trials"] >= 8 and rep["best_ms"] <= rep["base_ms"]
    best = rep["best"]
    assert best["kernel"] == kernel
    assert best["tile_rows"] in (8, 16) and best["mfma_chunk"] in (2, 4, 8)
    got, _ = ctx.self_join_count(dev, eps, **best)
    assert got == want
    # tune=True finds the cached settings for this shape class
    got, _ = ctx.self_join_count(dev, eps, kernel=kernel, tune=True)
    assert got == want
    dev.free()


def test_autotune_join_and_batch(ctx):
    A = hj.gen_mixture(42, 3000, 16, 10, 0.05)
    B = hj.gen_mixture(42, 5000, 16, 10, 0.05)[3000:] + 0.01
    dA, dB = ctx.to_device(A), ctx.to_device(B)
    rep = ctx.autotune_join(dA, dB, 0.12, kernel="bf16")
    got, _ = ctx.join_count(dA, dB, 0.12, **rep["best"])
    assert got == oracle.join_count(A, B, 0.12)
    sizes = [500, 1500, 64, 2000]
    off = np.zeros(5, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    P = hj.gen_mixture(43, int(off[-1]), 16, 10, 0.05)
    dP = ctx.to_device(P)
    rep = ctx.autotune_batch_self(dP, off, 0.1, kernel="f16")
    got, _ = ctx.batch_self_join_count(dP, off, 0.1, **rep["best"])
    want = [oracle.self_count(P[off[k]:off[k + 1]], 0.1) for k in range(4)]
    assert list(got) == want
    got, _ = ctx.batch_self_join_count(dP, off, 0.1, kernel="f16", tune=True)
    assert list(got) == want


@pytest.mark.parametrize("tile_rows,kblock", [(8, 4), (8, 8), (16, 4)])
@pytest.mark.parametrize("chunk", [2, 8])
def test_tile_shape_variants_keep_the_count(ctx, tile_rows, kblock, chunk):
    X = hj.gen_mixture(44, 4000, 20, 8, 0.06)
    eps = 0.2
    want = oracle.self_count(X, eps)
    got, _ = ctx.self_join_count(X, eps, kernel="fp64", tile_rows=tile_rows, kblock=kblock)
    assert got == want
    got, _ = ctx.self_join_count(X, eps, kernel="f16", mfma_chunk=chunk)
    assert got == want
    Y = X[:1500] + 0.01
    got, _ = ctx.join_count(X, Y, eps, kernel="bf16", mfma_chunk=chunk, grid=512)
    assert got == oracle.join_count(X, Y, eps)


def test_candidate_pairs_agree_across_kernels(ctx):
    # the filter counts its candidates from the work list, the FP64 kernel
    # while it runs; both describe the same list (the active dimensions are
    # fixed: left open, each kernel weighs its own cost per tile pair)
    for n, d, eps in [(3001, 28, 0.3), (130, 5, 0.4), (4096, 64, 2.6)]:
        X = hj.gen_mixture(n, n, d, 6, 0.05)
        for a in (1, 3):
            st = {k: ctx.self_join_count(X, eps, kernel=k, active_dims=a)[1]
                  for k in ("fp64", "f16", "bf16")}
            assert st["fp64"].candidate_pairs == st["f16"].candidate_pairs == \
                st["bf16"].candidate_pairs > 0
            Y = X[: n // 2] + 0.01
            sj = {k: ctx.join_count(X, Y, eps, kernel=k, active_dims=a)[1]
                  for k in ("fp64", "f16")}
            assert sj["fp64"].candidate_pairs == sj["f16"].candidate_pairs > 0


def test_active_choice_follows_the_estimated_cost(ctx):
    # left open, the active dimensions minimize the estimated work-list cost:
    # the result is one of the fixed choices, and its tile pairs are those of
    # that fixed choice
    X = hj.gen_uniform(9, 60000, 16)
    for kernel in ("fp64", "f16"):
        c_auto, st = ctx.self_join_count(X, 0.35, kernel=kernel)
        assert 0 <= st.active_dims <= 6
        c_fix, st_fix = ctx.self_join_count(X, 0.35, kernel=kernel, active_dims=st.active_dims)
        assert c_auto == c_fix
        assert st.tile_pairs == st_fix.tile_pairs
        assert st.candidate_pairs == st_fix.candidate_pairs


@pytest.mark.parametrize("n,d,eps", [(5000, 20, 0.1), (3001, 7, 0.05), (64, 130, 0.3)])
def test_dim_statistics_and_reorder(ctx, n, d, eps):
    rng = np.random.default_rng(n + d)
    X = rng.random((n, d)) * rng.uniform(0.1, 3.0, size=d)  # dimensions of different spread
    order, cost = ctx.dim_statistics(X, eps)
    want = oracle.dim_costs(X, eps)
    assert np.array_equal(cost, want)
    # reorder_dim[k] = rank of dimension k by increasing cost, ties by index
    ranks = np.empty(d, dtype=np.int64)
    ranks[np.argsort(want, kind="stable")] = np.arange(d)
    assert np.array_equal(order, ranks)
    Y = ctx.reorder_dimensions(X, order)
    Z = np.empty_like(X)
    Z[:, order] = X
    assert np.array_equal(Y, Z)
    # the most selective dimension comes first
    assert oracle.dim_costs(Y, eps)[0] == want.min()


def test_reference_reorder_shims(ctx):
    # outputStatistics / reorder_dimensions by their C++ names (egojoin.h:59-60)
    import ctypes
    lib = ctypes.CDLL(str(hj.library_path()))
    stats = getattr(lib, "_Z16outputStatisticsiidPdPi")
    reorder = getattr(lib, "_Z18reorder_dimensionsiiPdPi")
    stats.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                      ctypes.c_void_p]
    reorder.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(9)
    d = 21  # d % 8 != 0: the reference leaves the last 5 columns unmoved
    X = np.ascontiguousarray(rng.random((2000, d)) * rng.uniform(0.1, 2.0, size=d))
    rd = np.zeros(d + 8, dtype=np.int32)
    stats(2000, d, 0.1, X.ctypes.data, rd.ctypes.data)
    assert list(rd[d:]) == list(range(d, d + 8))
    assert sorted(rd[:d]) == list(range(d))
    Y = X.copy()
    reorder(2000, d, Y.ctypes.data, rd.ctypes.data)
    Z = np.empty_like(X)
    Z[:, rd[:d]] = X
    assert np.array_equal(Y, Z)


def test_filter_counts_its_mfmas(ctx):
    # the filter reports the K slots its MFMAs multiplied: 16 x 32 x 32 per
    # MFMA, at least four per non-diagonal tile pair (the first K chunk)
    X = hj.gen_uniform(3, 20000, 40)
    for kernel in ("f16", "bf16"):
        _, st = ctx.self_join_count(X, 0.6, kernel=kernel)
        assert st.pair_dims % (16 * 1024) == 0
        mfmas = st.pair_dims // (16 * 1024)
        assert mfmas >= 4 * (st.tile_pairs - (20000 + 63) // 64)
        assert st.early_exits <= st.tile_pairs
