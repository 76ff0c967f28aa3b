"""GPU parity: the HIP join (through the C ABI) against the FP64 oracle.

Counts must be bit-exact: both sides evaluate the same fixed-order FP64
predicate.  Sizes are small enough for the brute-force oracle to finish in
seconds.
"""
import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle

pytestmark = pytest.mark.gpu

KERNELS = ["fp64", "f16", "bf16"]

# (n, d, eps, seed): planted dense cases with many pairs, ragged n, tiny n
SELF_CASES = [
    (3000, 8, 0.3, 1),
    (4000, 64, 2.6, 2),
    (2001, 3, 0.05, 3),
    (1003, 1, 0.001, 4),
    (5000, 20, 0.9, 5),
    (4096, 28, 0.7, 6),
    (130, 5, 0.5, 7),
    (64, 5, 0.6, 8),
    (63, 5, 0.6, 9),
    (65, 5, 0.6, 10),
    (2, 3, 10.0, 11),
    (7, 2, 0.4, 12),
    (3000, 130, 3.9, 13),
]


def _uniform(seed, n, d):
    return hj.gen_uniform(seed, n, d)


@pytest.mark.parametrize("n,d,eps,seed", SELF_CASES)
@pytest.mark.parametrize("kernel", KERNELS)
def test_self_count_matches_oracle(ctx, n, d, eps, seed, kernel):
    X = _uniform(seed, n, d)
    want = oracle.self_count(X, eps)
    print(f"self n={n} d={d} eps={eps} seed={seed} kernel={kernel}", flush=True)
    got, st = ctx.self_join_count(X, eps, kernel=kernel)
    assert got == want
    assert st.candidate_pairs <= n * (n - 1) // 2


@pytest.mark.parametrize("active", [0, 1, 2, 3, 4, 6])
@pytest.mark.parametrize("order", ["hilbert", "row"])
@pytest.mark.parametrize("kblock", [4, 8])
def test_self_count_independent_of_knobs(ctx, active, order, kblock):
    X = hj.gen_mixture(21, 6000, 12, 16, 0.05)
    eps = 0.08
    want = oracle.self_count(X, eps)
    assert want > 0
    got, _ = ctx.self_join_count(X, eps, active_dims=active, order=order, kblock=kblock)
    assert got == want


@pytest.mark.parametrize("kernel", ["f16", "bf16"])
@pytest.mark.parametrize("split", [-1, 0, 3, 10, 12, 30])
@pytest.mark.parametrize("n,d,eps", [(3000, 64, 2.6), (4000, 28, 0.5), (2500, 40, 1.2)])
def test_filter_chunk_split(ctx, kernel, split, n, d, eps):
    # the early exit after the first K chunk never changes the count
    X = hj.gen_mixture(n + d, n, d, 8, 0.08)
    want = oracle.self_count(X, eps)
    got, st = ctx.self_join_count(X, eps, kernel=kernel, filter_split=split)
    assert got == want
    Y = hj.gen_mixture(n + d + 1, n // 2, d, 8, 0.08)
    got, _ = ctx.join_count(X, Y, eps, kernel=kernel, filter_split=split)
    assert got == oracle.join_count(X, Y, eps)


@pytest.mark.parametrize("nA,nB,d,eps", [(3000, 2000, 8, 0.3), (1000, 5000, 20, 0.9),
                                         (70, 63, 4, 0.4), (1, 500, 3, 0.5), (4000, 4000, 64, 2.6)])
@pytest.mark.parametrize("kernel", KERNELS)
def test_join_count_matches_oracle(ctx, nA, nB, d, eps, kernel):
    A = _uniform(100 + nA, nA, d)
    B = _uniform(200 + nB, nB, d)
    want = oracle.join_count(A, B, eps)
    print(f"join nA={nA} nB={nB} d={d} eps={eps} kernel={kernel}", flush=True)
    got, _ = ctx.join_count(A, B, eps, kernel=kernel)
    assert got == want
    got2, _ = ctx.join_count(B, A, eps, kernel=kernel)
    assert got2 == want


@pytest.mark.parametrize("active", [0, 2, 4])
def test_join_count_active_dims(ctx, active):
    A = hj.gen_mixture(31, 5000, 10, 8, 0.05)
    B = hj.gen_mixture(32, 4000, 10, 8, 0.05)
    eps = 0.1
    want = oracle.join_count(A, B, eps)
    got, _ = ctx.join_count(A, B, eps, active_dims=active)
    assert got == want


def _random_case(seed):
    # seeded shapes across the kernels' regimes: tiny and ragged n, d from 1
    # to 70, uniform or clustered data with an offset and scale, eps from the
    # distances of random point pairs so that counts are neither 0 nor all
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.choice([1, 2, 63, 64, 65, 200, 777, 1500, 3000]))
    m = int(rng.choice([1, 64, 129, 900, 2500]))
    d = int(rng.choice([1, 2, 3, 5, 8, 13, 20, 28, 33, 47, 64, 70]))
    if rng.random() < 0.5:
        X = rng.random((n + m, d))
    else:
        X = hj.gen_mixture(int(rng.integers(1 << 30)), n + m, d, int(rng.integers(2, 12)), 0.05)
    X = X * float(rng.choice([1.0, 13.0, 0.01])) + float(rng.choice([0.0, -50.0, 1e3]))
    # eps at a low quantile of the distances of random point pairs
    i, j = rng.integers(0, n + m, size=(2, 4000))
    dist = np.sqrt(((X[i] - X[j]) ** 2).sum(axis=1))
    eps = float(np.quantile(dist[i != j], rng.uniform(0.002, 0.05))) if (i != j).any() else 1.0
    return X[:n], X[n:], max(eps, 1e-9)


@pytest.mark.parametrize("seed", range(24))
def test_random_shapes_match_oracle(ctx, seed):
    A, B, eps = _random_case(seed)
    print(f"random case seed={seed} n={len(A)} m={len(B)} d={A.shape[1]} eps={eps:.4g}",
          flush=True)
    want_self = oracle.self_count(A, eps)
    want_join = oracle.join_count(A, B, eps)
    for kernel in KERNELS:
        got, _ = ctx.self_join_count(A, eps, kernel=kernel)
        assert got == want_self, (kernel, got, want_self)
        got, _ = ctx.join_count(A, B, eps, kernel=kernel)
        assert got == want_join, (kernel, got, want_join)


@pytest.mark.parametrize("seed", range(0, 24, 2))
def test_random_shapes_csr_match_oracle(ctx, seed):
    A, B, eps = _random_case(seed)
    print(f"random csr seed={seed} n={len(A)} m={len(B)} d={A.shape[1]}", flush=True)
    wsr, wsc = oracle.self_csr(A, eps)
    wjr, wjc = oracle.join_csr(A, B, eps)
    for kernel in KERNELS:
        rp, cols, _ = ctx.self_join_csr(A, eps, kernel=kernel)
        assert np.array_equal(rp, wsr) and np.array_equal(cols, wsc), kernel
        rp, cols, _ = ctx.join_csr(A, B, eps, kernel=kernel)
        assert np.array_equal(rp, wjr) and np.array_equal(cols, wjc), kernel


def test_identical_points_closed_form(ctx):
    # all pairs within eps: the reference drops the n % 8 tail (499500 for
    # n = 1003, egojoin.h:325); the full count is n (n - 1) / 2
    n = 1003
    X = np.tile(np.array([[0.25, 0.5, 0.75]]), (n, 1))
    got, _ = ctx.self_join_count(X, 0.1)
    assert got == n * (n - 1) // 2
    got, _ = ctx.join_count(X, X, 0.1)
    assert got == n * n


def test_self_equals_half_join_minus_n(ctx):
    X = _uniform(41, 3000, 6)
    eps = 0.2
    s, _ = ctx.self_join_count(X, eps)
    j, _ = ctx.join_count(X, X, eps)
    assert j == 2 * s + X.shape[0]


def test_large_clustered_join_identities(ctx):
    # 4M clustered points with tens of millions of pairs (sizes the brute-force
    # oracle cannot check): every kernel gives the same self count, and the
    # A x A join equals twice the self count plus the n diagonal pairs
    X = ctx.to_device(hj.gen_mixture(89, 4_000_000, 16, 64, 0.02))
    eps = 0.045
    counts = {}
    for kernel in KERNELS:
        print(f"large clustered self count kernel={kernel}", flush=True)
        counts[kernel], st = ctx.self_join_count(X, eps, kernel=kernel)
    assert len(set(counts.values())) == 1 and counts["fp64"] > 10_000_000
    print("large clustered A x A join f16", flush=True)
    j, _ = ctx.join_count(X, X, eps, kernel="f16")
    assert j == 2 * counts["fp64"] + 4_000_000
    X.free()


def test_lattice_closed_form(ctx):
    # integer lattice k^3 with spacing 1 and eps = 1: each point pairs with its
    # axis neighbours only -> 3 k^2 (k-1) pairs
    k = 14
    g = np.arange(k, dtype=np.float64)
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    got, _ = ctx.self_join_count(X, 1.0)
    assert got == 3 * k * k * (k - 1)


def test_band_pairs_exact(ctx):
    # pairs planted at distance eps (1 +- 2^-k) straddling cell boundaries
    rng = np.random.default_rng(5)
    eps = 0.125
    d = 8
    base = rng.random((400, d)) * 0.9
    base[:, :4] = np.round(base[:, :4] / eps) * eps  # on cell boundaries
    dirs = rng.normal(size=(400, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ks = rng.integers(20, 52, size=400)
    sign = rng.choice([-1.0, 1.0], size=400)
    r = eps * (1 + sign * np.ldexp(1.0, -ks))
    X = np.concatenate([base, base + dirs * r[:, None]])
    want = oracle.self_count(X, eps)
    for kernel in KERNELS:
        got, st = ctx.self_join_count(X, eps, kernel=kernel)
        assert got == want, kernel
        if kernel != "fp64":
            assert st.band_pairs > 0, kernel  # the FP64 recheck decided them


def test_negative_and_offset_coordinates(ctx):
    X = _uniform(51, 3000, 5) * 7.0 - 1000.0
    eps = 0.9
    want = oracle.self_count(X, eps)
    got, _ = ctx.self_join_count(X, eps)
    assert got == want


def test_ego_sort_order(ctx):
    X = _uniform(61, 5000, 6)
    eps = 0.2
    perm, keys, kd, kb = ctx.ego_sort(X, eps)
    assert kd == 6 and kb >= 3
    assert sorted(perm.tolist()) == list(range(5000))
    assert np.all(np.diff(keys.astype(np.uint64)) >= 0) if keys.size else True
    want = oracle.ego_sort(X, eps)  # stable lexicographic cell order
    assert np.array_equal(perm.astype(np.int64), want)


def test_ego_sort_prefix_key(ctx):
    # d = 64 at eps = 0.2 does not fit 64 bits: the key holds a prefix of the
    # dimensions and the order is lexicographic on that prefix, ties by index
    X = _uniform(62, 3000, 64)
    eps = 0.2
    perm, keys, kd, kb = ctx.ego_sort(X, eps)
    assert kd < 64
    cells = np.floor(X[:, :kd] * (1.0 / eps)).astype(np.int64)
    order = np.lexsort(tuple(cells[:, k] for k in range(kd - 1, -1, -1)))
    assert np.array_equal(perm.astype(np.int64), order)


@pytest.mark.parametrize("n", [1_500_000, 4_500_000])
def test_ego_sort_many_tiles(ctx, n):
    # a thousand and more sort tiles per digit (both tile sizes of the radix
    # sort): the look-back between tiles must give the stable lexicographic
    # cell order; a coarse grid leaves long runs of equal keys (ties by index)
    rng = np.random.default_rng(63)
    X = rng.random((n, 3))
    eps = 0.05
    perm, keys, kd, kb = ctx.ego_sort(X, eps)
    assert kd == 3
    cells = np.floor(X * (1.0 / eps)).astype(np.int64)
    order = np.lexsort((cells[:, 2], cells[:, 1], cells[:, 0]))
    assert np.array_equal(perm.astype(np.int64), order)
    assert np.all(np.diff(keys.astype(np.uint64)) >= 0)


def test_worklist_covers_every_pair(ctx):
    X = hj.gen_mixture(71, 4000, 10, 8, 0.06)
    eps = 0.12
    pairs, perm = ctx.self_tile_pairs(X, eps)
    have = set(map(tuple, pairs.tolist()))
    assert all(I <= J for I, J in have)
    assert len(have) == len(pairs)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size, dtype=perm.dtype)
    rowptr, cols = oracle.self_csr(X, eps)
    for i in range(X.shape[0]):
        for j in cols[rowptr[i]:rowptr[i + 1]]:
            a, b = sorted((int(inv[i]), int(inv[j])))
            assert (a // 64, b // 64) in have


def test_worklist_hilbert_order(ctx):
    X = hj.gen_mixture(72, 20000, 8, 16, 0.05)
    pairs, _ = ctx.self_tile_pairs(X, 0.05, order="hilbert")
    side = int(np.ceil((X.shape[0]) / 64))
    L = max(0, int(np.ceil(np.log2(side))))
    h = [oracle.hilbert_xy2d(L, int(I), int(J)) for I, J in pairs]
    assert all(a <= b for a, b in zip(h, h[1:]))


def test_hilbert_index_matches_textbook(ctx):
    rng = np.random.default_rng(3)
    for order in (1, 3, 10, 16):
        x = rng.integers(0, 1 << order, size=2000, dtype=np.uint32)
        y = rng.integers(0, 1 << order, size=2000, dtype=np.uint32)
        got = ctx.hilbert_index(order, x, y)
        want = [oracle.hilbert_xy2d(order, int(a), int(b)) for a, b in zip(x, y)]
        assert got.tolist() == want


def test_self_csr_matches_oracle(ctx):
    X = _uniform(81, 3000, 8)
    eps = 0.3
    rp, cols, _ = ctx.self_join_csr(X, eps)
    wrp, wcols = oracle.self_csr(X, eps)
    assert np.array_equal(rp, wrp)
    assert np.array_equal(cols, wcols)


def test_join_csr_matches_oracle(ctx):
    A = _uniform(82, 2000, 8)
    B = _uniform(83, 2500, 8)
    eps = 0.3
    rp, cols, _ = ctx.join_csr(A, B, eps)
    wrp, wcols = oracle.join_csr(A, B, eps)
    assert np.array_equal(rp, wrp)
    assert np.array_equal(cols, wcols)


@pytest.mark.parametrize("kernel", ["f16", "bf16"])
def test_filter_csr_clustered_matches_oracle(ctx, kernel):
    # four tight clusters: most pairs within eps, many of them in the filter's
    # band, so each wavefront stages and flushes both of its queues (certain
    # pairs and band pairs) several times
    X = hj.gen_mixture(85, 6000, 16, 4, 0.02)
    eps = 0.06
    rp, cols, st = ctx.self_join_csr(X, eps, kernel=kernel)
    wrp, wcols = oracle.self_csr(X, eps)
    assert st.band_pairs > 0
    assert np.array_equal(rp, wrp)
    assert np.array_equal(cols, wcols)
    A, B = X[:2500], hj.gen_mixture(86, 3000, 16, 4, 0.02)
    rp, cols, _ = ctx.join_csr(A, B, eps, kernel=kernel)
    wrp, wcols = oracle.join_csr(A, B, eps)
    assert np.array_equal(rp, wrp)
    assert np.array_equal(cols, wcols)


@pytest.mark.parametrize("kernel", ["f16", "bf16"])
def test_filter_csr_band_queue_overflow(ctx, kernel):
    # integers on a line at eps = 1 (as in the count test below): every pair
    # is a band pair and the queue overflows its initial capacity during the
    # pair pass too; row i lists i - 1 and i + 1 (closed form)
    n = 300_000
    X = np.arange(n, dtype=np.float64)[:, None]
    rp, cols, _ = ctx.self_join_csr(X, 1.0, kernel=kernel)
    deg = np.full(n, 2, dtype=np.int64)
    deg[0] = deg[-1] = 1
    assert np.array_equal(rp, np.concatenate([[0], np.cumsum(deg)]))
    i = np.arange(n, dtype=np.int64)
    want = np.stack([i - 1, i + 1], axis=1).ravel()
    want = want[(want >= 0) & (want < n)]
    assert np.array_equal(cols, want)


@pytest.mark.parametrize("kernel", ["fp64", "f16"])
def test_device_csr_matches_host(ctx, kernel):
    # device-resident lists (points in, rowptr / colidx out, all in HBM) equal
    # the host-memory lists; an empty result and a single point give rowptr 0
    X = hj.gen_mixture(87, 5000, 12, 6, 0.03)
    eps = 0.08
    hr, hc, _ = ctx.self_join_csr(X, eps, kernel=kernel)
    dX = ctx.to_device(X)
    dr, dc, _ = ctx.self_join_csr(dX, eps, kernel=kernel)
    assert dr.shape == (5001, 1) and dc.shape == (hc.size, 1)
    assert np.array_equal(dr.numpy().ravel(), hr)
    assert np.array_equal(dc.numpy().ravel(), hc)
    B = X[1000:4000] + 0.01  # shifted copies: partners within eps
    hr, hc, _ = ctx.join_csr(X, B, eps, kernel=kernel)
    dB = ctx.to_device(B)
    dr, dc, _ = ctx.join_csr(dX, dB, eps, kernel=kernel)
    assert np.array_equal(dr.numpy().ravel(), hr)
    assert np.array_equal(dc.numpy().ravel(), hc)
    assert hc.size > 0
    far = ctx.to_device(X[:100] + 50.0)
    dr, dc, _ = ctx.join_csr(far, dB, eps, kernel=kernel)
    assert dc.shape == (0, 1) and not dr.numpy().any()
    one = ctx.to_device(X[:1])
    dr, dc, _ = ctx.self_join_csr(one, eps, kernel=kernel)
    assert dr.numpy().ravel().tolist() == [0, 0] and dc.shape == (0, 1)


def test_device_inputs_match_host(ctx):
    X = _uniform(91, 5000, 16)
    Y = _uniform(92, 4000, 16)
    eps = 0.55
    dX, dY = ctx.to_device(X), ctx.to_device(Y)
    assert np.array_equal(dX.numpy(), X)
    h, _ = ctx.self_join_count(X, eps)
    dev, _ = ctx.self_join_count(dX, eps)
    assert h == dev == oracle.self_count(X, eps)
    h, _ = ctx.join_count(X, Y, eps)
    dev, _ = ctx.join_count(dX, dY, eps)
    assert h == dev == oracle.join_count(X, Y, eps)
    dX.free()
    dY.free()


def test_repeated_calls_reuse_buffers(ctx):
    # larger, then smaller, then larger inputs on one context
    for n, d, eps in [(6000, 10, 0.35), (500, 3, 0.05), (6000, 10, 0.35), (70, 70, 2.0)]:
        X = _uniform(n + d, n, d)
        got, _ = ctx.self_join_count(X, eps)
        assert got == oracle.self_count(X, eps)


def test_invalid_arguments_raise(ctx):
    X = _uniform(1, 100, 3)
    with pytest.raises(hj.HJError):
        ctx.self_join_count(X, 0.0)
    with pytest.raises(hj.HJError):
        ctx.self_join_count(X, -1.0)
    Y = X.copy()
    Y[5, 1] = np.nan
    with pytest.raises(hj.HJError):
        ctx.self_join_count(Y, 0.1)
    with pytest.raises(hj.HJError):
        ctx.self_join_count(X, 0.1, kblock=3)
    # the context stays usable after errors
    got, _ = ctx.self_join_count(X, 0.3)
    assert got == oracle.self_count(X, 0.3)


def test_empty_inputs(ctx):
    E = np.zeros((0, 4))
    X = _uniform(2, 10, 4)
    assert ctx.self_join_count(E, 0.1)[0] == 0
    assert ctx.self_join_count(X[:1], 0.1)[0] == 0
    assert ctx.join_count(E, X, 0.1)[0] == 0
    assert ctx.join_count(X, E, 0.1)[0] == 0


def _fixture(name):
    import json
    from pathlib import Path
    g = Path(__file__).resolve().parent / "golden"
    man = json.loads((g / "manifest.json").read_text())
    fx = next(f for f in man["fixtures"] if f["name"] == name)
    return g, fx


@pytest.mark.parametrize("name", ["self_uniform_d8", "self_dense_d64", "self_band_d6",
                                  "self_dup_ragged_1003", "self_lowd_d1"])
@pytest.mark.parametrize("kernel", KERNELS)
def test_golden_self_fixtures(ctx, name, kernel):
    g, fx = _fixture(name)
    X = np.fromfile(g / fx["a"], dtype=np.float64).reshape(-1, fx["d"])
    got, _ = ctx.self_join_count(X, fx["eps"], kernel=kernel)
    assert got == fx["count"]


@pytest.mark.parametrize("name", ["join_d20", "join_d3_ragged"])
@pytest.mark.parametrize("kernel", KERNELS)
def test_golden_join_fixtures(ctx, name, kernel):
    g, fx = _fixture(name)
    A = np.fromfile(g / fx["a"], dtype=np.float64).reshape(-1, fx["d"])
    B = np.fromfile(g / fx["b"], dtype=np.float64).reshape(-1, fx["d"])
    got, _ = ctx.join_count(A, B, fx["eps"], kernel=kernel)
    assert got == fx["count"]


def _run_tool(tool, *args):
    import subprocess
    from pathlib import Path
    root = Path(__file__).resolve().parent.parent
    out = subprocess.run([str(root / "hilbertjoin_amd/bin" / tool), *map(str, args)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout.strip().split(";")


def test_cli_tools_match_fixtures():
    # the reference's default flags: dimension reorder pre-pass on, -s 3
    # (STRIPES = (3^3 + 1) / 2 = 14), KBLOCK 4 (main.cpp:43-55, :83-87)
    g, fx = _fixture("self_uniform_d8")
    base = ("-n", fx["nA"], "-e", fx["eps"], "-d", fx["d"], "-f", g / fx["a"])
    fields = _run_tool("hilbertSelfJoinCardinality", *base)
    assert len(fields) == 15 and int(fields[12]) == fx["count"]
    assert int(fields[5]) == 14 and int(fields[6]) == 4
    assert float(fields[11]) > 0  # REORDERTIME: the pre-pass ran
    assert 0 < float(fields[13]) <= 1  # LOADPERCENT (64 x 64 tiles)
    # no pair of this fixture lies within rounding of eps, so neither the
    # summation order of the reorder nor the knobs can change the count
    for extra in (("--no-reorder",), ("--kernel=f16",), ("--kblock=16",), ("--kblock=8", "-s", "1"),
                  ("--auto-dims", "--kernel=bf16")):
        fields = _run_tool("hilbertSelfJoinCardinality", *base, *extra)
        assert int(fields[12]) == fx["count"], extra
        if extra == ("--no-reorder",):
            assert float(fields[11]) == 0.0
        if "-s" in extra:
            assert int(fields[5]) == 2  # (3^1 + 1) / 2
        if "--kblock=16" in extra:
            assert int(fields[6]) == 16
    g, fx = _fixture("join_d20")
    for kernel in ("fp64", "f16"):
        fields = _run_tool("hilbertJoinCardinality", "-n", fx["nA"], "-m", fx["nB"], "-e",
                           fx["eps"], "-d", fx["d"], "-f", g / fx["a"], "-g", g / fx["b"],
                           f"--kernel={kernel}")
        assert len(fields) == 16 and int(fields[13]) == fx["count"]
        assert int(fields[6]) == 14 and float(fields[12]) > 0


def test_reference_signature_shims():
    import ctypes
    lib = hj.load_library()
    X = _uniform(301, 3003, 8)
    Y = _uniform(302, 2000, 8)
    eps = 0.3
    Xc = X.copy()
    cnt = ctypes.c_size_t()
    st, it, lp = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    f = getattr(lib, "_Z20test_ego_loop3_macrommdPdPmiS_S_S_")
    f.restype = None
    f.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]
    f(X.shape[0], X.shape[1], eps, Xc.ctypes.data, ctypes.byref(cnt), 14, ctypes.byref(st),
      ctypes.byref(it), ctypes.byref(lp))
    assert cnt.value == oracle.self_count(X, eps)
    assert np.array_equal(Xc, X)  # inputs are not modified
    assert 0 < lp.value <= 1 and st.value >= 0 and it.value >= 0
    g = getattr(lib, "_Z21test_ego_loop3_noselfmmidPdS_PmiS_S_S_")
    g.restype = None
    g.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_double,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    g(X.shape[0], Y.shape[0], X.shape[1], eps, Xc.ctypes.data, Y.ctypes.data, ctypes.byref(cnt),
      2, ctypes.byref(st), ctypes.byref(it), ctypes.byref(lp))
    assert cnt.value == oracle.join_count(X, Y, eps)


def test_active_dims_autotuned(ctx):
    X = _uniform(311, 20000, 16)
    eps = 0.3
    got, st = ctx.self_join_count(X, eps)
    assert 0 <= st.active_dims <= 6
    assert got == oracle.self_count(X, eps)


@pytest.mark.parametrize("kernel", ["f16", "bf16"])
def test_band_queue_overflow(ctx, kernel):
    # integers on a line at eps = 1: every neighbour pair lies exactly at eps
    # (count n - 1 in closed form), and at this coordinate range the filter
    # cannot decide any pair, so its band queue overflows the initial capacity
    # (2^20 pairs) and the join runs again with the exact size
    n = 300_000
    X = np.arange(n, dtype=np.float64)[:, None]
    got, st = ctx.self_join_count(X, 1.0, kernel=kernel)
    assert got == n - 1
    assert st.band_pairs > (1 << 20)


@pytest.mark.parametrize("offset,eps,d", [(100.0, 0.5, 4), (1e6, 0.05, 4), (3.0, 0.3, 24)])
def test_fp64_expanded_band_and_difference_fallback(ctx, offset, eps, d):
    # Pairs planted at eps (1 +- 2^-k) around a common offset.  Near the
    # origin the FP64 kernel's expanded form decides most pairs and sends the
    # ones within its rounding band to the fixed-order predicate; far from it
    # (offset 1e6 against eps 0.05) the band would not be thin and the kernel
    # takes the difference form.  Either way the count is the oracle's.
    rng = np.random.default_rng(int(offset) % 1000 + d)
    base = rng.random((600, d)) + offset
    dirs = rng.normal(size=(600, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ks = rng.integers(26, 53, size=600)
    r = eps * (1 + rng.choice([-1.0, 1.0], size=600) * np.ldexp(1.0, -ks))
    X = np.concatenate([base, base + dirs * r[:, None], rng.random((800, d)) + offset])
    want = oracle.self_count(X, eps)
    assert want > 200
    for kb in (4, 8):
        got, _ = ctx.self_join_count(X, eps, kernel="fp64", kblock=kb)
        assert got == want, kb
    Y = X[::2].copy()
    got, _ = ctx.join_count(X, Y, eps, kernel="fp64")
    assert got == oracle.join_count(X, Y, eps)
