"""Pins the CPU oracle (no GPU): closed forms, invariants, an independent
numpy brute force, the CPU EGO port against brute force, the reference's
expanded-form arithmetic (disagreements only within rounding of eps), the
Hilbert curve and the committed golden fixtures.

Each check is chosen so that a plausible mistake in the oracle (a dropped
dimension, a wrong comparison, a transposed operand, an off-by-one in the
pair loop) fails at least one of them.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"


def _uniform(seed, n, d):
    return np.random.default_rng(seed).random((n, d))


def _numpy_self_count(X, eps):
    # independent formulation: Gram-free pairwise distances in numpy
    diff = X[:, None, :] - X[None, :, :]
    d2 = np.einsum("ijk,ijk->ij", diff, diff)
    iu = np.triu_indices(X.shape[0], 1)
    return int(np.count_nonzero(d2[iu] <= eps * eps)), d2[iu]


# ---- cell rule and comparator (egojoin.cpp:9-17) --------------------------------
def test_cell_rule_is_reciprocal_multiply():
    for eps in (0.2, 0.03, 0.1, 1.0 / 3.0, 2.6):
        r = 1.0 / eps
        for x in np.linspace(-3, 3, 1001):
            assert oracle.cell(x, eps) == np.floor(x * r)


def test_cell_boundaries():
    eps = 0.25  # exact in binary: x = k*eps lands on a cell boundary exactly
    for k in range(-8, 9):
        x = k * eps
        assert oracle.cell(x, eps) == k
        assert oracle.cell(np.nextafter(x, -np.inf), eps) == k - 1
        assert oracle.cell(np.nextafter(x, np.inf), eps) == k


def test_ego_compare_lexicographic():
    eps = 0.5
    a = np.array([0.1, 0.9, 0.2])
    assert oracle.ego_compare(a, a, eps) == 0
    assert oracle.ego_compare(a, np.array([0.6, 0.0, 0.0]), eps) < 0
    assert oracle.ego_compare(a, np.array([0.1, 0.6, 0.4]), eps) == 0  # same cells
    assert oracle.ego_compare(a, np.array([0.1, 0.4, 0.9]), eps) > 0
    assert oracle.ego_compare(a, np.array([0.4, 0.9, 0.7]), eps) < 0


# ---- the fixed-order predicate ---------------------------------------------------
def test_dist2_fixed_order_fma():
    from fractions import Fraction
    rng = np.random.default_rng(1)
    for d in (1, 2, 7, 64):
        a, b = rng.random(d), rng.random(d)
        s = 0.0
        for k in range(d):
            t = a[k] - b[k]  # one IEEE subtraction
            s = float(Fraction(t) * Fraction(t) + Fraction(s))  # fma: a single rounding
        assert oracle.dist2(a, b) == s


def test_identical_points():
    for n in (1, 2, 7, 8, 9, 1003):
        X = np.tile(np.array([[0.3, 0.7]]), (n, 1))
        assert oracle.self_count(X, 0.01) == n * (n - 1) // 2
        assert oracle.join_count(X, X, 0.01) == n * n


def test_lattice():
    k = 9
    g = np.arange(k, dtype=np.float64)
    X = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    assert oracle.self_count(X, 1.0) == 2 * k * (k - 1)          # axis neighbours
    assert oracle.self_count(X, np.sqrt(2.0) * (1 + 1e-12)) == 2 * k * (k - 1) + 2 * (k - 1) ** 2
    X3 = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    assert oracle.self_count(X3, 1.0) == 3 * k * k * (k - 1)
    # a dropped last dimension would count pairs that differ only in dim 2
    assert oracle.self_count(X3, 0.5) == 0


def test_separated_clusters():
    a = _uniform(2, 300, 4) * 0.1
    b = _uniform(3, 200, 4) * 0.1 + 10.0
    eps = 0.05
    assert oracle.self_count(np.vstack([a, b]), eps) == oracle.self_count(a, eps) + \
        oracle.self_count(b, eps)
    assert oracle.join_count(a, b, eps) == 0


@pytest.mark.parametrize("n,d,eps", [(400, 3, 0.1), (300, 16, 0.9), (250, 64, 2.6),
                                     (500, 1, 0.003)])
def test_matches_numpy_brute_force(n, d, eps):
    X = _uniform(n + d, n, d)
    want, d2 = _numpy_self_count(X, eps)
    # the formulations may differ only for pairs within rounding of eps
    assert np.count_nonzero(np.abs(d2 - eps * eps) < 1e-12) == 0
    assert oracle.self_count(X, eps) == want
    assert oracle.self_count(X, eps, early_exit=False) == want


def test_invariants():
    X = _uniform(5, 600, 6)
    Y = _uniform(6, 500, 6)
    eps = 0.3
    s = oracle.self_count(X, eps)
    rng = np.random.default_rng(0)
    assert oracle.self_count(X[rng.permutation(600)], eps) == s
    assert oracle.self_count(X[:, rng.permutation(6)], eps) == s
    assert oracle.join_count(X, X, eps) == 2 * s + 600
    assert oracle.join_count(X, Y, eps) == oracle.join_count(Y, X, eps)
    counts = [oracle.self_count(X, e) for e in (0.1, 0.2, 0.3, 0.4)]
    assert counts == sorted(counts) and counts[0] < counts[-1]


# ---- CPU EGO port (restatement of the reference algorithm) -----------------------
@pytest.mark.parametrize("n,d,eps", [(3000, 8, 0.3), (2001, 3, 0.05), (1500, 64, 2.6),
                                     (1003, 1, 0.001), (777, 20, 0.9)])
@pytest.mark.parametrize("active", [0, 1, 3])
def test_ego_port_equals_brute_force(n, d, eps, active):
    X = _uniform(n * 7 + d, n, d)
    want = oracle.self_count(X, eps)
    got, _, evaluated = oracle.ego_self_count(X, eps, active_dims=active)
    assert got == want
    assert evaluated <= 64 * ((n + 7) // 8) ** 2


def test_ego_port_join_equals_brute_force():
    A, B = _uniform(11, 1500, 5), _uniform(12, 1200, 5)
    for active in (0, 2, 4):
        assert oracle.ego_join_count(A, B, 0.2, active_dims=active)[0] == \
            oracle.join_count(A, B, 0.2)


def test_ego_port_boundary_pairs():
    # partner pairs straddling cell boundaries in every dimension at a tiny
    # distance: the reference's exclusive upper bound misses some of these
    eps = 0.125
    rng = np.random.default_rng(9)
    base = np.round(rng.random((500, 8)) * 7) * eps
    X = np.vstack([base - 1e-4, base + 1e-4])
    want = oracle.self_count(X, eps)
    assert oracle.ego_self_count(X, eps, active_dims=3)[0] == want


def test_ego_sort_is_lexicographic_and_stable():
    X = _uniform(13, 2000, 4)
    eps = 0.3
    perm = oracle.ego_sort(X, eps)
    cells = np.floor(X * (1.0 / eps)).astype(np.int64)
    want = np.lexsort(tuple(cells[:, k] for k in range(3, -1, -1)))
    assert np.array_equal(perm, want)


# ---- the reference's expanded-form arithmetic ------------------------------------
def test_expanded_form_disagrees_only_within_rounding():
    X = _uniform(17, 1500, 16)
    eps = 0.6
    bad = oracle.self_disagreements(X, eps, kblock=4)
    for i, j in bad:
        assert abs(oracle.dist2(X[i], X[j]) - eps * eps) <= 1e-12
    assert oracle.self_count_expanded(X, eps) - oracle.self_count(X, eps) in \
        range(-len(bad), len(bad) + 1)


def test_expanded_form_counts_far_and_near_pairs():
    # exact arithmetic of the expanded form on clear cases
    a = np.zeros(8)
    assert oracle.lib().hjo_pair_within_expanded(a.ctypes.data, (a + 0.01).ctypes.data, 8, 0.1,
                                                 4) == 1
    b = a + 0.1
    assert oracle.lib().hjo_pair_within_expanded(a.ctypes.data, b.ctypes.data, 8, 0.1, 4) == 0


# ---- Hilbert curve -------------------------------------------------------------
@pytest.mark.parametrize("order", [1, 2, 3, 5])
def test_hilbert_bijection_and_adjacency(order):
    side = 1 << order
    seen = {}
    for x in range(side):
        for y in range(side):
            h = oracle.hilbert_xy2d(order, x, y)
            assert 0 <= h < side * side
            seen[h] = (x, y)
            assert oracle.hilbert_d2xy(order, h) == (x, y)
    assert len(seen) == side * side
    for h in range(side * side - 1):
        (x0, y0), (x1, y1) = seen[h], seen[h + 1]
        assert abs(x0 - x1) + abs(y0 - y1) == 1
    assert seen[0] == (0, 0)
    assert seen[side * side - 1] == (side - 1, 0)


def test_hilbert_known_values():
    # order 1: (0,0)->0, (0,1)->1, (1,1)->2, (1,0)->3
    assert [oracle.hilbert_xy2d(1, x, y) for x, y in [(0, 0), (0, 1), (1, 1), (1, 0)]] == \
        [0, 1, 2, 3]


# ---- CSR ------------------------------------------------------------------------
def test_csr_consistent_with_count():
    X = _uniform(19, 800, 5)
    eps = 0.25
    rp, cols = oracle.self_csr(X, eps)
    assert rp[-1] == 2 * oracle.self_count(X, eps)
    for i in range(800):
        row = cols[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0) and i not in row
        for j in row:
            assert i in cols[rp[j]:rp[j + 1]]


# ---- golden fixtures ---------------------------------------------------------------
def _fixtures():
    man = GOLDEN / "manifest.json"
    return json.loads(man.read_text())["fixtures"] if man.exists() else []


@pytest.mark.parametrize("fx", _fixtures(), ids=lambda f: f["name"])
def test_golden_fixture(fx):
    X = np.fromfile(GOLDEN / fx["a"], dtype=np.float64).reshape(-1, fx["d"])
    if fx["kind"] == "self":
        assert oracle.self_count(X, fx["eps"]) == fx["count"]
    else:
        Y = np.fromfile(GOLDEN / fx["b"], dtype=np.float64).reshape(-1, fx["d"])
        assert oracle.join_count(X, Y, fx["eps"]) == fx["count"]


def test_dim_costs_brute_force_and_closed_forms():
    # cost[k] = pairs whose cells in dimension k differ by at most one: checked
    # against a direct count over all pairs, and against closed forms
    rng = np.random.default_rng(3)
    X = rng.random((300, 5)) * np.array([1.0, 3.0, 0.2, 1.0, 10.0])
    eps = 0.17
    got = oracle.dim_costs(X, eps)
    cells = np.floor(X * (1.0 / eps))
    for k in range(5):
        diff = np.abs(cells[:, None, k] - cells[None, :, k])
        iu = np.triu_indices(len(X), 1)
        assert got[k] == int((diff[iu] <= 1).sum())
    # n identical points: n(n-1)/2 in every dimension; cells 3 apart: none
    assert list(oracle.dim_costs(np.zeros((7, 3)), 0.5)) == [21, 21, 21]
    Y = np.array([[0.0], [0.3], [0.6], [0.9]])
    assert oracle.dim_costs(Y, 0.1)[0] == 0
    assert oracle.dim_costs(Y, 0.3)[0] == 3  # cells 0,1,2,3: three neighbour pairs


# ---- A x B pair lists (hjo_join_csr) ----------------------------------------------
def _numpy_join_pairs(A, B, eps):
    # independent formulation: broadcast differences, einsum sum of squares
    diff = A[:, None, :] - B[None, :, :]
    d2 = np.einsum("ijk,ijk->ij", diff, diff)
    return d2 <= eps * eps, d2


def test_join_csr_matches_numpy_and_count():
    A, B = _uniform(31, 700, 6), _uniform(32, 500, 6)
    eps = 0.3
    rp, cols = oracle.join_csr(A, B, eps)
    mask, d2 = _numpy_join_pairs(A, B, eps)
    assert np.count_nonzero(np.abs(d2 - eps * eps) < 1e-12) == 0
    assert rp[0] == 0 and rp[-1] == oracle.join_count(A, B, eps) == np.count_nonzero(mask)
    for i in range(len(A)):
        row = cols[rp[i]:rp[i + 1]]
        assert np.array_equal(row, np.nonzero(mask[i])[0])  # ascending, exact set


def test_join_csr_is_transpose_of_swapped_join():
    A, B = _uniform(33, 400, 5), _uniform(34, 350, 5)
    eps = 0.35
    rab, cab = oracle.join_csr(A, B, eps)
    rba, cba = oracle.join_csr(B, A, eps)
    ab = {(i, int(j)) for i in range(len(A)) for j in cab[rab[i]:rab[i + 1]]}
    ba = {(int(j), i) for i in range(len(B)) for j in cba[rba[i]:rba[i + 1]]}
    assert ab == ba and len(ab) > 0


def test_join_csr_of_a_set_with_itself_is_self_csr_plus_diagonal():
    X = _uniform(35, 500, 4)
    eps = 0.2
    rs, cs = oracle.self_csr(X, eps)
    rj, cj = oracle.join_csr(X, X, eps)
    for i in range(len(X)):
        want = np.sort(np.append(cs[rs[i]:rs[i + 1]], i))
        assert np.array_equal(cj[rj[i]:rj[i + 1]], want)


# ---- the CPU baseline's sampled join loop (hjo_ego_self_count_sampled) ---------------
@pytest.mark.parametrize("stride", [1, 2, 3, 7])
def test_ego_sampled_equals_brute_force_on_sampled_blocks(stride):
    # the join loop runs on every stride-th 8-point block of the EGO-sorted
    # set; a pair is counted under the block of its smaller sorted position
    X = _uniform(41, 1600, 6)
    eps = 0.3
    got, _, ev = oracle.ego_self_count_sampled(X, eps, stride, active_dims=2)
    perm = oracle.ego_sort(X, eps)
    pos = np.empty(len(X), dtype=np.int64)
    pos[perm] = np.arange(len(X))
    _, d2 = _numpy_self_count(X, eps)
    iu, ju = np.triu_indices(len(X), 1)
    close = d2 <= eps * eps
    first = np.minimum(pos[iu[close]], pos[ju[close]])
    assert got == int(np.count_nonzero((first // 8) % stride == 0))
    full, _, ev_full = oracle.ego_self_count(X, eps, active_dims=2)
    if stride == 1:
        assert got == full and ev == ev_full
    else:
        assert 0 < ev < ev_full


# ---- pairs near eps (hjo_self_band) --------------------------------------------------
def test_self_band_matches_numpy():
    X = _uniform(43, 600, 5)
    eps, tol = 0.4, 0.02
    got = {tuple(p) for p in oracle.self_band(X, eps, tol)}
    _, d2 = _numpy_self_count(X, eps)
    iu, ju = np.triu_indices(len(X), 1)
    gap = np.abs(d2 - eps * eps)
    assert np.count_nonzero(np.abs(gap - tol) < 1e-12) == 0
    want = {(int(i), int(j)) for i, j in zip(iu[gap <= tol], ju[gap <= tol])}
    assert got == want and len(want) > 0


# ---- the reference's A x B tile arithmetic (hjo_join_disagreements) ----------------
def _band_pairs(seed, n, d, eps, ks):
    # pairs at eps (1 + s 2^-k): the expanded form's cancellation can flip them
    rng = np.random.default_rng(seed)
    a = rng.random((n, d)) * 0.5 + 0.25
    w = rng.normal(size=(n, d))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    k = rng.choice(ks, size=n)
    sgn = rng.choice([-1.0, 1.0], size=n)
    b = a + (eps * (1.0 + sgn * 2.0 ** -k))[:, None] * w
    return a, b


def test_join_disagreements_only_within_rounding_and_symmetric():
    eps = 0.3
    A, B = _band_pairs(51, 1500, 16, eps, np.arange(40, 54))
    bad = oracle.join_disagreements(A, B, eps)
    assert len(bad) > 0  # the band pairs expose the expanded form's rounding
    for i, j in bad:
        assert abs(oracle.dist2(A[i], B[j]) - eps * eps) <= 1e-13
    # the NOSELF arithmetic is symmetric in A and B (the products and the
    # block-norm sums commute), so the lists are transposes
    bad_t = oracle.join_disagreements(B, A, eps)
    assert {tuple(p) for p in bad} == {(int(j), int(i)) for i, j in bad_t}
    # the end-evaluated count differs from the fixed-order one by at most
    # the listed pairs
    diff = oracle.join_count_expanded(A, B, eps) - oracle.join_count(A, B, eps)
    assert abs(diff) <= len(bad)


def test_join_disagreements_empty_far_from_eps_and_consistent_with_self():
    eps = 0.3
    A, B = _band_pairs(52, 400, 16, eps, np.arange(4, 8))  # 2^-7 away from eps at least
    assert len(oracle.join_disagreements(A, B, eps)) == 0
    X = np.vstack(_band_pairs(53, 300, 8, eps, np.arange(44, 52)))
    sd = {tuple(p) for p in oracle.self_disagreements(X, eps)}
    jd = {(int(i), int(j)) for i, j in oracle.join_disagreements(X, X, eps) if i < j}
    assert sd == jd


def test_expanded_form_pair_rule_on_clear_cases():
    # exact cases of the pair rule (sign bit of the expanded sum): far in,
    # far out, and the join count of the expanded form on them
    a = np.zeros((1, 8))
    near, far = a + 0.01, a + 0.1
    assert oracle.pair_within_expanded(a[0], near[0], 0.1)
    assert not oracle.pair_within_expanded(a[0], far[0], 0.1)
    assert oracle.pair_within(a[0], near[0], 0.1) and not oracle.pair_within(a[0], far[0], 0.1)
    assert oracle.join_count_expanded(a, np.vstack([near, far]), 0.1) == 1


# ---- every oracle function is pinned by a CPU test here ------------------------------
def test_every_oracle_function_has_a_cpu_check():
    import re
    text = Path(__file__).read_text()
    called = set(re.findall(r"oracle\.(\w+)\(", text)) | set(re.findall(r"\b(hjo_\w+)\b", text))
    wrappers = {n[len("hjo_"):] for n in oracle._SIGS}
    missing = [n for n in oracle._SIGS if n not in called and n[len("hjo_"):] not in called]
    assert not missing, missing
    assert wrappers  # the signature table is not empty


def test_thread_count_does_not_change_results():
    X = _uniform(45, 1200, 5)
    threads = oracle.max_threads()
    assert threads >= 1
    full = oracle.ego_self_count(X, 0.25, active_dims=2)[0]
    assert oracle.ego_self_count(X, 0.25, active_dims=2, threads=1)[0] == full
    assert oracle.max_threads() == 1  # the setting persists, as omp_set_num_threads does
    assert full == oracle.self_count(X, 0.25)
    # restore the thread count for the other tests
    oracle.ego_self_count(X[:8], 0.25, threads=threads)
    assert oracle.max_threads() == threads
