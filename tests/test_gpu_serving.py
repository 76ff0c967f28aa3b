"""GPU parity of the resident index and the serving loop: every request's
count and pair list equal the FP64 oracle's A x B join of that request with
the indexed set, alone or batched with any other requests."""
import threading

import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle

pytestmark = pytest.mark.gpu

KERNELS = ["fp64", "f16", "bf16"]
EPS_MAX = 0.12


@pytest.fixture(scope="module")
def base():
    # indexed set: 20000 points of an 8-d Gaussian mixture; queries from the
    # same mixture (seed shifted) so they have partners
    X = hj.gen_mixture(31, 20000, 8, 24, 0.05)
    Qall = hj.gen_mixture(31, 26000, 8, 24, 0.05)[20000:] + 0.01
    return X, Qall


def _requests(Qall, sizes, seed):
    rng = np.random.default_rng(seed)
    off = np.zeros(len(sizes) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    Q = Qall[:int(off[-1])]
    eps = rng.uniform(0.02, EPS_MAX, size=len(sizes))
    return Q, off, eps


@pytest.fixture(scope="module")
def indexes(ctx, base):
    X, _ = base
    out = {k: hj.Index(ctx, X, EPS_MAX, kernel=k) for k in KERNELS}
    yield out
    for ix in out.values():
        ix.close()


@pytest.mark.parametrize("kernel", KERNELS)
def test_index_query_count_device_queries(ctx, base, indexes, kernel):
    # query points already in HBM (hj_index_query_count_device)
    X, Qall = base
    sizes = [3, 0, 200, 64, 1000]
    Q, off, eps = _requests(Qall, sizes, 5)
    dQ = ctx.to_device(Q)
    counts, _ = indexes[kernel].query_count(dQ, off, eps)
    want = [oracle.join_count(Q[off[k]:off[k + 1]], X, eps[k]) if sizes[k] else 0
            for k in range(len(sizes))]
    assert list(counts) == want and sum(want) > 0
    with pytest.raises(ValueError):  # offsets past the device rows: rejected on the host
        indexes[kernel].query_count(dQ, [0, len(Q) + 64], [0.05])
    dQ.free()


@pytest.mark.parametrize("kernel", KERNELS)
def test_index_query_count_matches_oracle(ctx, base, indexes, kernel):
    X, Qall = base
    sizes = [0, 1, 5, 64, 65, 300, 2000, 130]
    Q, off, eps = _requests(Qall, sizes, 1)
    print(f"index count kernel={kernel}", flush=True)
    counts, st = indexes[kernel].query_count(Q, off, eps)
    want = [oracle.join_count(Q[off[k]:off[k + 1]], X, eps[k]) if sizes[k] else 0
            for k in range(len(sizes))]
    assert list(counts) == want
    assert sum(want) > 0
    info = indexes[kernel].info()
    assert info["n"] == 20000 and info["d"] == 8 and info["eps_max"] == EPS_MAX


@pytest.mark.parametrize("kernel", KERNELS)
def test_index_query_csr_matches_oracle(ctx, base, indexes, kernel):
    X, Qall = base
    sizes = [3, 700, 0, 64, 1200]
    Q, off, eps = _requests(Qall, sizes, 2)
    rowptr, colidx, _ = indexes[kernel].query_csr(Q, off, eps)
    assert rowptr.shape == (int(off[-1]) + 1,)
    for k in range(len(sizes)):
        a, b = int(off[k]), int(off[k + 1])
        wr, wc = oracle.join_csr(Q[a:b], X, eps[k])
        got_r = rowptr[a:b + 1] - rowptr[a]
        assert np.array_equal(got_r, wr)
        assert np.array_equal(colidx[rowptr[a]:rowptr[b]], wc)
    assert colidx.size > 0


def test_index_queries_outside_the_index_range(ctx, base, indexes):
    # far queries (beyond the filter operands' range) take the FP64 kernel;
    # queries just outside the index's box are clamped into its grid
    X, Qall = base
    far = Qall[:50] + 100.0
    near = Qall[50:250] * 1.05 - 0.02
    Q = np.concatenate([far, near])
    off = np.array([0, 50, 250], dtype=np.uint64)
    eps = np.array([EPS_MAX, 0.08])
    wr, wc = oracle.join_csr(near, X, 0.08)
    for kernel in KERNELS:
        counts, _ = indexes[kernel].query_count(Q, off, eps)
        assert counts[0] == 0
        assert counts[1] == oracle.join_count(near, X, 0.08)
        # the pair lists of such a batch come from the kernel it ran with
        rowptr, colidx, _ = indexes[kernel].query_csr(Q, off, eps)
        assert np.all(rowptr[:51] == 0)
        assert np.array_equal(rowptr[50:] - rowptr[50], wr)
        assert np.array_equal(colidx, wc)


def test_index_rejects_eps_above_max(ctx, base, indexes):
    X, Qall = base
    with pytest.raises(hj.HJError):
        indexes["f16"].query_count(Qall[:10], [0, 10], [EPS_MAX * 1.01])


def test_index_fp64_override(ctx, base, indexes):
    X, Qall = base
    Q, off, eps = _requests(Qall, [400, 900], 3)
    a, _ = indexes["f16"].query_count(Q, off, eps)
    b, _ = indexes["f16"].query_count(Q, off, eps, kernel="fp64")
    assert np.array_equal(a, b)


@pytest.mark.parametrize("kernel", ["f16", "fp64"])
def test_server_results_same_alone_and_in_mix(ctx, base, indexes, kernel):
    # requests from three threads, count and pair-list requests mixed, are
    # batched continuously by two workers; each result equals the request
    # run alone against the index and the oracle's count
    X, Qall = base
    rng = np.random.default_rng(4)
    nreq = 60
    sizes = np.exp(rng.uniform(0, np.log(3000), size=nreq)).astype(int)
    starts = rng.integers(0, Qall.shape[0] - 3000, size=nreq)
    eps = rng.uniform(0.02, EPS_MAX, size=nreq)
    pairs = rng.random(nreq) < 0.5
    reqs = [(Qall[s:s + n], e, p) for s, n, e, p in zip(starts, sizes, eps, pairs)]
    srv = hj.Server(indexes["f16"], workers=2, kernel=kernel, max_batch_points=8192)
    tickets = [None] * nreq

    def submit(part):
        for i in part:
            tickets[i] = srv.submit(reqs[i][0], reqs[i][1], pairs=bool(reqs[i][2]))

    threads = [threading.Thread(target=submit, args=(range(t, nreq, 3),)) for t in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    results = [srv.wait(tk, timeout_ms=120000) for tk in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == nreq and st["failed"] == 0
    assert st["p99_ms"] >= st["p50_ms"] > 0
    assert any(r.batch_requests > 1 for r in results)
    for (q, e, p), r in zip(reqs, results):
        assert r.status == 0
        alone, _ = indexes["f16"].query_count(q, [0, len(q)], [e], kernel=kernel)
        assert r.count == alone[0] == oracle.join_count(q, X, e)
        if p:
            wr, wc = oracle.join_csr(q, X, e)
            assert np.array_equal(r.rowptr, wr)
            assert np.array_equal(r.colidx, wc)
        else:
            assert r.rowptr is None


def test_server_admission_by_result_memory(ctx, base, indexes):
    X, Qall = base
    q1, q2, q3 = Qall[:1000], Qall[1000:2000], Qall[2000:3000]
    e = 0.1
    c2 = oracle.join_count(q2, X, e)
    c3 = oracle.join_count(q3, X, e)
    assert c2 > 0 and c3 > 0
    # room for a count result and two pair lists' rowptrs, and for q2's and
    # q3's pairs only one at a time
    budget = 8 + (1001 * 8 + 8 * c2) + (1001 * 8 + 8 * c3) - 1
    srv = hj.Server(indexes["f16"], workers=2, result_bytes=budget)
    t1 = srv.submit(q1, e, pairs=False)
    t2 = srv.submit(q2, e, pairs=True)
    r1 = srv.wait(t1, release=False)
    r2 = srv.wait(t2, release=False)
    assert r1.status == 0 and r2.status == 0 and r2.count == c2
    t3 = srv.submit(q3, e, pairs=True)
    with pytest.raises(hj.HJError) as ex:
        srv.wait(t3, timeout_ms=1500, release=False)
    assert ex.value.code == 6  # still waiting for result memory
    assert srv.stats()["pending"] == 1
    srv.release(t2)
    r3 = srv.wait(t3, timeout_ms=60000)
    assert r3.status == 0 and r3.count == c3
    wr, wc = oracle.join_csr(q3, X, e)
    assert np.array_equal(r3.rowptr, wr) and np.array_equal(r3.colidx, wc)
    srv.release(t1)
    srv.close()
    # a pair-list request whose rowptr alone exceeds the budget fails
    small = hj.Server(indexes["f16"], workers=1, result_bytes=1000)
    t4 = small.submit(Qall[:200], e, pairs=True)
    t5 = small.submit(Qall[:200], e, pairs=False)
    assert small.wait(t4, timeout_ms=60000).status == 4
    r5 = small.wait(t5, timeout_ms=60000)
    assert r5.status == 0 and r5.count == oracle.join_count(Qall[:200], X, e)
    small.close()
