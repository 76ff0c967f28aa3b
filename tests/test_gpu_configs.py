"""The measured configurations at full size, on the HIP path, with planted pairs.

Every configuration of the benchmark is joined at its real size, with a few
hundred pairs planted at eps * (1 +- 2^-k) (tests/planted.py) so that counts
are not 0 and every filter kernel sends pairs through its FP64 band recheck.
Counts must be identical on the FP64 kernel and both MFMA filters.  They are
pinned, depending on what finishes on the host in seconds, by the oracle's
CPU EGO join on the same points (the exact count), by brute force on small
jobs and requests, and otherwise by the pair lists: every listed pair is
checked with the fixed-order FP64 predicate and every planted pair within
eps must be listed.  Identities: a self-join's A x A join equals twice its
count plus n, and an A x B count does not depend on the order of the sets.

Workloads (seeds as the benchmark's):
  self n=200,000 d=64 uniform eps=0.2 (seed 1), also through the command-line
      tool with the reference's default flags (unplanted: 0 pairs, as the
      reference's own known answer, FAQ.md:32-33)
  A x B 2M x 2M d=20 uniform eps=0.2 (seeds 1, 2)
  self n=11M d=28, 64-component Gaussian mixture, sigma 0.04, eps=0.03 (seed 3)
  self n=1M d=64 uniform, eps 0.15 / 0.2 / 0.3 (seed 4)
  256 self-joins of 1e3..2e5 uniform d=64 points, eps 0.2, one launch (seed 5)
  serving: resident 4M x 32 mixture (128 components, sigma 0.05), a burst of
      requests of 1..65536 query points, eps U[0.02, 0.10], half pair lists
"""
import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle
import planted

pytestmark = pytest.mark.gpu

KERNELS = ["fp64", "f16", "bf16"]


def _run_tool(tool, *args):
    import subprocess
    from pathlib import Path
    root = Path(__file__).resolve().parent.parent
    out = subprocess.run([str(root / "hilbertjoin_amd/bin" / tool), *map(str, args)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout.strip().split(";")


def _self_counts(ctx, dev, eps, kernels=KERNELS):
    counts = {}
    for kernel in kernels:
        print(f"  self count kernel={kernel}", flush=True)
        counts[kernel], st = ctx.self_join_count(dev, eps, kernel=kernel)
        if kernel != "fp64":
            assert st.band_pairs > 0, kernel  # the planted pairs went through the recheck
    assert len(set(counts.values())) == 1, counts
    return counts["fp64"]


def _check_self_csr(ctx, dev, X, eps, pairs, inside, kernel, count):
    rp, cols, _ = ctx.self_join_csr(dev, eps, kernel=kernel)
    rp, cols = rp.numpy().ravel(), cols.numpy().ravel()
    listed = planted.csr_pairs(rp, cols)
    assert len(listed) == 2 * count
    und = {(i, j) for i, j in listed if i < j}
    assert len(und) == count and all((j, i) in listed for i, j in und)
    planted.verify_listed(und, X, X, eps)
    want = {tuple(sorted(p)) for p, ok in zip(pairs.tolist(), inside) if ok}
    assert want <= und
    return und


def test_cfg1_self_200k_d64(ctx, tmp_path):
    eps = 0.2
    X = hj.gen_uniform(1, 200_000, 64)
    # the reference's known answer (FAQ.md:32-33): 0 pairs, through the
    # command-line tool with its default flags (reorder pre-pass, -s 3)
    X.tofile(tmp_path / "cfg1.bin")
    fields = _run_tool("hilbertSelfJoinCardinality", "-n", 200000, "-e", eps, "-d", 64, "-f",
                       tmp_path / "cfg1.bin")
    assert int(fields[12]) == 0 and int(fields[5]) == 14 and float(fields[11]) > 0
    assert 0 < float(fields[13]) < 0.2
    Xp, pairs, inside = planted.plant_self(X, eps, 300, seed=101)
    assert inside.sum() > 100
    want = oracle.ego_self_count(Xp, eps, active_dims=3)[0]  # the exact count
    dev = ctx.to_device(Xp)
    got = _self_counts(ctx, dev, eps)
    assert got == want >= inside.sum()
    _check_self_csr(ctx, dev, Xp, eps, pairs, inside, "f16", got)
    j, _ = ctx.join_count(dev, dev, eps, kernel="f16")
    assert j == 2 * got + Xp.shape[0]
    dev.free()


def test_cfg2_join_2M_x_2M_d20(ctx):
    eps = 0.2
    A = hj.gen_uniform(1, 2_000_000, 20)
    B = hj.gen_uniform(2, 2_000_000, 20)
    Ap, Bp, pairs, inside = planted.plant_join(A, B, eps, 300, seed=202)
    del A, B
    assert inside.sum() > 100
    dA, dB = ctx.to_device(Ap), ctx.to_device(Bp)
    counts = {}
    for kernel in KERNELS:
        print(f"  join count kernel={kernel}", flush=True)
        counts[kernel], st = ctx.join_count(dA, dB, eps, kernel=kernel)
        if kernel != "fp64":
            assert st.band_pairs > 0
        c2, _ = ctx.join_count(dB, dA, eps, kernel=kernel)
        assert c2 == counts[kernel]
    assert len(set(counts.values())) == 1, counts
    rp, cols, _ = ctx.join_csr(dA, dB, eps, kernel="f16")
    listed = planted.csr_pairs(rp.numpy().ravel(), cols.numpy().ravel())
    assert len(listed) == counts["fp64"]
    planted.verify_listed(listed, Ap, Bp, eps)
    assert {tuple(p) for p, ok in zip(pairs.tolist(), inside) if ok} <= listed
    dA.free()
    dB.free()


def test_cfg3_self_11M_d28_mixture(ctx):
    eps = 0.03
    X = hj.gen_mixture(3, 11_000_000, 28, 64, 0.04)
    Xp, pairs, inside = planted.plant_self(X, eps, 300, seed=303)
    del X
    assert inside.sum() > 100
    dev = ctx.to_device(Xp)
    got = _self_counts(ctx, dev, eps)
    _check_self_csr(ctx, dev, Xp, eps, pairs, inside, "bf16", got)
    j, _ = ctx.join_count(dev, dev, eps, kernel="bf16")
    assert j == 2 * got + Xp.shape[0]
    dev.free()


@pytest.mark.parametrize("eps", [0.15, 0.2, 0.3])
def test_cfg4_self_1M_d64(ctx, eps):
    X = hj.gen_uniform(4, 1_000_000, 64)
    Xp, pairs, inside = planted.plant_self(X, eps, 300, seed=400 + int(eps * 100))
    del X
    assert inside.sum() > 100
    dev = ctx.to_device(Xp)
    got = _self_counts(ctx, dev, eps)
    if eps < 0.25:  # the CPU EGO join finishes in seconds here
        print("  oracle CPU EGO join", flush=True)
        assert got == oracle.ego_self_count(Xp, eps, active_dims=3)[0]
    _check_self_csr(ctx, dev, Xp, eps, pairs, inside, "f16", got)
    j, _ = ctx.join_count(dev, dev, eps, kernel="f16")
    assert j == 2 * got + Xp.shape[0]
    dev.free()


def test_cfg4_batch_256_self_joins(ctx):
    eps = 0.2
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(2e5), size=256)).astype(np.int64)
    jobs, planted_in = [], []
    for k, n in enumerate(sizes):
        Xk = hj.gen_uniform(1000 + k, int(n), 64)
        Xk, _, inside = planted.plant_self(Xk, eps, 4, seed=5000 + k)
        jobs.append(Xk)
        planted_in.append(int(inside.sum()))
    off = np.zeros(257, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    P = np.concatenate(jobs)
    dev = ctx.to_device(P)
    del P
    counts = {}
    for name, kw in (("fp64", {"kernel": "fp64"}), ("f16", {"kernel": "f16", "tune": True}),
                     ("bf16", {"kernel": "bf16"})):
        print(f"  batch {name}", flush=True)
        counts[name], st = ctx.batch_self_join_count(dev, off, eps, **kw)
        if name != "fp64":
            assert st.band_pairs > 0
    for name in ("f16", "bf16"):
        assert np.array_equal(counts[name], counts["fp64"]), name
    # brute force on the small jobs; the planted closed form on all (a pair of
    # unplanted uniform points at d = 64 lies within 0.2 with probability
    # about 1e-53)
    for k, n in enumerate(sizes):
        if n <= 4000:
            assert counts["fp64"][k] == oracle.self_count(jobs[k], eps), k
    assert counts["fp64"].tolist() == planted_in
    dev.free()


def test_cfg5_serving_burst_with_planted_neighbours(ctx):
    d, n_index, eps_max = 32, 4_000_000, 0.10
    allpts = hj.gen_mixture(6, n_index + 1_000_000, d, 128, 0.05)
    X, pool = allpts[:n_index], allpts[n_index:]
    rng = np.random.default_rng(7)
    nreq = 96
    sizes = np.maximum(np.minimum(np.exp(rng.uniform(0.0, np.log(65536.0), size=nreq))
                                  .astype(np.int64), 65536), 1)
    eps = rng.uniform(0.02, eps_max, size=nreq)
    want_pairs = rng.random(nreq) < 0.5
    reqs, plants = [], []
    for i in range(nreq):
        s0 = int(rng.integers(0, pool.shape[0] - 65536))
        Q = pool[s0:s0 + int(sizes[i])].copy()
        m = min(int(sizes[i]), 4)
        near, rows = planted.near_copies(X, float(eps[i]), m, rng)
        Q[:m] = near
        inside = [(q, int(rows[q])) for q in range(m)
                  if oracle.pair_within(Q[q], X[rows[q]], float(eps[i]))]
        reqs.append(Q)
        plants.append(inside)
    index = hj.Index(ctx, X, eps_max, kernel="f16")
    results = {}
    for kernel in ("f16", "fp64"):
        print(f"  server kernel={kernel}", flush=True)
        with hj.Server(index, workers=2, kernel=kernel, result_bytes=1 << 30) as srv:
            tickets = [srv.submit(reqs[i], float(eps[i]), pairs=bool(want_pairs[i]))
                       for i in range(nreq)]
            results[kernel] = [srv.wait(t) for t in tickets]
            st = srv.stats()
        assert all(r.status == 0 for r in results[kernel])
        # the burst arrived while batches were on the device: requests were
        # admitted together (continuous batching)
        assert st["mean_batch_requests"] > 1.0, st
    for i in range(nreq):
        rf, r64 = results["f16"][i], results["fp64"][i]
        assert rf.count == r64.count, i
        e = float(eps[i])
        if sizes[i] <= 32:
            assert r64.count == oracle.join_count(reqs[i], X, e), i
        if want_pairs[i]:
            for r in (rf, r64):
                listed = planted.csr_pairs(r.rowptr, r.colidx)
                assert len(listed) == r.count
                planted.verify_listed(listed, reqs[i], X, e)
                assert set(plants[i]) <= listed, i
        else:
            assert r64.count >= len(plants[i])
    assert sum(len(p) for p in plants) > 100
    index.close()
