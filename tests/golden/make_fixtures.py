"""Writes the golden fixtures of tests/golden (run from the repo root).

The reference ships no tests, fixtures or golden vectors (its only
known-answer datum is the FAQ line COUNTS=0 for n=200000, d=64, eps=0.2,
FAQ.md:32-33), and it cannot be built here (Boost and MKL headers are
absent), so the expected counts come from the FP64 oracle and are confirmed
by an independent numpy brute force; fixtures whose pairs sit within
rounding of eps record the exact-rational count as well.

Every fixture is raw row-major float64 (`.bin`, the reference's input
format, dataIo.cpp:128-137) plus an entry in manifest.json.
"""
from __future__ import annotations

import json
import sys
from fractions import Fraction
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402


def numpy_count(A, B, eps, self):
    diff = A[:, None, :] - B[None, :, :]
    d2 = np.einsum("ijk,ijk->ij", diff, diff)
    if self:
        d2 = d2[np.triu_indices(A.shape[0], 1)]
    return int(np.count_nonzero(d2 <= eps * eps)), d2


def exact_count(X, eps):
    e2 = Fraction(eps) * Fraction(eps)
    c = 0
    for i in range(X.shape[0]):
        for j in range(i + 1, X.shape[0]):
            s = sum((Fraction(a) - Fraction(b)) ** 2 for a, b in zip(X[i], X[j]))
            c += s <= e2
    return c


def band_set(seed, npairs, d, eps):
    rng = np.random.default_rng(seed)
    base = rng.random((npairs, d)) * 0.8
    base[:, : d // 2] = np.round(base[:, : d // 2] / eps) * eps
    dirs = rng.normal(size=(npairs, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ks = rng.integers(20, 50, size=npairs)
    sign = rng.choice([-1.0, 1.0], size=npairs)
    r = eps * (1 + sign * np.ldexp(1.0, -ks))
    return np.concatenate([base, base + dirs * r[:, None]])


def main():
    rng = np.random.default_rng
    fixtures = []

    def add(name, kind, eps, A, B=None, exact=False):
        d = A.shape[1]
        fa = f"{name}_a.bin"
        A.astype(np.float64).tofile(HERE / fa)
        entry = {"name": name, "kind": kind, "d": d, "eps": eps, "a": fa, "nA": A.shape[0]}
        if kind == "self":
            count = oracle.self_count(A, eps)
            npc, d2 = numpy_count(A, A, eps, True)
        else:
            fb = f"{name}_b.bin"
            B.astype(np.float64).tofile(HERE / fb)
            entry.update(b=fb, nB=B.shape[0])
            count = oracle.join_count(A, B, eps)
            npc, d2 = numpy_count(A, B, eps, False)
        near = int(np.count_nonzero(np.abs(d2 - eps * eps) <= 1e-12))
        if near == 0:
            assert npc == count, (name, npc, count)
        entry["count"] = count
        entry["pairs_within_1e-12_of_eps2"] = near
        if exact:
            entry["exact_rational_count"] = exact_count(A, eps)
        fixtures.append(entry)
        print(name, entry["count"], near)

    add("self_uniform_d8", "self", 0.3, rng(101).random((2000, 8)))
    add("self_dense_d64", "self", 2.6, rng(102).random((1000, 64)))
    add("self_band_d6", "self", 0.125, band_set(103, 150, 6, 0.125), exact=True)
    X = np.repeat(rng(104).random((3, 5)), [500, 300, 203], axis=0)
    add("self_dup_ragged_1003", "self", 0.05, X)
    add("self_lowd_d1", "self", 0.001, rng(105).random((1003, 1)))
    add("join_d20", "join", 0.9, rng(106).random((800, 20)), rng(107).random((700, 20)))
    add("join_d3_ragged", "join", 0.05, rng(108).random((1001, 3)), rng(109).random((999, 3)))
    (HERE / "manifest.json").write_text(json.dumps({"fixtures": fixtures}, indent=1) + "\n")


if __name__ == "__main__":
    main()
