"""Planted neighbour pairs for the full-size configuration tests.

The configurations the join is measured on have no pairs within eps (at d = 64
the chance that two uniform points lie within 0.2 is about 1e-53), so a count
of 0 says little.  These helpers overwrite a few hundred rows with pairs at
distance eps * (1 + s * 2^-k), s = +-1, k drawn from [8, 50]: far inside the
reduced-precision filter's error band, so every one of them goes through the
FP64 recheck, and up to FP64 rounding of the distance itself.  Half of the
pairs are moved so that their midpoint sits on a cell boundary in every
dimension (the pair straddles the boundary everywhere, the case the
reference's exclusive stripe bound can miss).  Which planted pairs are within
eps is decided by the oracle's fixed-order FP64 predicate, pair by pair.
"""
from __future__ import annotations

import numpy as np

import oracle


def _offsets(rng, npairs, d, eps, boundary_frac, lo, hi):
    w = rng.normal(size=(npairs, d))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    k = rng.integers(8, 51, size=npairs)
    sgn = rng.choice([-1.0, 1.0], size=npairs)
    r = eps * (1.0 + sgn * np.ldexp(1.0, -k))
    return w * r[:, None], rng.random(npairs) < boundary_frac


def _place(a, delta, on_boundary, eps, lo, hi):
    b = a + delta
    if on_boundary.any():
        # midpoint on the cell boundary j * eps nearest to it, inside [lo, hi]
        mid = 0.5 * (a[on_boundary] + b[on_boundary])
        jmin, jmax = np.ceil(lo / eps) + 1, np.floor(hi / eps) - 1
        j = np.clip(np.round(mid / eps), jmin, max(jmin, jmax))
        half = 0.5 * delta[on_boundary]
        a[on_boundary] = j * eps - half
        b[on_boundary] = j * eps + half
    return a, b


def plant_self(X, eps, npairs, seed, boundary_frac=0.5, lo=0.0, hi=1.0):
    """Copy of X with 2 * npairs rows replaced by planted pairs.  Returns
    (X', pairs (npairs x 2 row indices), inside (bool per pair))."""
    rng = np.random.default_rng(seed)
    n, d = X.shape
    rows = rng.choice(n, size=2 * npairs, replace=False)
    ia, ib = rows[:npairs], rows[npairs:]
    delta, onb = _offsets(rng, npairs, d, eps, boundary_frac, lo, hi)
    a, b = _place(X[ia].copy(), delta, onb, eps, lo, hi)
    Xp = X.copy()
    Xp[ia] = a
    Xp[ib] = b
    inside = np.array([oracle.pair_within(a[p], b[p], eps) for p in range(npairs)])
    return Xp, np.stack([ia, ib], axis=1), inside


def plant_join(A, B, eps, npairs, seed, boundary_frac=0.5, lo=0.0, hi=1.0):
    """Copies of A and B with npairs rows of each replaced by planted pairs
    (a in A, b in B).  Returns (A', B', pairs (npairs x 2: row of A, row of
    B), inside)."""
    rng = np.random.default_rng(seed)
    d = A.shape[1]
    ia = rng.choice(A.shape[0], size=npairs, replace=False)
    ib = rng.choice(B.shape[0], size=npairs, replace=False)
    delta, onb = _offsets(rng, npairs, d, eps, boundary_frac, lo, hi)
    a, b = _place(A[ia].copy(), delta, onb, eps, lo, hi)
    Ap, Bp = A.copy(), B.copy()
    Ap[ia] = a
    Bp[ib] = b
    inside = np.array([oracle.pair_within(a[p], b[p], eps) for p in range(npairs)])
    return Ap, Bp, np.stack([ia, ib], axis=1), inside


def near_copies(P, eps, m, rng):
    """m points at distance eps * (1 + s 2^-k) from m distinct rows of P.
    Returns (points, rows of P)."""
    rows = rng.choice(P.shape[0], size=m, replace=False)
    delta, onb = _offsets(rng, m, P.shape[1], eps, 0.0, 0.0, 1.0)
    return P[rows] + delta, rows


def csr_pairs(rowptr, cols):
    """Set of (row, col) pairs of a CSR list."""
    rows = np.repeat(np.arange(rowptr.size - 1), np.diff(rowptr))
    return set(zip(rows.tolist(), np.asarray(cols).tolist()))


def verify_listed(pairs, P, Q, eps):
    """Every listed pair (i in P, j in Q) is within eps by the fixed-order
    predicate (no false positive)."""
    bad = [(i, j) for i, j in pairs if not oracle.pair_within(P[i], Q[j], eps)]
    assert not bad, bad[:10]
