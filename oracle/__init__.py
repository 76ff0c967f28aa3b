"""ctypes binding of the CPU oracle (oracle/hj_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use it as the checker (or, for cpu_baseline, as
the timed CPU restatement of the reference's algorithm).  The product package
hilbertjoin_amd never imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

P = ctypes.c_void_p
I64, U64, INT, DBL = ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
_SIGS = {
    "hjo_cell": (DBL, [DBL, DBL]),
    "hjo_ego_compare": (INT, [P, P, INT, DBL]),
    "hjo_dist2": (DBL, [P, P, INT]),
    "hjo_pair_within": (INT, [P, P, INT, DBL]),
    "hjo_self_count": (U64, [P, I64, INT, DBL, INT]),
    "hjo_join_count": (U64, [P, I64, P, I64, INT, DBL, INT]),
    "hjo_pair_within_expanded": (INT, [P, P, INT, DBL, INT]),
    "hjo_self_count_expanded": (U64, [P, I64, INT, DBL, INT]),
    "hjo_join_count_expanded": (U64, [P, I64, P, I64, INT, DBL, INT]),
    "hjo_self_disagreements": (I64, [P, I64, INT, DBL, INT, P, I64]),
    "hjo_join_disagreements": (I64, [P, I64, P, I64, INT, DBL, INT, P, I64]),
    "hjo_self_band": (I64, [P, I64, INT, DBL, DBL, P, I64]),
    "hjo_self_csr": (U64, [P, I64, INT, DBL, P, P]),
    "hjo_join_csr": (U64, [P, I64, P, I64, INT, DBL, P, P]),
    "hjo_ego_sort": (None, [P, I64, INT, DBL, P]),
    "hjo_ego_self_count": (U64, [P, I64, INT, DBL, INT, INT, INT, P, P]),
    "hjo_ego_join_count": (U64, [P, I64, P, I64, INT, DBL, INT, INT, INT, P, P]),
    "hjo_ego_self_count_sampled": (U64, [P, I64, INT, DBL, INT, INT, INT, I64, P, P]),
    "hjo_hilbert_xy2d": (U64, [INT, ctypes.c_uint32, ctypes.c_uint32]),
    "hjo_hilbert_d2xy": (None, [INT, U64, P, P]),
    "hjo_dim_costs": (INT, [P, I64, INT, DBL, P]),
    "hjo_max_threads": (INT, []),
}

_lib = None


def build() -> Path:
    """Compiles the oracle with its Makefile (gcc, no GPU needed)."""
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f64(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.ndim == 2
    return a


def cell(x: float, eps: float) -> float:
    return lib().hjo_cell(float(x), float(eps))


def ego_compare(a, b, eps: float) -> int:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().hjo_ego_compare(a.ctypes.data, b.ctypes.data, a.size, float(eps))


def dist2(a, b) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().hjo_dist2(a.ctypes.data, b.ctypes.data, a.size)


def self_count(X, eps: float, early_exit: bool = True) -> int:
    X = _f64(X)
    return lib().hjo_self_count(X.ctypes.data, X.shape[0], X.shape[1], float(eps),
                                int(early_exit))


def join_count(A, B, eps: float, early_exit: bool = True) -> int:
    A, B = _f64(A), _f64(B)
    return lib().hjo_join_count(A.ctypes.data, A.shape[0], B.ctypes.data, B.shape[0],
                                A.shape[1], float(eps), int(early_exit))


def self_count_expanded(X, eps: float, kblock: int = 4) -> int:
    X = _f64(X)
    return lib().hjo_self_count_expanded(X.ctypes.data, X.shape[0], X.shape[1], float(eps),
                                         kblock)


def pair_within(a, b, eps: float) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return bool(lib().hjo_pair_within(a.ctypes.data, b.ctypes.data, a.size, float(eps)))


def pair_within_expanded(a, b, eps: float, kblock: int = 4) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return bool(lib().hjo_pair_within_expanded(a.ctypes.data, b.ctypes.data, a.size, float(eps),
                                               kblock))


def join_count_expanded(A, B, eps: float, kblock: int = 4) -> int:
    A, B = _f64(A), _f64(B)
    return lib().hjo_join_count_expanded(A.ctypes.data, A.shape[0], B.ctypes.data, B.shape[0],
                                         A.shape[1], float(eps), kblock)


def join_disagreements(A, B, eps: float, kblock: int = 4, cap: int = 1 << 16) -> np.ndarray:
    """Pairs (a, b) on which the reference's A x B tile arithmetic can
    disagree with the fixed-order predicate (hjo_join_disagreements)."""
    A, B = _f64(A), _f64(B)
    out = np.zeros((cap, 2), dtype=np.int64)
    k = lib().hjo_join_disagreements(A.ctypes.data, A.shape[0], B.ctypes.data, B.shape[0],
                                     A.shape[1], float(eps), kblock, out.ctypes.data, cap)
    return out[:min(k, cap)]


def self_disagreements(X, eps: float, kblock: int = 4, cap: int = 1 << 16) -> np.ndarray:
    X = _f64(X)
    out = np.zeros((cap, 2), dtype=np.int64)
    k = lib().hjo_self_disagreements(X.ctypes.data, X.shape[0], X.shape[1], float(eps), kblock,
                                     out.ctypes.data, cap)
    return out[:min(k, cap)]


def self_band(X, eps: float, tol: float, cap: int = 1 << 16) -> np.ndarray:
    X = _f64(X)
    out = np.zeros((cap, 2), dtype=np.int64)
    k = lib().hjo_self_band(X.ctypes.data, X.shape[0], X.shape[1], float(eps), float(tol),
                            out.ctypes.data, cap)
    return out[:min(k, cap)]


def self_csr(X, eps: float):
    X = _f64(X)
    n = X.shape[0]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    tot = lib().hjo_self_csr(X.ctypes.data, n, X.shape[1], float(eps), rowptr.ctypes.data, None)
    cols = np.zeros(max(tot, 1), dtype=np.int64)
    lib().hjo_self_csr(X.ctypes.data, n, X.shape[1], float(eps), rowptr.ctypes.data,
                       cols.ctypes.data)
    return rowptr, cols[:tot]


def join_csr(A, B, eps: float):
    A, B = _f64(A), _f64(B)
    n = A.shape[0]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    args = (A.ctypes.data, n, B.ctypes.data, B.shape[0], A.shape[1], float(eps),
            rowptr.ctypes.data)
    tot = lib().hjo_join_csr(*args, None)
    cols = np.zeros(max(tot, 1), dtype=np.int64)
    lib().hjo_join_csr(*args, cols.ctypes.data)
    return rowptr, cols[:tot]


def ego_sort(X, eps: float) -> np.ndarray:
    X = _f64(X)
    perm = np.zeros(max(X.shape[0], 1), dtype=np.int64)
    lib().hjo_ego_sort(X.ctypes.data, X.shape[0], X.shape[1], float(eps), perm.ctypes.data)
    return perm[:X.shape[0]]


def ego_self_count(X, eps: float, active_dims: int = 3, kblock: int = 4, threads: int = 0,
                   stride: int = 1):
    """(count, (sort_s, index_s, join_s), evaluated pairs) of the CPU EGO join."""
    X = _f64(X)
    t = np.zeros(3)
    ev = U64()
    if stride != 1:
        return ego_self_count_sampled(X, eps, stride, active_dims, kblock, threads)
    c = lib().hjo_ego_self_count(X.ctypes.data, X.shape[0], X.shape[1], float(eps),
                                 active_dims, kblock, threads, t.ctypes.data, ctypes.byref(ev))
    return c, tuple(t), ev.value


def ego_self_count_sampled(X, eps: float, stride: int, active_dims: int = 3, kblock: int = 4,
                           threads: int = 0):
    """(count, times, evaluated pairs) of the CPU EGO join with the join loop run
    on every stride-th 8-point block of the sorted set only."""
    X = _f64(X)
    t = np.zeros(3)
    ev = U64()
    c = lib().hjo_ego_self_count_sampled(X.ctypes.data, X.shape[0], X.shape[1], float(eps),
                                         active_dims, kblock, threads, int(stride), t.ctypes.data,
                                         ctypes.byref(ev))
    return c, tuple(t), ev.value


def ego_join_count(A, B, eps: float, active_dims: int = 2, kblock: int = 4, threads: int = 0):
    A, B = _f64(A), _f64(B)
    t = np.zeros(3)
    ev = U64()
    c = lib().hjo_ego_join_count(A.ctypes.data, A.shape[0], B.ctypes.data, B.shape[0],
                                 A.shape[1], float(eps), active_dims, kblock, threads,
                                 t.ctypes.data, ctypes.byref(ev))
    return c, tuple(t), ev.value


def dim_costs(X, eps: float) -> np.ndarray:
    """Exact per-dimension cost of outputStatistics (hjo_dim_costs)."""
    X = _f64(X)
    out = np.zeros(X.shape[1], dtype=np.uint64)
    rc = lib().hjo_dim_costs(X.ctypes.data, X.shape[0], X.shape[1], float(eps), out.ctypes.data)
    if rc != 0:
        raise ValueError("a dimension spans too many cells")
    return out


def hilbert_xy2d(order: int, x: int, y: int) -> int:
    return lib().hjo_hilbert_xy2d(order, x, y)


def hilbert_d2xy(order: int, h: int) -> tuple[int, int]:
    x, y = ctypes.c_uint32(), ctypes.c_uint32()
    lib().hjo_hilbert_d2xy(order, h, ctypes.byref(x), ctypes.byref(y))
    return x.value, y.value


def max_threads() -> int:
    return lib().hjo_max_threads()
