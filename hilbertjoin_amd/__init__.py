"""hilbertjoin_amd -- epsilon similarity join on one AMD Instinct MI355X.

A thin ctypes binding of the C ABI in include/hilbertjoin_amd.h, shaped like
the reference's two join entry points (test_ego_loop3_macro /
test_ego_loop3_noself, hilbertjoin/egojoin.cpp:777-907): a self-join count and
an A x B join count of FP64 points, plus CSR pair lists, the EGO order and the
tile-pair work list for inspection.

Inputs are numpy arrays (host) or DeviceArray objects (point sets kept
resident in HBM through the library's own device buffers; the join then starts
from HBM).  Nothing here computes a join on the CPU: without the built library
every call raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = [
    "HJError", "Context", "Stats", "DeviceArray", "KERNEL_FP64", "KERNEL_MFMA_F16", "KERNEL_MFMA_BF16",
    "ORDER_HILBERT", "ORDER_ROW", "library_path", "load_library", "gen_uniform",
    "gen_mixture", "read_points", "exported_symbols", "Index", "Server", "Result", "filter_band",
]

KERNEL_FP64, KERNEL_MFMA_F16, KERNEL_MFMA_BF16 = 0, 1, 2
ORDER_HILBERT, ORDER_ROW = 0, 1
KERNEL_NAMES = {"fp64": KERNEL_FP64, "f16": KERNEL_MFMA_F16, "bf16": KERNEL_MFMA_BF16}

_ERRORS = {1: "invalid argument", 2: "HIP error", 3: "out of memory", 4: "capacity",
           5: "internal error", 6: "timeout"}

_PKG = Path(__file__).resolve().parent
_INCLUDE = _PKG.parent / "include"


def library_path() -> Path:
    return _PKG / "libhilbertjoin_amd.so"


class HJError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


class _Opts(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int32), ("active_dims", ctypes.c_int32),
                ("order", ctypes.c_int32), ("kblock", ctypes.c_int32), ("grid", ctypes.c_int32),
                ("filter_split", ctypes.c_int32), ("tile_rows", ctypes.c_int32),
                ("mfma_chunk", ctypes.c_int32), ("tune", ctypes.c_int32),
                ("filter_frames", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2)]


class _Stats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("grid_ms", ctypes.c_double),
                ("sort_ms", ctypes.c_double), ("index_ms", ctypes.c_double),
                ("join_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("tile_pairs", ctypes.c_uint64),
                ("candidate_pairs", ctypes.c_uint64), ("pair_dims", ctypes.c_uint64),
                ("band_pairs", ctypes.c_uint64), ("load_fraction", ctypes.c_double),
                ("key_dims", ctypes.c_int32), ("key_bits", ctypes.c_int32),
                ("active_dims", ctypes.c_int32), ("kernel", ctypes.c_int32),
                ("early_exits", ctypes.c_uint64), ("frames", ctypes.c_int32),
                ("sampled_ranges", ctypes.c_int32)]


@dataclass
class Stats:
    total_ms: float
    grid_ms: float
    sort_ms: float
    index_ms: float
    join_ms: float
    kernel_ms: float
    tile_pairs: int
    candidate_pairs: int
    pair_dims: int
    band_pairs: int
    load_fraction: float
    key_dims: int
    key_bits: int
    active_dims: int
    kernel: int
    early_exits: int
    frames: int
    sampled_ranges: int

    @classmethod
    def _from(cls, s: _Stats) -> "Stats":
        return cls(**{f: getattr(s, f) for f, _ in _Stats._fields_})


_lib = None

P = ctypes.c_void_p
class _TuneReport(ctypes.Structure):
    _fields_ = [("trials", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("base_ms", ctypes.c_double), ("best_ms", ctypes.c_double), ("best", _Opts)]


_KERNEL_BY_ID = {0: "fp64", 1: "f16", 2: "bf16"}


def _opts_dict(o: _Opts) -> dict:
    return {"kernel": _KERNEL_BY_ID[o.kernel], "active_dims": o.active_dims,
            "order": "hilbert" if o.order == ORDER_HILBERT else "row", "kblock": o.kblock,
            "grid": o.grid, "filter_split": o.filter_split, "tile_rows": o.tile_rows,
            "mfma_chunk": o.mfma_chunk, "filter_frames": o.filter_frames == 0}


class _ServerOpts(ctypes.Structure):
    _fields_ = [("workers", ctypes.c_int32), ("kernel", ctypes.c_int32),
                ("max_batch_requests", ctypes.c_uint32), ("reserved0", ctypes.c_uint32),
                ("max_batch_points", ctypes.c_uint64), ("result_bytes", ctypes.c_uint64),
                ("order", ctypes.c_int32), ("reserved1", ctypes.c_int32 * 3)]


class _Result(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("status", ctypes.c_int32),
                ("pairs", ctypes.c_int32), ("nq", ctypes.c_uint64), ("count", ctypes.c_uint64),
                ("rowptr", ctypes.POINTER(ctypes.c_int64)),
                ("colidx", ctypes.POINTER(ctypes.c_int64)), ("queue_ms", ctypes.c_double),
                ("run_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("batch", ctypes.c_uint64), ("batch_requests", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("pending", ctypes.c_uint64), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("requests_per_s", ctypes.c_double),
                ("mean_batch_requests", ctypes.c_double),
                ("result_bytes_used", ctypes.c_uint64)]


U64, U32, I32, DBL = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double
_SIGS = {
    "hj_opts_default": (None, [P]),
    "hj_ctx_create": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_int]),
    "hj_ctx_destroy": (ctypes.c_int, [P]),
    "hj_last_error": (ctypes.c_char_p, [P]),
    "hj_device_info": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p,
                                      ctypes.c_int]),
    "hj_self_join_count": (ctypes.c_int, [P, P, U64, U32, DBL, P, P, P]),
    "hj_join_count": (ctypes.c_int, [P, P, U64, P, U64, U32, DBL, P, P, P]),
    "hj_self_join_count_device": (ctypes.c_int, [P, P, U64, U32, DBL, P, P, P]),
    "hj_join_count_device": (ctypes.c_int, [P, P, U64, P, U64, U32, DBL, P, P, P]),
    "hj_self_join_count_device_noalloc": (ctypes.c_int, [P, P, U64, U32, DBL, P, P, P]),
    "hj_batch_self_join_count": (ctypes.c_int, [P, P, P, U32, U32, P, P, P, P]),
    "hj_batch_join_count": (ctypes.c_int, [P, P, P, P, P, U32, U32, P, P, P, P]),
    "hj_batch_self_join_count_device": (ctypes.c_int, [P, P, P, U32, U32, P, P, P, P]),
    "hj_batch_join_count_device": (ctypes.c_int, [P, P, P, P, P, U32, U32, P, P, P, P]),
    "hj_self_join_csr": (ctypes.c_int, [P, P, U64, U32, DBL, P, P, P, U64, P, P]),
    "hj_join_csr": (ctypes.c_int, [P, P, U64, P, U64, U32, DBL, P, P, P, U64, P, P]),
    "hj_self_join_csr_device": (ctypes.c_int, [P, P, U64, U32, DBL, P, P, P, U64, P, P]),
    "hj_join_csr_device": (ctypes.c_int, [P, P, U64, P, U64, U32, DBL, P, P, P, U64, P, P]),
    "hj_ego_sort": (ctypes.c_int, [P, P, U64, U32, DBL, P, P, P, P]),
    "hj_self_tile_pairs": (ctypes.c_int, [P, P, U64, U32, DBL, P, P, U64, P, P]),
    "hj_join_tile_pairs": (ctypes.c_int, [P, P, U64, P, U64, U32, DBL, P, P, U64, P, P, P]),
    "hj_hilbert_index": (ctypes.c_int, [P, ctypes.c_int, P, P, U64, P]),
    "hj_read_points": (ctypes.c_int, [ctypes.c_char_p, U64, U32, P]),
    "hj_device_alloc": (ctypes.c_int, [P, U64, ctypes.POINTER(P)]),
    "hj_device_free": (ctypes.c_int, [P, P]),
    "hj_memcpy_to_device": (ctypes.c_int, [P, P, P, U64]),
    "hj_memcpy_to_host": (ctypes.c_int, [P, P, P, U64]),
    "hj_device_copy": (ctypes.c_int, [P, P, P, U64]),
    "hj_index_create": (ctypes.c_int, [P, P, U64, U32, DBL, P, ctypes.POINTER(P)]),
    "hj_index_create_device": (ctypes.c_int, [P, P, U64, U32, DBL, P, ctypes.POINTER(P)]),
    "hj_index_destroy": (ctypes.c_int, [P]),
    "hj_index_info": (ctypes.c_int, [P, P, P, P, P, P, P]),
    "hj_index_query_count": (ctypes.c_int, [P, P, P, P, U32, P, P, P, P]),
    "hj_index_query_count_device": (ctypes.c_int, [P, P, P, P, U32, P, P, P, P]),
    "hj_index_query_csr": (ctypes.c_int, [P, P, P, P, U32, P, P, P, P, U64, P, P]),
    "hj_server_opts_default": (None, [P]),
    "hj_server_create": (ctypes.c_int, [P, ctypes.c_int, P, ctypes.POINTER(P)]),
    "hj_server_submit": (ctypes.c_int, [P, P, U64, DBL, ctypes.c_int, P]),
    "hj_server_wait": (ctypes.c_int, [P, U64, DBL, P]),
    "hj_server_release": (ctypes.c_int, [P, U64]),
    "hj_server_stats_get": (ctypes.c_int, [P, P]),
    "hj_server_stats_reset": (ctypes.c_int, [P]),
    "hj_server_destroy": (ctypes.c_int, [P]),
    "hj_server_last_error": (ctypes.c_char_p, [P]),
    "hj_autotune_self": (ctypes.c_int, [P, P, U64, U32, DBL, P, P]),
    "hj_autotune_join": (ctypes.c_int, [P, P, U64, P, U64, U32, DBL, P, P]),
    "hj_autotune_batch_self": (ctypes.c_int, [P, P, P, U32, U32, P, P, P]),
    "hj_dim_statistics": (ctypes.c_int, [P, P, U64, U32, DBL, P, P]),
    "hj_reorder_dimensions": (ctypes.c_int, [P, P, U64, U32, P]),
    "hj_filter_band": (ctypes.c_int, [I32, U32, DBL, DBL, I32, P]),
    "hj_filter_pair_band": (ctypes.c_int, [I32, U32, DBL, DBL, I32, ctypes.c_float,
                                           ctypes.c_float, ctypes.c_float, ctypes.c_float, P, P]),
    "hj_gen_uniform": (None, [U64, U64, U32, P]),
    "hj_gen_mixture": (None, [U64, U64, U32, U32, DBL, DBL, P]),
}


def load_library(path: str | os.PathLike | None = None):
    """Loads libhilbertjoin_amd.so (built by hilbertjoin_amd.build); raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else library_path()
    if not p.exists():
        raise HJError(5, f"{p} is not built; run `python -m hilbertjoin_amd.build`")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def exported_symbols() -> list[str]:
    """Names of the functions declared with HJ_API in include/hilbertjoin_amd.h."""
    import re
    text = (_INCLUDE / "hilbertjoin_amd.h").read_text()
    return re.findall(r"HJ_API\s+[\w\s\*]+?\b(hj_\w+)\s*\(", text)


def _host_f64(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError("points must be a 2-D array (n x d)")
    return a


class DeviceArray:
    """An n x d array resident in device memory (hj_device_alloc): an FP64
    point set, or (dtype int64) a pair list's rowptr / colidx."""

    def __init__(self, ctx: "Context", shape: tuple[int, int], dtype=np.float64):
        self._ctx = ctx
        self.shape = (int(shape[0]), int(shape[1]))
        self.dtype = np.dtype(dtype)
        if self.dtype.itemsize != 8:
            raise ValueError("device arrays hold 8-byte elements")
        self.ptr = P()
        ctx._check(ctx._lib.hj_device_alloc(ctx._h, max(self.nbytes, 8), ctypes.byref(self.ptr)))

    @property
    def nbytes(self) -> int:
        return self.shape[0] * self.shape[1] * 8

    def data_ptr(self) -> int:
        return self.ptr.value or 0

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        self._ctx._check(self._ctx._lib.hj_memcpy_to_host(self._ctx._h, out.ctypes.data,
                                                          self.ptr, self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr and self._ctx._h:
            self._ctx._lib.hj_device_free(self._ctx._h, self.ptr)
        self.ptr = P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _is_device(a) -> bool:
    return isinstance(a, DeviceArray)


def _dev_f64(t: DeviceArray) -> DeviceArray:
    return t


def _opts(kernel="fp64", active_dims=-1, order=ORDER_ROW, kblock=0, grid=0,
          filter_split=0, tile_rows=0, mfma_chunk=0, tune=False, filter_frames=True) -> _Opts:
    o = _Opts()
    load_library().hj_opts_default(ctypes.byref(o))
    o.kernel = KERNEL_NAMES[kernel] if isinstance(kernel, str) else int(kernel)
    o.active_dims = int(active_dims)
    o.order = ORDER_HILBERT if order in ("hilbert", ORDER_HILBERT) else ORDER_ROW
    o.kblock = int(kblock)
    o.grid = int(grid)
    o.filter_split = int(filter_split)
    o.tile_rows = int(tile_rows)
    o.mfma_chunk = int(mfma_chunk)
    o.tune = 1 if tune else 0
    o.filter_frames = 0 if filter_frames else -1
    return o


class Context:
    """One HIP device + stream + grow-only device buffers (hj_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        self._h = P()
        rc = self._lib.hj_ctx_create(ctypes.byref(self._h), int(device))
        if rc != 0:
            raise HJError(rc, f"cannot create a context on device {device}")

    def close(self) -> None:
        if self._h:
            self._lib.hj_ctx_destroy(self._h)
            self._h = P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise HJError(rc, self._lib.hj_last_error(self._h).decode())

    def to_device(self, pts) -> DeviceArray:
        """Copies an n x d array into device memory owned by this context."""
        a = _host_f64(pts)
        dev = DeviceArray(self, a.shape)
        self._check(self._lib.hj_memcpy_to_device(self._h, dev.ptr, a.ctypes.data, dev.nbytes))
        return dev

    def device_copy(self, dst: DeviceArray, src: DeviceArray) -> None:
        """Copies src into dst (same shape) with the library's streaming copy kernel."""
        if dst.shape != src.shape:
            raise ValueError("shapes differ")
        self._check(self._lib.hj_device_copy(self._h, dst.ptr, src.ptr, src.nbytes))

    def device_info(self) -> tuple[int, str]:
        cus = ctypes.c_int()
        name = ctypes.create_string_buffer(128)
        self._check(self._lib.hj_device_info(self._h, ctypes.byref(cus), name, 128))
        return cus.value, name.value.decode()

    # ---- counts -------------------------------------------------------------
    def self_join_count(self, pts, eps: float, **kw) -> tuple[int, Stats]:
        """Unordered pairs i<j with ||x_i - x_j|| <= eps (test_ego_loop3_macro)."""
        o, st, cnt = _opts(**kw), _Stats(), U64()
        if _is_device(pts):
            t = _dev_f64(pts)
            rc = self._lib.hj_self_join_count_device(self._h, t.data_ptr(), t.shape[0],
                                                     t.shape[1], float(eps), ctypes.byref(o),
                                                     ctypes.byref(cnt), ctypes.byref(st))
        else:
            a = _host_f64(pts)
            rc = self._lib.hj_self_join_count(self._h, a.ctypes.data, a.shape[0], a.shape[1],
                                              float(eps), ctypes.byref(o), ctypes.byref(cnt),
                                              ctypes.byref(st))
        self._check(rc)
        return cnt.value, Stats._from(st)

    def self_join_count_noalloc(self, pts: DeviceArray, eps: float, **kw) -> tuple[int, Stats]:
        """self_join_count of device-resident points without any allocation
        (hj_self_join_count_device_noalloc): the library's fixed per-GPU scratch
        array holds every internal buffer; joins that need more fail with a
        capacity error."""
        if not _is_device(pts) or pts.dtype != np.float64:
            raise ValueError("the no-allocation join takes float64 points in a DeviceArray")
        o, st, cnt = _opts(**kw), _Stats(), U64()
        rc = self._lib.hj_self_join_count_device_noalloc(self._h, pts.data_ptr(), pts.shape[0],
                                                         pts.shape[1], float(eps),
                                                         ctypes.byref(o), ctypes.byref(cnt),
                                                         ctypes.byref(st))
        self._check(rc)
        return cnt.value, Stats._from(st)

    def join_count(self, A, B, eps: float, **kw) -> tuple[int, Stats]:
        """Pairs (a, b) in A x B with ||a - b|| <= eps (test_ego_loop3_noself)."""
        o, st, cnt = _opts(**kw), _Stats(), U64()
        if _is_device(A):
            a, b = _dev_f64(A), _dev_f64(B)
            if a.shape[1] != b.shape[1]:
                raise ValueError("A and B must have the same dimensionality")
            rc = self._lib.hj_join_count_device(self._h, a.data_ptr(), a.shape[0], b.data_ptr(),
                                                b.shape[0], a.shape[1], float(eps),
                                                ctypes.byref(o), ctypes.byref(cnt),
                                                ctypes.byref(st))
        else:
            a, b = _host_f64(A), _host_f64(B)
            if a.shape[1] != b.shape[1]:
                raise ValueError("A and B must have the same dimensionality")
            rc = self._lib.hj_join_count(self._h, a.ctypes.data, a.shape[0], b.ctypes.data,
                                         b.shape[0], a.shape[1], float(eps), ctypes.byref(o),
                                         ctypes.byref(cnt), ctypes.byref(st))
        self._check(rc)
        return cnt.value, Stats._from(st)

    # ---- batched joins ------------------------------------------------------
    @staticmethod
    def _batch_args(offsets, eps, njobs_hint=None):
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        if off.ndim != 1 or off.size < 1:
            raise ValueError("offsets must be a 1-D array of njobs + 1 row offsets")
        nj = off.size - 1
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, dtype=np.float64), (nj,)))
        return off, e, nj

    def batch_self_join_count(self, pts, offsets, eps, **kw) -> tuple[np.ndarray, Stats]:
        """Self-join counts of the row ranges [offsets[k], offsets[k+1]) of pts,
        job k under radius eps[k] (a scalar applies to every job), in one launch."""
        off, e, nj = self._batch_args(offsets, eps)
        o, st = _opts(**kw), _Stats()
        counts = np.zeros(nj, dtype=np.uint64)
        if _is_device(pts):
            t = _dev_f64(pts)
            if off[-1] != t.shape[0]:
                raise ValueError("offsets[-1] must equal the number of rows")
            rc = self._lib.hj_batch_self_join_count_device(
                self._h, t.data_ptr(), off.ctypes.data, nj, t.shape[1], e.ctypes.data,
                ctypes.byref(o), counts.ctypes.data, ctypes.byref(st))
        else:
            a = _host_f64(pts)
            if off[-1] != a.shape[0]:
                raise ValueError("offsets[-1] must equal the number of rows")
            rc = self._lib.hj_batch_self_join_count(
                self._h, a.ctypes.data, off.ctypes.data, nj, a.shape[1], e.ctypes.data,
                ctypes.byref(o), counts.ctypes.data, ctypes.byref(st))
        self._check(rc)
        return counts, Stats._from(st)

    def batch_join_count(self, A, a_offsets, B, b_offsets, eps, **kw) -> tuple[np.ndarray, Stats]:
        """A x B join counts of the row-range pairs ([a_offsets[k], a_offsets[k+1]),
        [b_offsets[k], b_offsets[k+1])) under radius eps[k], in one launch."""
        aoff, e, nj = self._batch_args(a_offsets, eps)
        boff = np.ascontiguousarray(b_offsets, dtype=np.uint64)
        if boff.shape != aoff.shape:
            raise ValueError("a_offsets and b_offsets must have the same length")
        o, st = _opts(**kw), _Stats()
        counts = np.zeros(nj, dtype=np.uint64)
        if _is_device(A):
            a, b = _dev_f64(A), _dev_f64(B)
            fn, pa, pb = self._lib.hj_batch_join_count_device, a.data_ptr(), b.data_ptr()
        else:
            a, b = _host_f64(A), _host_f64(B)
            fn, pa, pb = self._lib.hj_batch_join_count, a.ctypes.data, b.ctypes.data
        if a.shape[1] != b.shape[1]:
            raise ValueError("A and B must have the same dimensionality")
        if aoff[-1] != a.shape[0] or boff[-1] != b.shape[0]:
            raise ValueError("the last offsets must equal the numbers of rows")
        rc = fn(self._h, pa, aoff.ctypes.data, pb, boff.ctypes.data, nj, a.shape[1],
                e.ctypes.data, ctypes.byref(o), counts.ctypes.data, ctypes.byref(st))
        self._check(rc)
        return counts, Stats._from(st)

    # ---- autotuner -----------------------------------------------------------
    def autotune_self(self, pts: DeviceArray, eps: float, **kw) -> dict:
        """Times candidate settings of this self-join (device points) and returns
        {"best": opts dict, "trials", "base_ms", "best_ms"}; the result is also
        cached in the context for calls with tune=True."""
        o, rep = _opts(**kw), _TuneReport()
        self._check(self._lib.hj_autotune_self(self._h, pts.data_ptr(), pts.shape[0],
                                               pts.shape[1], float(eps), ctypes.byref(o),
                                               ctypes.byref(rep)))
        return {"best": _opts_dict(rep.best), "trials": rep.trials, "base_ms": rep.base_ms,
                "best_ms": rep.best_ms}

    def autotune_join(self, A: DeviceArray, B: DeviceArray, eps: float, **kw) -> dict:
        o, rep = _opts(**kw), _TuneReport()
        self._check(self._lib.hj_autotune_join(self._h, A.data_ptr(), A.shape[0], B.data_ptr(),
                                               B.shape[0], A.shape[1], float(eps),
                                               ctypes.byref(o), ctypes.byref(rep)))
        return {"best": _opts_dict(rep.best), "trials": rep.trials, "base_ms": rep.base_ms,
                "best_ms": rep.best_ms}

    def autotune_batch_self(self, pts: DeviceArray, offsets, eps, **kw) -> dict:
        off, e, nj = self._batch_args(offsets, eps)
        o, rep = _opts(**kw), _TuneReport()
        self._check(self._lib.hj_autotune_batch_self(self._h, pts.data_ptr(), off.ctypes.data,
                                                     nj, pts.shape[1], e.ctypes.data,
                                                     ctypes.byref(o), ctypes.byref(rep)))
        return {"best": _opts_dict(rep.best), "trials": rep.trials, "base_ms": rep.base_ms,
                "best_ms": rep.best_ms}

    # ---- dimension statistics ------------------------------------------------
    def dim_statistics(self, pts, eps: float):
        """(reorder_dim, cost): cost[k] = pairs whose cells in dimension k are equal
        or adjacent; reorder_dim[k] = position of dimension k by increasing cost
        (outputStatistics)."""
        a = _host_f64(pts)
        order = np.zeros(a.shape[1], dtype=np.int32)
        cost = np.zeros(a.shape[1], dtype=np.uint64)
        self._check(self._lib.hj_dim_statistics(self._h, a.ctypes.data, a.shape[0], a.shape[1],
                                                float(eps), order.ctypes.data,
                                                cost.ctypes.data))
        return order, cost

    def reorder_dimensions(self, pts, reorder_dim) -> np.ndarray:
        """A copy of pts with column k moved to column reorder_dim[k]
        (reorder_dimensions, done on the device)."""
        a = _host_f64(pts).copy()
        perm = np.ascontiguousarray(reorder_dim, dtype=np.int32)
        if perm.shape != (a.shape[1],):
            raise ValueError("reorder_dim must have d entries")
        self._check(self._lib.hj_reorder_dimensions(self._h, a.ctypes.data, a.shape[0],
                                                    a.shape[1], perm.ctypes.data))
        return a

    # ---- pair lists ---------------------------------------------------------
    def self_join_csr(self, pts, eps: float, **kw):
        """CSR over original indices: row i lists every j != i within eps, ascending.

        colidx is sized from a count join first (far cheaper than a pair pass),
        so the pair lists are built once.  Points in a DeviceArray give
        (rowptr, colidx) as int64 DeviceArrays of shape (n + 1, 1) / (total, 1),
        written in HBM (hj_self_join_csr_device)."""
        dev = _is_device(pts)
        a = pts if dev else _host_f64(pts)
        n = a.shape[0]
        cnt, _ = self.self_join_count(a, eps, **kw)
        o, st, tot = _opts(**kw), _Stats(), U64()
        if dev:
            rowptr = DeviceArray(self, (n + 1, 1), np.int64)
            fn, src = self._lib.hj_self_join_csr_device, a.data_ptr()
            alloc = lambda m: DeviceArray(self, (m, 1), np.int64)  # noqa: E731
        else:
            rowptr = np.zeros(n + 1, dtype=np.int64)
            fn, src = self._lib.hj_self_join_csr, a.ctypes.data
            alloc = lambda m: np.empty(m, dtype=np.int64)  # noqa: E731
        ptr = (lambda x: x.data_ptr()) if dev else (lambda x: x.ctypes.data)

        def run(cols, size):
            return fn(self._h, src, n, a.shape[1], float(eps), ctypes.byref(o), ptr(rowptr),
                      ptr(cols), size, ctypes.byref(tot), ctypes.byref(st))
        cols = alloc(max(2 * cnt, 1))
        rc = run(cols, 2 * cnt)
        if rc == 4:  # HJ_ERR_CAPACITY
            cols = alloc(max(tot.value, 1))
            rc = run(cols, tot.value)
        self._check(rc)
        if dev:
            cols.shape = (tot.value, 1)
            return rowptr, cols, Stats._from(st)
        return rowptr, cols[:tot.value], Stats._from(st)

    def join_csr(self, A, B, eps: float, **kw):
        """CSR over original indices: row a lists every b within eps, ascending.

        colidx is sized from a count join first, so the lists are built once.
        DeviceArray inputs give int64 DeviceArray outputs (hj_join_csr_device)."""
        dev = _is_device(A)
        if dev != _is_device(B):
            raise ValueError("A and B must both be host arrays or both DeviceArrays")
        a, b = (A, B) if dev else (_host_f64(A), _host_f64(B))
        cnt, _ = self.join_count(a, b, eps, **kw)
        o, st, tot = _opts(**kw), _Stats(), U64()
        if dev:
            rowptr = DeviceArray(self, (a.shape[0] + 1, 1), np.int64)
            fn, pa, pb = self._lib.hj_join_csr_device, a.data_ptr(), b.data_ptr()
            alloc = lambda m: DeviceArray(self, (m, 1), np.int64)  # noqa: E731
            ptr = lambda x: x.data_ptr()  # noqa: E731
        else:
            rowptr = np.zeros(a.shape[0] + 1, dtype=np.int64)
            fn, pa, pb = self._lib.hj_join_csr, a.ctypes.data, b.ctypes.data
            alloc = lambda m: np.empty(m, dtype=np.int64)  # noqa: E731
            ptr = lambda x: x.ctypes.data  # noqa: E731

        def run(cols, size):
            return fn(self._h, pa, a.shape[0], pb, b.shape[0], a.shape[1], float(eps),
                      ctypes.byref(o), ptr(rowptr), ptr(cols), size, ctypes.byref(tot),
                      ctypes.byref(st))
        cols = alloc(max(cnt, 1))
        rc = run(cols, cnt)
        if rc == 4:  # HJ_ERR_CAPACITY
            cols = alloc(max(tot.value, 1))
            rc = run(cols, tot.value)
        self._check(rc)
        if dev:
            cols.shape = (tot.value, 1)
            return rowptr, cols, Stats._from(st)
        return rowptr, cols[:tot.value], Stats._from(st)

    # ---- building blocks ----------------------------------------------------
    def ego_sort(self, pts, eps: float):
        """(perm, keys, key_dims, key_bits) of the EGO order (epsilonGridOrdering)."""
        a = _host_f64(pts)
        n = a.shape[0]
        perm = np.zeros(max(n, 1), dtype=np.uint32)
        keys = np.zeros(max(n, 1), dtype=np.uint64)
        kd, kb = I32(), I32()
        self._check(self._lib.hj_ego_sort(self._h, a.ctypes.data, n, a.shape[1], float(eps),
                                          perm.ctypes.data, keys.ctypes.data, ctypes.byref(kd),
                                          ctypes.byref(kb)))
        return perm[:n], keys[:n], kd.value, kb.value

    def self_tile_pairs(self, pts, eps: float, **kw):
        """(pairs (k x 2: I, J), perm) of the self-join work list in drain order."""
        a = _host_f64(pts)
        n = a.shape[0]
        o, npairs = _opts(**kw), U64()
        perm = np.zeros(max(n, 1), dtype=np.uint32)
        rc = self._lib.hj_self_tile_pairs(self._h, a.ctypes.data, n, a.shape[1], float(eps),
                                          ctypes.byref(o), None, 0, ctypes.byref(npairs), None)
        if rc not in (0, 4):
            self._check(rc)
        out = np.zeros((max(npairs.value, 1), 2), dtype=np.uint32)
        self._check(self._lib.hj_self_tile_pairs(self._h, a.ctypes.data, n, a.shape[1],
                                                 float(eps), ctypes.byref(o), out.ctypes.data,
                                                 npairs.value, ctypes.byref(npairs),
                                                 perm.ctypes.data))
        return out[:npairs.value], perm[:n]

    def join_tile_pairs(self, A, B, eps: float, **kw):
        """(pairs (k x 2: I over A, J over B), permA, permB) of an A x B work list."""
        a, b = _host_f64(A), _host_f64(B)
        o, npairs = _opts(**kw), U64()
        pa = np.zeros(max(a.shape[0], 1), dtype=np.uint32)
        pb = np.zeros(max(b.shape[0], 1), dtype=np.uint32)
        args = (self._h, a.ctypes.data, a.shape[0], b.ctypes.data, b.shape[0], a.shape[1],
                float(eps), ctypes.byref(o))
        rc = self._lib.hj_join_tile_pairs(*args, None, 0, ctypes.byref(npairs), None, None)
        if rc not in (0, 4):
            self._check(rc)
        out = np.zeros((max(npairs.value, 1), 2), dtype=np.uint32)
        self._check(self._lib.hj_join_tile_pairs(*args, out.ctypes.data, npairs.value,
                                                 ctypes.byref(npairs), pa.ctypes.data,
                                                 pb.ctypes.data))
        return out[:npairs.value], pa[:a.shape[0]], pb[:b.shape[0]]

    def hilbert_index(self, order: int, x, y) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.uint32)
        y = np.ascontiguousarray(y, dtype=np.uint32)
        out = np.zeros(max(x.size, 1), dtype=np.uint64)
        self._check(self._lib.hj_hilbert_index(self._h, int(order), x.ctypes.data, y.ctypes.data,
                                               x.size, out.ctypes.data))
        return out[:x.size]


class Index:
    """A point set EGO-sorted once into HBM (hj_index_create); batches of
    requests (query point sets with their own eps <= eps_max) are joined
    against it."""

    def __init__(self, ctx: Context, pts, eps_max: float, **kw):
        self._ctx = ctx
        self._lib = ctx._lib
        self._h = P()
        o = _opts(**kw)
        if _is_device(pts):
            rc = self._lib.hj_index_create_device(ctx._h, pts.data_ptr(), pts.shape[0],
                                                  pts.shape[1], float(eps_max), ctypes.byref(o),
                                                  ctypes.byref(self._h))
            self.d = pts.shape[1]
        else:
            a = _host_f64(pts)
            rc = self._lib.hj_index_create(ctx._h, a.ctypes.data, a.shape[0], a.shape[1],
                                           float(eps_max), ctypes.byref(o), ctypes.byref(self._h))
            self.d = a.shape[1]
        ctx._check(rc)
        self.eps_max = float(eps_max)

    def info(self) -> dict:
        n, d, e = U64(), U32(), DBL()
        kd, kb, act = I32(), I32(), I32()
        self._lib.hj_index_info(self._h, ctypes.byref(n), ctypes.byref(d), ctypes.byref(e),
                                ctypes.byref(kd), ctypes.byref(kb), ctypes.byref(act))
        return {"n": n.value, "d": d.value, "eps_max": e.value, "key_dims": kd.value,
                "key_bits": kb.value, "active_dims": act.value}

    def _batch(self, Q, offsets, eps):
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        nreq = off.size - 1
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, dtype=np.float64), (nreq,)))
        return off, e, nreq

    def query_count(self, Q, offsets, eps, ctx: Context | None = None, **kw):
        """Counts of the requests rows [offsets[k], offsets[k+1]) of Q under eps[k]."""
        ctx = ctx or self._ctx
        off, e, nreq = self._batch(Q, offsets, eps)
        o, st = _opts(**kw), _Stats()
        counts = np.zeros(max(nreq, 0), dtype=np.uint64)
        if _is_device(Q):
            # the library cannot see the size of a device buffer: check the
            # shape here, before any kernel reads it
            if Q.dtype != np.float64 or Q.shape[1] != self.d:
                raise ValueError(f"query points must be float64 with d = {self.d} columns")
            if off.size < 1 or off[-1] != Q.shape[0] or np.any(np.diff(off.astype(np.int64)) < 0):
                raise ValueError("offsets must be non-decreasing and end at the number of rows")
            rc = self._lib.hj_index_query_count_device(
                ctx._h, self._h, Q.data_ptr(), off.ctypes.data, nreq, e.ctypes.data,
                ctypes.byref(o), counts.ctypes.data, ctypes.byref(st))
        else:
            a = _host_f64(Q) if len(Q) else np.zeros((0, self.d))
            if a.shape[1] != self.d:
                raise ValueError(f"query points must have d = {self.d} columns")
            if off[-1] != a.shape[0]:
                raise ValueError("offsets[-1] must equal the number of query rows")
            rc = self._lib.hj_index_query_count(ctx._h, self._h, a.ctypes.data, off.ctypes.data,
                                                nreq, e.ctypes.data, ctypes.byref(o),
                                                counts.ctypes.data, ctypes.byref(st))
        ctx._check(rc)
        return counts, Stats._from(st)

    def query_csr(self, Q, offsets, eps, ctx: Context | None = None, **kw):
        """Pair lists of a batch: (rowptr over all query rows, colidx of index ids)."""
        ctx = ctx or self._ctx
        off, e, nreq = self._batch(Q, offsets, eps)
        a = _host_f64(Q) if len(Q) else np.zeros((0, self.d))
        if off[-1] != a.shape[0]:
            raise ValueError("offsets[-1] must equal the number of query rows")
        o, st, total = _opts(**kw), _Stats(), U64()
        rowptr = np.zeros(a.shape[0] + 1, dtype=np.int64)
        cap = 1 << 16
        colidx = np.zeros(cap, dtype=np.int64)
        rc = self._lib.hj_index_query_csr(ctx._h, self._h, a.ctypes.data, off.ctypes.data, nreq,
                                          e.ctypes.data, ctypes.byref(o), rowptr.ctypes.data,
                                          colidx.ctypes.data, cap, ctypes.byref(total),
                                          ctypes.byref(st))
        if rc == 4:
            cap = total.value
            colidx = np.zeros(max(cap, 1), dtype=np.int64)
            rc = self._lib.hj_index_query_csr(ctx._h, self._h, a.ctypes.data, off.ctypes.data,
                                              nreq, e.ctypes.data, ctypes.byref(o),
                                              rowptr.ctypes.data, colidx.ctypes.data, cap,
                                              ctypes.byref(total), ctypes.byref(st))
        ctx._check(rc)
        return rowptr, colidx[:total.value].copy(), Stats._from(st)

    def close(self) -> None:
        if self._h:
            self._lib.hj_index_destroy(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Result:
    ticket: int
    status: int
    count: int
    nq: int
    rowptr: np.ndarray | None
    colidx: np.ndarray | None
    queue_ms: float
    run_ms: float
    latency_ms: float
    batch: int
    batch_requests: int


class Server:
    """Serving loop over an Index (hj_server_*): submit requests from any
    thread, wait for their results; worker threads batch the waiting requests
    continuously."""

    def __init__(self, index: Index, device: int = 0, **opts):
        self._lib = index._lib
        self._index = index  # kept alive while the server runs
        o = _ServerOpts()
        self._lib.hj_server_opts_default(ctypes.byref(o))
        for k, v in opts.items():
            if k == "kernel" and isinstance(v, str):
                v = KERNEL_NAMES[v]
            if k == "order" and isinstance(v, str):
                v = ORDER_HILBERT if v == "hilbert" else ORDER_ROW
            setattr(o, k, v)
        self._h = P()
        rc = self._lib.hj_server_create(index._h, int(device), ctypes.byref(o),
                                        ctypes.byref(self._h))
        if rc != 0:
            raise HJError(rc, "cannot create the server")

    def _check(self, rc):
        if rc != 0:
            raise HJError(rc, self._lib.hj_server_last_error(self._h).decode())

    def submit(self, q, eps: float, pairs: bool = False) -> int:
        a = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, self._index.d)
        t = U64()
        self._check(self._lib.hj_server_submit(self._h, a.ctypes.data, a.shape[0], float(eps),
                                               1 if pairs else 0, ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0, release: bool = True) -> Result:
        r = _Result()
        self._check(self._lib.hj_server_wait(self._h, ticket, float(timeout_ms),
                                             ctypes.byref(r)))
        rowptr = colidx = None
        if r.pairs:
            rowptr = np.ctypeslib.as_array(r.rowptr, shape=(r.nq + 1,)).copy()
            colidx = (np.ctypeslib.as_array(r.colidx, shape=(r.count,)).copy() if r.count
                      else np.zeros(0, dtype=np.int64))
        res = Result(r.ticket, r.status, r.count, r.nq, rowptr, colidx, r.queue_ms, r.run_ms,
                     r.latency_ms, r.batch, r.batch_requests)
        if release:
            self._lib.hj_server_release(self._h, ticket)
        return res

    def release(self, ticket: int) -> None:
        self._check(self._lib.hj_server_release(self._h, ticket))

    def stats(self) -> dict:
        st = _ServerStats()
        self._check(self._lib.hj_server_stats_get(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in _ServerStats._fields_}

    def reset_stats(self) -> None:
        self._check(self._lib.hj_server_stats_reset(self._h))

    def close(self) -> None:
        if self._h:
            self._lib.hj_server_destroy(self._h)
            self._h = P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _FilterInfo(ctypes.Structure):
    _fields_ = [("scale", ctypes.c_double), ("t_in", ctypes.c_float), ("t_out", ctypes.c_float),
                ("split", ctypes.c_int32), ("steps1", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("parts", ctypes.c_int32)]


def filter_band(kernel: str, d: int, eps: float, radius: float, filter_split: int = 0) -> dict:
    """Operand scale, K layout and classification thresholds of the MFMA
    filter for points within `radius` of the operand centre (hj_filter_band;
    host only)."""
    fi = _FilterInfo()
    rc = load_library().hj_filter_band(KERNEL_NAMES[kernel], int(d), float(eps), float(radius),
                                       int(filter_split), ctypes.byref(fi))
    if rc != 0:
        raise HJError(rc, "invalid filter parameters")
    return {f: getattr(fi, f) for f, _ in _FilterInfo._fields_}


def filter_pair_band(kernel: str, d: int, eps: float, radius: float, rho_a: float,
                     rho_b: float, dd_bound: float, dd_shift: float,
                     filter_split: int = 0) -> tuple[float, float]:
    """(t_in, t_out) of one tile pair with tile frames (hj_filter_pair_band;
    host only): rho_a, rho_b bound the tiles' operand norms in their frames,
    dd_bound / dd_shift the squared anchor difference over the tested slots."""
    t_in, t_out = ctypes.c_float(), ctypes.c_float()
    rc = load_library().hj_filter_pair_band(KERNEL_NAMES[kernel], int(d), float(eps),
                                            float(radius), int(filter_split), float(rho_a),
                                            float(rho_b), float(dd_bound), float(dd_shift),
                                            ctypes.byref(t_in), ctypes.byref(t_out))
    if rc != 0:
        raise HJError(rc, "invalid filter parameters")
    return t_in.value, t_out.value


def gen_uniform(seed: int, n: int, d: int) -> np.ndarray:
    """Seeded uniform [0,1) points (hj_gen_uniform)."""
    out = np.empty((n, d), dtype=np.float64)
    load_library().hj_gen_uniform(seed, n, d, out.ctypes.data)
    return out


def gen_mixture(seed: int, n: int, d: int, ncomp: int, sigma: float,
                alpha: float = 0.5) -> np.ndarray:
    """Seeded Gaussian-mixture points clipped to [0,1) (hj_gen_mixture)."""
    out = np.empty((n, d), dtype=np.float64)
    load_library().hj_gen_mixture(seed, n, d, ncomp, sigma, alpha, out.ctypes.data)
    return out


def read_points(path: str | os.PathLike, n: int, d: int) -> np.ndarray:
    out = np.empty((n, d), dtype=np.float64)
    rc = load_library().hj_read_points(str(path).encode(), n, d, out.ctypes.data)
    if rc != 0:
        raise HJError(rc, f"cannot read {n} x {d} points from {path}")
    return out
