"""CPU checks of the C ABI library: it loads, exports every declared entry
point and the reference-compatible C++ symbols, and its host-side services
(generators, reader, command-line argument handling) behave as documented.
No device calls are made."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import hilbertjoin_amd as hj

ROOT = Path(__file__).resolve().parent.parent


def _dynsyms():
    out = subprocess.run(["nm", "-D", "--defined-only", str(hj.library_path())],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_symbol():
    hj.load_library()
    declared = hj.exported_symbols()
    assert len(declared) >= 18
    syms = _dynsyms()
    missing = [s for s in declared if s not in syms]
    assert not missing


def test_compat_symbols_have_reference_signatures():
    # egojoin.h:56 and :64, C++ linkage
    syms = _dynsyms()
    assert "_Z20test_ego_loop3_macrommdPdPmiS_S_S_" in syms
    assert "_Z21test_ego_loop3_noselfmmidPdS_PmiS_S_S_" in syms
    # egojoin.h:59-60: outputStatistics(int, int, double, double*, int*),
    # reorder_dimensions(int, int, double*, int*)
    assert "_Z16outputStatisticsiidPdPi" in syms
    assert "_Z18reorder_dimensionsiiPdPi" in syms
    text = (ROOT / "include" / "egojoin_compat.h").read_text()
    assert "void test_ego_loop3_macro(size_t n, size_t d, double epsilon, double *array," in text


def test_every_header_function_cites_its_reference_or_says_it_has_none():
    text = (ROOT / "include" / "hilbertjoin_amd.h").read_text()
    blocks = re.split(r"\n\s*\n", text)
    for b in blocks:
        if "HJ_API int" in b or "HJ_API void" in b or "HJ_API const" in b:
            fns = re.findall(r"\b(hj_\w+)\s*\(", b)
            assert fns


def _mix64(z):
    z = np.uint64(z)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def test_uniform_generator_formula():
    with np.errstate(over="ignore"):
        seed, n, d = 5, 7, 3
        got = hj.gen_uniform(seed, n, d)
        k = np.arange(n * d, dtype=np.uint64)
        s = _mix64(np.uint64(seed))
        x = _mix64(s + (k + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15))
        want = (x >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    assert np.array_equal(got.ravel(), want)
    assert got.min() >= 0 and got.max() < 1
    assert not np.array_equal(hj.gen_uniform(1, 4, 4), hj.gen_uniform(2, 4, 4))


def test_mixture_generator_shape_and_range():
    X = hj.gen_mixture(3, 20000, 28, 64, 0.04)
    assert X.shape == (20000, 28)
    assert X.min() >= 0.0 and X.max() < 1.0
    assert np.array_equal(X, hj.gen_mixture(3, 20000, 28, 64, 0.04))
    # clustered: nearest-centre spread close to sigma
    assert 0.02 < np.median(np.abs(X - np.median(X, axis=0))) < 0.4


def test_read_points_text_tokens_like_strtok(tmp_path):
    # the reference tokenizes each fgets line with strtok(line, ","): runs of
    # separators collapse, '%' lines are comments, the newline left after a
    # trailing ',' (or alone on a line) is a token of value 0
    (tmp_path / "t.csv").write_text("%comment\n1,,2\n,3,4\n5,6,\n,,\n7\n")
    got = hj.read_points(tmp_path / "t.csv", 5, 2)
    assert got.tolist() == [[1, 2], [3, 4], [5, 6], [0, 0], [7, 0]]
    (tmp_path / "u.csv").write_text("1,\n2.5,-1e3\n")
    assert hj.read_points(tmp_path / "u.csv", 2, 3).tolist() == [[1, 0, 0], [2.5, -1000, 0]]


def test_read_points_bin_and_csv(tmp_path):
    X = hj.gen_uniform(9, 50, 4)
    X.tofile(tmp_path / "p.bin")
    assert np.array_equal(hj.read_points(tmp_path / "p.bin", 50, 4), X)
    lines = ["% comment"] + [",".join(repr(float(v)) for v in row) for row in X[:10]]
    (tmp_path / "p.csv").write_text("\n".join(lines) + "\n")
    assert np.array_equal(hj.read_points(tmp_path / "p.csv", 10, 4), X[:10])
    with pytest.raises(hj.HJError):
        hj.read_points(tmp_path / "p.bin", 51, 4)  # short file
    with pytest.raises(hj.HJError):
        hj.read_points(tmp_path / "missing.bin", 5, 4)


def test_cli_usage_and_bad_flags():
    for tool in ("hilbertSelfJoinCardinality", "hilbertJoinCardinality"):
        exe = ROOT / "hilbertjoin_amd" / "bin" / tool
        r = subprocess.run([str(exe)], capture_output=True, text=True)
        assert r.returncode == 1 and "obligatory" in r.stderr
        r = subprocess.run([str(exe), "-n", "10", "-e", "0.1", "-d", "2", "-s", "9"],
                           capture_output=True, text=True)
        assert r.returncode == 1 and "active dimensions" in r.stderr
        r = subprocess.run([str(exe), "-n", "10", "-e", "0.1", "-d", "2", "--kernel=xx"],
                           capture_output=True, text=True)
        assert r.returncode == 1


def test_context_on_missing_device_reports_error():
    # device 4096 does not exist on any box; the error is a status, not a crash
    with pytest.raises(hj.HJError):
        hj.Context(4096)


def test_device_query_shapes_checked_before_any_device_call():
    # a device query batch whose rows, columns or dtype disagree with the
    # offsets or the index is rejected in Python: the C entry point cannot see
    # the size of a device buffer
    class _Idx:
        d = 8
        _ctx = None
        _lib = None
        _h = None
        _batch = hj.Index._batch
    dq = object.__new__(hj.DeviceArray)
    dq.shape, dq.dtype = (100, 8), np.dtype(np.float64)
    for shape, dtype, off in (((100, 8), np.float64, [0, 50, 101]),
                              ((100, 9), np.float64, [0, 100]),
                              ((100, 8), np.int64, [0, 100]),
                              ((100, 8), np.float64, [0, 60, 40, 100])):
        dq.shape, dq.dtype = shape, np.dtype(dtype)
        with pytest.raises(ValueError):
            hj.Index.query_count(_Idx(), dq, off, 0.1, ctx=object())
