"""Build recipe of the MI355X epsilon-join library and its command-line tools.

Produces, in-tree:
  hilbertjoin_amd/libhilbertjoin_amd.so        C ABI (include/hilbertjoin_amd.h) and the
                                               reference-compatible entry points
                                               (include/egojoin_compat.h)
  hilbertjoin_amd/bin/hilbertSelfJoinCardinality
  hilbertjoin_amd/bin/hilbertJoinCardinality

Device code is compiled for gfx950 only.  Run:  python -m hilbertjoin_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libhilbertjoin_amd.so"
BIN = PKG / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_SOURCES = ["hj_sort.hip", "hj_prep.hip", "hj_fp64.hip", "hj_mfma.hip"]
HIP_HOST_SOURCES = ["hj_join.cpp", "hj_api.cpp", "hj_serve.cpp", "hj_noalloc.cpp"]  # host code that calls the HIP runtime
CXX_SOURCES = ["hj_host.cpp", "hj_compat.cpp"]
CLIS = {"hilbertSelfJoinCardinality": "cli_self.cpp", "hilbertJoinCardinality": "cli_join.cpp"}

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
# host code: x86-64-v3 (AVX2, FMA; floor/ceil are single instructions instead
# of libm calls -- the per-join grid set-up of a batch runs them per dimension
# and job); -ffp-contract=off keeps every host operation the one written
HOST_ARCH = "-march=x86-64-v3"
# packed FP32 instructions are kept out of the device code
HIP_FLAGS = COMMON + [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                      "-packed-fp32-ops", "-munsafe-fp-atomics", "-Xarch_host", HOST_ARCH]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")


# MFMA accumulators in VGPRs (gfx950 reads them there directly), so the
# filter's threshold test needs no AGPR -> VGPR copies
MFMA_FLAGS = ["-mllvm", "-amdgpu-mfma-vgpr-form"]


def _compile(src: str) -> Path:
    out = BUILD / (Path(src).stem + ".o")
    path = str(CSRC / src)
    if src.endswith(".hip"):
        extra = MFMA_FLAGS if src == "hj_mfma.hip" else []
        cmd = [HIPCC, *HIP_FLAGS, *extra, "-c", path, "-o", str(out)]
    elif src in HIP_HOST_SOURCES:
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", path, "-o", str(out)]
    else:
        cmd = [CXX, *COMMON, HOST_ARCH, "-I/opt/rocm/include", "-c", path, "-o", str(out)]
    _run(cmd)
    return out


def build(verbose: bool = True) -> Path:
    """Compiles every source (always) and links the library and the tools."""
    BUILD.mkdir(exist_ok=True)
    BIN.mkdir(exist_ok=True)
    srcs = [s for s in HIP_SOURCES + HIP_HOST_SOURCES + CXX_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    for name, src in CLIS.items():
        _run([CXX, *COMMON, "-o", str(BIN / name), str(CSRC / src), "-L", str(PKG),
              "-lhilbertjoin_amd", "-Wl,-rpath,$ORIGIN/..", "-lpthread"])
    if verbose:
        print(f"built {LIB.relative_to(PKG.parent)} and {len(CLIS)} tools")
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build()
