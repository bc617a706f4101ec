"""Build libtriton_amd.so (HIP for gfx950 + C++ host) in-tree.

Every HIP translation unit is compiled for gfx950 with FMA contraction off
(the FP64 path reproduces the CPU build bit for bit) and packed-FP32 ops
disabled; FP32 kernels opt back into contraction with a pragma.  Objects go
to triton_amd/_build/, the library to triton_amd/_lib/libtriton_amd.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libtriton_amd.so"

ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["kernels_ref.hip", "step_f64.hip", "step_f32.hip", "batch.hip", "sim.hip"]
HOST_SOURCES = ["host/case.cpp", "host/terrain.cpp", "host/serve.cpp", "capi.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{ROOT / 'include'}",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def _obj(src: str) -> Path:
    return BUILD / (src.replace("/", "_") + ".o")


def build(verbose: bool = False) -> Path:
    """Compile all sources and link the shared library; returns its path."""
    if not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(_obj(s))])
    cxx = shutil.which("g++") or "g++"
    for s in HOST_SOURCES:
        jobs.append([cxx, *HOST_FLAGS, "-c", str(CSRC / s), "-o", str(_obj(s))])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        for cmd, fut in [(c, ex.submit(_run, c)) for c in jobs]:
            fut.result()
            if verbose:
                print(" ".join(cmd[:1] + cmd[-3:]), flush=True)
    objs = [str(_obj(s)) for s in HIP_SOURCES + HOST_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    """Build the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle_triton.so"


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_oracle())
    sys.exit(0)
