"""Build libnaga_amd.so in-tree with hipcc for gfx950.

    python -m naga_amd.build        (or naga_amd.build.build())

Every HIP/C++ source under csrc/ is compiled to an object file in parallel,
then linked into naga_amd/libnaga_amd.so.  All device code is compiled with
packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnaga_amd.so")
ARCH = os.environ.get("NAGA_OFFLOAD_ARCH", "gfx950")

SOURCES = ["step_kernels.hip", "setup_kernels.hip", "calculus.hip", "operators.hip", "engine.cpp",
           "server.cpp", "capi.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-result",
    "-I", os.path.join(HERE, "..", "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build naga_amd)")


def _compile(src: str, obj: str) -> str:
    cmd = [hipcc()] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile and link; returns the library path."""
    with tempfile.TemporaryDirectory(prefix="naga_build_") as tmp:
        objs = [os.path.join(tmp, s + ".o") for s in SOURCES]
        with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
            list(ex.map(_compile, SOURCES, objs))
        out_tmp = LIB + ".tmp"
        cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", out_tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
        os.replace(out_tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
