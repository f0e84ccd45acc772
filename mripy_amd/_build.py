"""Build libgaussnufft.so in-tree for gfx950.

The HIP sources are compiled to objects with hipcc and linked with g++ against
the HIP runtime and rocFFT that PyTorch-ROCm bundles, so one process never maps
two HIP runtimes (torch's libamdhip64 carries the soname libamdhip64.so.7, which
the dynamic loader then reuses).  Outside a torch install the system ROCm
libraries are used instead.

Usage: ``python -m mripy_amd._build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgaussnufft.so")
SOURCES = ["gn_api.hip", "gn_kernels.hip", "gn_solver.hip", "gn_batch.hip", "gn_fft.hip", "gn_cfft.hip"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
    # no packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def runtime_libdir():
    """Directory of the HIP runtime the process will use (torch's bundle if present)."""
    spec = importlib.util.find_spec("torch")
    if spec is not None and spec.origin:
        d = os.path.join(os.path.dirname(spec.origin), "lib")
        if os.path.exists(os.path.join(d, "libamdhip64.so")) and os.path.exists(os.path.join(d, "librocfft.so")):
            return d, os.path.join(d, "libamdhip64.so"), os.path.join(d, "librocfft.so")
    d = os.path.join(ROCM, "lib")
    return d, os.path.join(d, "libamdhip64.so"), os.path.join(d, "librocfft.so")


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    libdir, hip, rocfft = runtime_libdir()
    tmp = LIB + ".tmp"
    cmd = ["g++", "-shared", "-o", tmp, *objs, hip, rocfft, "-Wl,-rpath," + libdir, "-Wl,--no-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
