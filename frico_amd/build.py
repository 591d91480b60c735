"""Builds libfrico_hip.so in-tree with hipcc for gfx950 (no JIT cache, no pip install)."""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfrico_hip.so")
SOURCES = ["fill.hip", "lu.hip", "adjoint.hip", "capi.cpp", "adjoint_host.cpp", "maxwell.cpp", "batch.cpp", "planes.cpp", "outputs.cpp", "msh.cpp", "noalloc.hip", "gemm_f16.hip", "blocksolve.hip", "h5write.cpp"]
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 instructions are disabled on purpose (see the build notes).
FLAGS = ["-O3", "-std=c++23", "-fPIC", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
         "-I", CSRC, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-command-line-argument"]


def _compile(src: str) -> str:
    out = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return out


CAPI_SRC = os.path.join(ROOT, "tests", "capi", "run_context.cpp")
CAPI_BIN = os.path.join(ROOT, "tests", "capi", "run_context")


def build_capi_caller(verbose: bool = False) -> str:
    """The compiled C++ caller of the C ABI (tests/capi/run_context.cpp), with g++ against
    include/frico_gpu.h and the in-tree libfrico_hip.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), CAPI_SRC, "-L", HERE,
           "-lfrico_hip", "-Wl,-rpath,$ORIGIN/../../frico_amd", "-o", CAPI_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed on tests/capi/run_context.cpp")
    if verbose:
        print("built", CAPI_BIN)
    return CAPI_BIN


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libfrico_hip.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_capi_caller(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
