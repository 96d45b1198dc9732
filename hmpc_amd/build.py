"""Builds libhmpc_hip.so for gfx950 in-tree (hipcc, no JIT cache).

Usage: python -m hmpc_amd.build   (or hmpc_amd.build.build() from __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhmpc_hip.so")
SOURCES = ["kernels.hip", "kernels_fast.hip", "kernels_fast_l2.hip", "kernels_fast_l4.hip", "kernels_jobs_l2.hip",
           "kernels_jobs_l4.hip", "kernels_ext.hip", "api.hip"]
ARCH = os.environ.get("HMPC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    # packed FP32 instructions are disabled for all device code on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    if os.environ.get("HMPC_BUILD_TIMES"):
        print(f"  {src}: {time.time() - t0:.1f} s", flush=True)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    build_facade_test()
    if verbose:
        print(f"built {LIB}")
    return LIB


FACADE_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "facade_test.cpp")
FACADE_TEST_BIN = os.path.join(os.path.dirname(HERE), "tests", "cpp", "facade_test")


def build_facade_test() -> str:
    """The C++ program that drives include/hmpc_amd/comp.hpp (run by tests/test_gpu_facade.py)."""
    root = os.path.dirname(HERE)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{os.path.join(root, 'include')}", FACADE_TEST_SRC,
           f"-L{HERE}", "-lhmpc_hip", "-Wl,-rpath,$ORIGIN/../../hmpc_amd", "-o", FACADE_TEST_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("facade test build failed")
    return FACADE_TEST_BIN


if __name__ == "__main__":
    build(verbose=True)
