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
# the job kernels (jobs_impl.hpp) are instantiated in one unit per launch family and limb count, so
# that the slowest units compile in parallel
SOURCES = ["kernels_jobs_l2.hip", "kernels_jobs_l2_e16.hip", "kernels_jobs_l2_bgv.hip", "kernels_jobs_l2_mul.hip", "kernels_jobs_l4.hip",
           "kernels_jobs_l4_bgv.hip", "kernels_jobs_l4_mul.hip", "kernels.hip", "kernels_fast.hip", "kernels_fast_l2.hip",
           "kernels_fast_l4.hip", "kernels_fast_c4.hip", "kernels_fast_c5.hip", "kernels_ext.hip", "kernels_lhe.hip", "api.hip"]
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


def _up_to_date(src: str, obj: str) -> bool:
    """True when obj is newer than the source and every header of csrc/ and include/ (a coarse rule
    for the incremental developer build; build() itself always recompiles)."""
    if not os.path.exists(obj):
        return False
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    deps.append(os.path.join(os.path.dirname(HERE), "include", "hmpc_hip.h"))
    return all(os.path.getmtime(d) < t for d in deps)


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if incremental and _up_to_date(src, obj):
        return obj
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    if os.environ.get("HMPC_BUILD_TIMES"):
        print(f"  {src}: {time.time() - t0:.1f} s", flush=True)
    return obj


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every unit (incremental=True skips units whose object is newer than all sources) and
    link hmpc_amd/libhmpc_hip.so."""
    os.makedirs(BUILD, exist_ok=True)
    workers = max(1, min(len(SOURCES), os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(lambda src: _compile(src, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    build_facade_test()
    build_hmpc_source_test()
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


HMPC_SOURCE_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "hmpc_source_test.cpp")
HMPC_SOURCE_TEST_BIN = os.path.join(os.path.dirname(HERE), "tests", "cpp", "hmpc_source_test")


def build_hmpc_source_test() -> str:
    """The C++20 program written in hmpc's spelling over include/hmpc_amd/hmpc.hpp (run by
    tests/test_gpu_hmpc_source.py)."""
    root = os.path.dirname(HERE)
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", f"-I{os.path.join(root, 'include')}", HMPC_SOURCE_TEST_SRC,
           f"-L{HERE}", "-lhmpc_hip", "-Wl,-rpath,$ORIGIN/../../hmpc_amd", "-o", HMPC_SOURCE_TEST_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hmpc source test build failed")
    return HMPC_SOURCE_TEST_BIN


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
