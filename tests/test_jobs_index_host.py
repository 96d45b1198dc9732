"""CPU check of the fast kernels' indexing (tests/cpp/jobs_index_check.cpp), with the kernels' own
geometry code (hmpc_amd/csrc/tile_geometry.hpp) compiled for the host:
- the persistent job kernel: every operand, key, message, twiddle, scratch and LDS index of both
  job types stays within the job's polynomial, its scratch slot and the tile, for N = 2^2..2^17, 2
  and 4 limbs and every elements-per-thread choice;
- the address decompositions every tile kernel relies on: the LDS swizzles are permutations and
  linear over disjoint bits, each unit's elements are element 0 | unit_offset(k), the per-round
  twiddle base gives the per-stage indices, and tile offsets are additive, for every tile geometry,
  round, unit order and elements-per-thread choice;
- the streaming column passes of the radix-2^26 product: scalar-loaded twiddle indices are the same
  on every lane of a wave (or lane 0's / lane 63's where two are taken), and their direct round-0
  loads and last-round stores move every tile element once, inside its polynomial, at N = 2^14..2^17."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "jobs_index_check.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "jobs_index_check")


def test_job_kernel_indices_stay_in_bounds():
    subprocess.run(["g++", "-O2", "-std=c++17", "-w", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", SRC, "-o", BIN],
                   check=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    # whole-tile jobs of N < 4096: the valid guard skips the reads that would leave the N-element planes
    guarded = int(r.stdout.split("guarded=")[1])
    assert guarded > 0
