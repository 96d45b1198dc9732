"""CPU check of the serving loop's host policy (hmpc_amd/csrc/server_pool.hpp): the output pool with
an allocator that fails on demand, and the launch admission limits. Built with g++ (no GPU)."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "server_check.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "server_check")


def test_server_pool_and_admission_policy():
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-fsanitize=address,undefined", SRC, "-o", BIN],
                   check=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
