"""CPU check of the capability draw helpers of hmpc_amd/csrc/chacha.hpp built for the host.

The uniform ternary draw (width 0, not in hmpc) maps a 64-bit word w to floor(3 w / 2^64) - 1
with a 64-bit carry computation on the device; here it is compared with exact integer division at
every interval boundary k 2^64 / 3 and on random words, and the centred binomial widths keep
hmpc's popcount rule and word usage."""

import os
import random
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "sampler_check.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "sampler_check")
HDRS = [os.path.join(ROOT, "hmpc_amd", "csrc", h) for h in ("chacha.hpp", "modarith.hpp")]


@pytest.fixture(scope="module")
def tool():
    if not os.path.exists(BIN) or any(os.path.getmtime(BIN) < os.path.getmtime(f) for f in [SRC] + HDRS):
        subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", SRC, "-o", BIN],
                       check=True)
    p = subprocess.Popen([BIN], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
    yield p
    p.stdin.close()
    p.wait()


def draw(p, w, bits):
    p.stdin.write(f"{w & 0xFFFFFFFF} {w >> 32} {bits}\n")
    p.stdin.flush()
    v, two = (int(x) for x in p.stdout.readline().split())
    return v, two


def test_ternary_at_interval_boundaries(tool):
    words = [0, 1, (1 << 64) - 1, (1 << 63), (1 << 63) - 1]
    for k in (1, 2):
        edge = (k << 64) // 3  # floor(k 2^64 / 3): the last word below the boundary or the first on it
        words += [edge - 1, edge, edge + 1, edge + 2]
    for w in words:
        v, two = draw(tool, w, 0)
        assert v == (3 * w >> 64) - 1, hex(w)
        assert two == 1


def test_ternary_random_words_and_balance(tool):
    rng = random.Random(5)
    counts = {-1: 0, 0: 0, 1: 0}
    for _ in range(3000):
        w = rng.getrandbits(64)
        v, _ = draw(tool, w, 0)
        assert v == (3 * w >> 64) - 1
        counts[v] += 1
    assert all(900 < c < 1100 for c in counts.values()), counts


@pytest.mark.parametrize("bits", [2, 20, 32, 40, 64])
def test_centred_binomial_widths(tool, bits):
    rng = random.Random(bits)
    for _ in range(200):
        w = rng.getrandbits(64)
        v, two = draw(tool, w, bits)
        assert v == bin(w & ((1 << bits) - 1)).count("1") - bits // 2
        assert two == (1 if bits > 32 else 0)
