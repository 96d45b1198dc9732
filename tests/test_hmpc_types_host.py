"""CPU check of the compile-time integer layer of include/hmpc_amd/hmpc.hpp (hmpc's spelling of the
path): _int literals, mod<q>'s limb count L = ceil(bit_width(q - 1) / 32), its constants R mod q,
R^2 mod q and -q^-1 mod 2^32, the Montgomery form of converted values (including a value wider than
q), integer_traits<mod>::one, the Montgomery product, and shapes, all against python integers."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "hmpc_types_check.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "hmpc_types_check")
HDRS = [os.path.join(ROOT, "include", "hmpc_amd", h) for h in ("hmpc.hpp", "comp.hpp")]

MODULI = {
    "q279": 0x591F5B834C0D961F67343BCC8902BDEDA2771F54306FF151162FF8D2B40F4194DC0001,
    "p64": 0x8822D80623320001,
    "q150": 0x2FAEADBE7A0195C011AC195AD10269830E8001,
    "p125": 0x18F25CD9A75CCBD9C146D4ABAEC00001,
    "gold": (1 << 64) - (1 << 32) + 1,
    "small": 12289,
}


@pytest.fixture(scope="module")
def output():
    if not os.path.exists(BIN) or any(os.path.getmtime(BIN) < os.path.getmtime(f) for f in [SRC] + HDRS):
        subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", f"-I{os.path.join(ROOT, 'include')}", SRC, "-o", BIN],
                       check=True)
    lines = subprocess.run([BIN], capture_output=True, text=True, check=True).stdout.splitlines()
    return {line.split()[0]: line.split()[1:] for line in lines}


@pytest.mark.parametrize("name", list(MODULI))
def test_modulus_constants_and_conversions(output, name):
    q = MODULI[name]
    f = output[name]
    L = (q - 1).bit_length() // 32 + (1 if (q - 1).bit_length() % 32 else 0)
    R = 1 << (32 * L)
    assert int(f[0]) == L and int(f[1]) == (q - 1).bit_length()
    got = [int(v, 16) for v in f[2:5]]
    assert got == [q, R % q, R * R % q]
    assert int(f[5]) == (-pow(q, -1, 1 << 32)) % (1 << 32)
    values = [0, 1, 2, 12345, 0xFFFFFFFF, 0xFEDCBA98_01234567_89ABCDEF]
    words = [int(v, 16) for v in f[6:]]
    assert words[:6] == [v % q * R % q for v in values]
    assert words[6] == R % q  # integer_traits<mod>::one
    assert words[7] == 15 * R % q  # (3 R)(5 R) R^-1


def test_literal_widths_and_shapes(output):
    assert [int(v) for v in output["widths"]] == [MODULI["q279"].bit_length(), 64, 64, 14, 4]
    assert [int(v) for v in output["shape"]] == [24, 4, 3, 1, 3]
