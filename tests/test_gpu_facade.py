"""GPU test of the C++ interface (include/hmpc_amd/comp.hpp): the facade program built by
hmpc_amd.build runs products, hmpc's NTT rewrites and BGV through the C ABI; its outputs are
recomputed here with the oracle, bit for bit."""

import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from bigref import Q128A

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_test")
KEY = [42, 0, 0, 0, 0, 0, 0, 0]


def _read(path):
    out = []
    with open(path, "rb") as f:
        while True:
            h = f.read(8)
            if not h:
                return out
            n = int(np.frombuffer(h, np.uint64)[0])
            out.append(np.frombuffer(f.read(4 * n), np.uint32))


def test_cpp_facade_matches_oracle(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/facade_test is missing: run hmpc_amd.build first")
    out = tmp_path / "facade.bin"
    print("facade program", flush=True)
    sys.stdout.flush()
    r = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    a, b, c, t, B, s, e, A, words, c0, c1, words2, s0, s1, y, total, Mv = _read(out)
    om = oracle.Modulus(Q128A)
    L, log_n, n, k = 4, 12, 3, 2
    N = 1 << log_n
    pa = om.sample_uniform(KEY, 0, n * N)
    pb = om.sample_uniform(KEY, 1, n * N)
    assert np.array_equal(a, pa.reshape(-1)) and np.array_equal(b, pb.reshape(-1))
    want_c, _ = om.polymul_timed(log_n, pa, pb, threads=8)
    assert np.array_equal(c, want_c.reshape(-1))
    t_mont = om.from_unsigned(1 << 32)
    assert np.array_equal(t, oracle.int_to_limbs(t_mont, L))
    oB = om.sample_uniform(KEY, 10, N)
    os_ = om.ntt(om.sample_cbd(KEY, 11, 2, N), log_n)
    oe = om.ntt(om.sample_cbd(KEY, 12, 40, N), log_n)
    tcol = oracle.ints_to_planar([t_mont] * N, L)
    oA = om.pointwise("add", om.pointwise("mul", oB, os_), om.pointwise("mul", oe, tcol))
    for got, want in ((B, oB), (s, os_), (e, oe), (A, oA)):
        assert np.array_equal(got, want.reshape(-1))
    want0, want1, nonce = om.bgv_encrypt(log_n, oA, oB, words, t_mont, KEY, [0, 0, 0], threads=8)
    assert np.array_equal(c0, want0.reshape(-1)) and np.array_equal(c1, want1.reshape(-1))
    assert words.size == k * N
    # ct + ct2 (ct2 on the next capability nonce)
    w0, w1, _ = om.bgv_encrypt(log_n, oA, oB, words2, t_mont, KEY, nonce, threads=8)
    assert np.array_equal(s0, om.pointwise("add", want0, w0).reshape(-1))
    assert np.array_equal(s1, om.pointwise("add", want1, w1).reshape(-1))
    # tensor cipher of a with key {42}, nonce {1}
    assert np.array_equal(y, oracle.cipher_encrypt(pa, KEY, [1, 0, 0]).reshape(-1))
    # sum of a, and M v for the 4 x 4096 matrix M
    q = Q128A
    assert oracle.planar_to_ints(total.reshape(L, 1))[0] == sum(oracle.planar_to_ints(pa)) % q
    M = oracle.planar_to_ints(om.sample_uniform(KEY, 20, 4 * N))
    v = oracle.planar_to_ints(om.sample_uniform(KEY, 21, N))
    Rinv = pow(1 << 128, -1, q)
    got = oracle.planar_to_ints(Mv.reshape(L, 4))
    for r in range(4):
        assert got[r] == sum(M[r * N + j] * v[j] for j in range(N)) * Rinv % q
