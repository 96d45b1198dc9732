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
    (a, b, c, t, B, s, e, A, words, c0, c1, words2, s0, s1, y, total, Mv, nonce_bgv, mh, e0, e1, nonce_elem, ru, rs, ri,
     rd, sh0, sh1, xp, xq, Ap, nonce_plain, p0, p1) = _read(out)
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
    # examples/bgv.cpp:80-90 as expressions with a shared capability u: the same ciphertexts as one
    # enc call on the recorded nonce (u, v, w on nonces n0, n0 + kN, n0 + 2kN)
    want_mh = om.ntt(oracle.ints_to_planar([om.from_unsigned(int(w_)) for w_ in words], L), log_n)
    assert np.array_equal(mh, want_mh.reshape(-1))
    x0, x1, _ = om.bgv_encrypt(log_n, oA, oB, words, t_mont, KEY, nonce_bgv.tolist(), threads=8)
    assert np.array_equal(e0, x0.reshape(-1)) and np.array_equal(e1, x1.reshape(-1))
    # capability samplers in the planned kernels: one nonce per kernel, in order
    nz = int(nonce_elem[0])
    cnt = n * N
    def cap(dist, bits, layout, k):
        return om.sample_capability(KEY, [nz + k * cnt, 0, 0], dist, bits, layout, log_n, n)
    assert np.array_equal(ru, cap("uniform", 0, "element", 0).reshape(-1))
    assert np.array_equal(rs, om.pointwise("add", cap("centered_binomial", 40, "element", 1), pa).reshape(-1))
    assert np.array_equal(ri, om.ntt(cap("centered_binomial", 40, "intt", 2), log_n, inverse=True).reshape(-1))
    assert np.array_equal(rd, om.ntt(cap("drown_signed", oracle.drown_bits(2, True), "ntt", 3), log_n).reshape(-1))
    shared = cap("centered_binomial", 2, "element", 4)
    assert np.array_equal(sh0, om.pointwise("add", shared, pa).reshape(-1))
    assert np.array_equal(sh1, om.pointwise("mul", shared, pb).reshape(-1))
    # cast to q and BGV with an NTT-form plaintext over the Goldilocks prime
    gold = oracle.Modulus((1 << 64) - (1 << 32) + 1)
    oxp = gold.sample_uniform(KEY, 30, k * N)
    assert np.array_equal(xp, oxp.reshape(-1))
    assert np.array_equal(xq, gold.cast_to(om, oxp).reshape(-1))
    pcol = oracle.ints_to_planar([om.from_unsigned((1 << 64) - (1 << 32) + 1)] * N, L)
    oAp = om.pointwise("add", om.pointwise("mul", oB, os_), om.pointwise("mul", oe, pcol))
    assert np.array_equal(Ap, oAp.reshape(-1))
    w0_, w1_, _ = om.bgv_encrypt_plaintext(gold, log_n, oAp, oB, gold.ntt(oxp, log_n), True, KEY, nonce_plain.tolist(),
                                           threads=8)
    assert np.array_equal(p0, w0_.reshape(-1)) and np.array_equal(p1, w1_.reshape(-1))
