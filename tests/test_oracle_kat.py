"""Pins the CPU oracle against hmpc's own known-answer tests and against the mathematics.

Each check is chosen so that a plausible mistake in the oracle (wrong word order, a
dropped carry, a wrong bit-reversal, a transposed operand) fails it.
"""

import json
import os
import random

import numpy as np
import pytest

import oracle
from bigref import (P150, Q64, Q128A, Q128B, Stream, brv, capability_samples, centred, chacha20_block,
                    inverse_negacyclic_dft, is_probable_prime, limbs_of, negacyclic_dft, quarter_round, sample_mont,
                    schoolbook_negacyclic, stream_words, words_to_int)

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
# moduli of hmpc's linear-BGV test (tests/expr/crypto/lhe/enc.cpp:16-17) and of its BGV example
# (examples/bgv.cpp:25-29)
P64_ENC = int(VEC["lhe"]["enc_test"]["p"], 16)
Q279 = int(VEC["lhe"]["enc_test"]["q"], 16)
P64_EXAMPLE = int(VEC["lhe"]["bgv_example"]["p"], 16)


def _int(x):
    return int(x, 16) if isinstance(x, str) else int(x)


# ------------------------------------------------------------------ ChaCha20 (tests/crypto/chacha.cpp)
def test_quarter_round_rfc8439():
    qr = VEC["chacha20"]["quarter_round"]
    assert list(quarter_round(*qr["in"])) == qr["out"]


def test_block_function_rfc8439():
    b = VEC["chacha20"]["block"]
    out = oracle.chacha20_block(b["key"], b["counter"], b["nonce"])
    assert out.tolist() == b["block"]
    assert chacha20_block(b["key"], b["counter"], b["nonce"]) == b["block"]


@pytest.mark.parametrize("idx", range(5))
def test_additional_block_vectors(idx):
    v = VEC["chacha20"]["extra"][idx]
    assert oracle.chacha20_block(v["key"], v["counter"], v["nonce"]).tolist() == v["block"]


def test_stream_counter_advances_per_block():
    # hmpc/crypto/chacha.hpp:159-163: counter += 1 after each block
    v0, v1 = VEC["chacha20"]["extra"][0], VEC["chacha20"]["extra"][1]
    s = oracle.chacha20_stream(v0["key"], v0["nonce"], 0, 32)
    assert s[:16].tolist() == v0["block"] and s[16:].tolist() == v1["block"]


def test_chacha_random_keys_match_textbook():
    rng = random.Random(7)
    for _ in range(20):
        key = [rng.getrandbits(32) for _ in range(8)]
        nonce = [rng.getrandbits(32) for _ in range(3)]
        ctr = rng.getrandbits(32)
        assert oracle.chacha20_block(key, ctr, nonce).tolist() == chacha20_block(key, ctr, nonce)


# ------------------------------------------------------------------ Montgomery (tests/ints/mod.cpp)
def test_mod_constants_p150():
    v = VEC["mod"]
    p = _int(v["p"])
    m = oracle.Modulus(p)
    assert m.limbs == 5 and m.bit_size == 150
    nprime, r1, r2 = m.constants()
    assert r1 == _int(v["R_mod_p"]) == (1 << 160) % p
    assert r2 == (1 << 320) % p
    assert (p * nprime + 1) % (1 << 32) == 0
    assert m.from_unsigned(0) == 0
    assert m.from_unsigned(1) == r1
    assert m.from_unsigned(512) == _int(v["mont_512"])
    assert m.to_unsigned(_int(v["mont_512"])) == 512
    x = _int(v["mont_512"])
    for k in range(1, 6):
        assert m.to_unsigned(m.pow(x, k)) == 512 ** k
    assert m.to_unsigned(m.from_unsigned(p)) == 0
    assert m.to_unsigned(m.from_unsigned(p + 1)) == 1
    assert m.to_unsigned(m.from_unsigned(_int(v["big_unsigned"]))) == v["big_unsigned_value"]
    assert m.to_unsigned(m.from_small_signed(-1)) == _int(v["neg_one_value"])


@pytest.mark.parametrize("q", [Q64, Q128A, Q128B, P150, 1601, 0x8822D80623320001])
def test_montgomery_ops_closed_form(q):
    m = oracle.Modulus(q)
    R = 1 << (32 * m.limbs)
    Rinv = pow(R, -1, q)
    rng = random.Random(q & 0xFFFF)
    for _ in range(50):
        a, b = rng.randrange(q), rng.randrange(q)
        assert m.mul(a, b) == a * b * Rinv % q
        assert m.add(a, b) == (a + b) % q
        assert m.sub(a, b) == (a - b) % q
        assert m.from_unsigned(a) == a * R % q
        assert m.to_unsigned(a) == a * Rinv % q
        e = rng.getrandbits(70)
        assert m.pow(a, e) == pow(a * Rinv, e, q) * R % q
        if a:
            assert m.invert(a) == pow(a * Rinv, -1, q) * R % q
    for s in (-20, -1, 0, 1, 20):
        assert m.from_small_signed(s) == (s % q) * R % q
    # REDC with more rounds than limbs: v * 2^(-32 r) mod q (uniform sampling path)
    for r in (m.limbs, m.limbs + 3):
        v = rng.getrandbits(32 * r - 3)
        assert m.montgomery_reduce(v, r, r) == v * pow(1 << (32 * r), -1, q) % q


# ------------------------------------------------------------------ moduli of the workloads
def test_workload_moduli_are_the_stated_primes():
    assert is_probable_prime(Q64) and limbs_of(Q64) == 2
    for q, k, L in ((Q128A, 17, 4), (Q128B, 15, 4)):
        assert is_probable_prime(q) and q % (1 << k) == 1 and limbs_of(q) == L
        # smallest prime > 2^124 with q = 1 mod 2^k
        c = (1 << 124) + 1
        while c < q:
            assert not is_probable_prime(c), hex(c)
            c += 1 << k
    assert Q128B % (1 << 16) != 1  # cannot host N = 2^15 (needs 2N | q - 1)


# ------------------------------------------------------------------ roots (poly_mod.cpp, root_of_unity.cpp)
def python_root_search(q, log_degree):
    """Closed form of hmpc/ints/num/theory/root_of_unity.hpp:33-53 from the all-zero ChaCha stream."""
    L = limbs_of(q)
    bits = (q - 1).bit_length()
    R = 1 << (32 * L)
    words = stream_words([0] * 8, [0] * 3, 0, 64 * L)
    deg = 1 << log_degree
    for k in range(64):
        x = words_to_int(words[k * L:(k + 1) * L], bits)
        if x > q - 1:
            x -= q
        if x == 0:
            continue
        a = x * pow(R, -1, q) % q
        g = pow(a, (q - 1) // deg, q)
        if pow(g, deg // 2, q) != 1:
            return g * R % q, k
    raise AssertionError


def test_root_kat_p150():
    v = VEC["poly_mod"]
    p = _int(v["p"])
    m = oracle.Modulus(p)
    root = m.root_of_unity(v["log_n"] + 1)
    iroot = m.root_of_unity(v["log_n"] + 1, inverse=True)
    assert m.to_unsigned(root) == _int(v["root"])
    assert m.to_unsigned(iroot) == _int(v["inverse_root"])
    assert python_root_search(p, v["log_n"] + 1)[0] == root


def test_root_of_unity_1601():
    v = VEC["root_of_unity"]
    m = oracle.Modulus(v["p"])
    n = v["n"]
    root = m.root_of_unity(n.bit_length() - 1)
    one = m.from_unsigned(1)
    for i in range(1, n):
        assert m.pow(root, i) != one
    assert m.pow(root, n) == one
    assert m.pow(root, 0) == one


@pytest.mark.parametrize("q,log_n", [(Q64, 12), (Q64, 16), (Q128A, 16), (Q128A, 12), (Q128B, 14)])
def test_root_closed_form_workload_moduli(q, log_n):
    m = oracle.Modulus(q)
    assert m.root_of_unity(log_n + 1) == python_root_search(q, log_n + 1)[0]


def test_roots_table_layout():
    q, log_n = Q64, 5
    m = oracle.Modulus(q)
    R = 1 << (32 * m.limbs)
    N = 1 << log_n
    phi = m.root_of_unity(log_n + 1) * pow(R, -1, q) % q
    fwd = oracle.planar_to_ints(m.roots(log_n))
    inv = oracle.planar_to_ints(m.roots(log_n, inverse=True))
    ninv = pow(N, -1, q)
    assert fwd[0] == R % q
    assert inv[0] == ninv * R % q
    for i in range(1, N):
        j = brv(i, log_n)
        assert fwd[j] == pow(phi, i, q) * R % q
        expect = pow(phi, -i, q) * (ninv if j == 1 else 1) % q
        assert inv[j] == expect * R % q


# ------------------------------------------------------------------ NTT (poly_mod.cpp:92-103 + math)
def test_ntt_kat_p150():
    v = VEC["poly_mod"]
    p = _int(v["p"])
    m = oracle.Modulus(p)
    log_n = v["log_n"]
    N = 1 << log_n
    R = 1 << (32 * m.limbs)
    x = oracle.ints_to_planar([(i % N) * R % p for i in range(2 * N)], m.limbs)
    y = m.ntt(x, log_n, threads=4)
    vals = oracle.planar_to_ints(y)
    Rinv = pow(R, -1, p)
    for j in range(2):
        for k, want in v["ntt_of_iota"].items():
            assert vals[j * N + int(k)] * Rinv % p == _int(want)
    z = m.ntt(y, log_n, inverse=True, threads=4)
    assert np.array_equal(z, x)


def test_ntt_of_zero_is_zero():
    m = oracle.Modulus(P150)
    x = np.zeros((m.limbs, 2 << 10), np.uint32)
    assert not m.ntt(x, 10).any() and not m.ntt(x, 10, inverse=True).any()


@pytest.mark.parametrize("q", [Q64, Q128A, Q128B])
@pytest.mark.parametrize("log_n", [2, 3, 5])
def test_ntt_closed_form_dft(q, log_n):
    m = oracle.Modulus(q)
    R = 1 << (32 * m.limbs)
    Rinv = pow(R, -1, q)
    N = 1 << log_n
    rng = random.Random(log_n)
    a = [rng.randrange(q) for _ in range(3 * N)]
    x = oracle.ints_to_planar([v * R % q for v in a], m.limbs)
    y = oracle.planar_to_ints(m.ntt(x, log_n))
    phi = m.root_of_unity(log_n + 1) * Rinv % q
    for b in range(3):
        want = negacyclic_dft(a[b * N:(b + 1) * N], phi, q)
        assert [v * Rinv % q for v in y[b * N:(b + 1) * N]] == want


@pytest.mark.parametrize("q,log_n", [(Q64, 6), (Q128A, 6), (Q128B, 7), (P150, 5)])
def test_convolution_theorem_vs_schoolbook(q, log_n):
    m = oracle.Modulus(q)
    R = 1 << (32 * m.limbs)
    Rinv = pow(R, -1, q)
    N = 1 << log_n
    rng = random.Random(q % 1000 + log_n)
    a = [rng.randrange(q) for _ in range(N)]
    b = [rng.randrange(q) for _ in range(N)]
    xa = oracle.ints_to_planar([v * R % q for v in a], m.limbs)
    xb = oracle.ints_to_planar([v * R % q for v in b], m.limbs)
    prod = m.ntt(m.pointwise("mul", m.ntt(xa, log_n), m.ntt(xb, log_n)), log_n, inverse=True)
    want = schoolbook_negacyclic(a, b, q)
    assert [v * Rinv % q for v in oracle.planar_to_ints(prod)] == want
    assert np.array_equal(m.schoolbook_negacyclic(log_n, xa, xb), prod)


def test_c1_roundtrip_and_product_n4096():
    """Config: N=4096, q64, coefficients from the explicit-key stream (key 42)."""
    m = oracle.Modulus(Q64)
    log_n = 12
    key = [42]
    x = m.sample_uniform(key, 0, 4096)
    y = m.sample_uniform(key, 1, 4096)
    fx = m.ntt(x, log_n, threads=4)
    assert np.array_equal(m.ntt(fx, log_n, inverse=True, threads=4), x)
    prod = m.ntt(m.pointwise("mul", fx, m.ntt(y, log_n)), log_n, inverse=True)
    assert np.array_equal(prod, m.schoolbook_negacyclic(log_n, x, y))


# ------------------------------------------------------------------ samplers
@pytest.mark.parametrize("q", [Q64, Q128A])
def test_uniform_explicit_key_closed_form(q):
    m = oracle.Modulus(q)
    key = [42, 1, 2, 3, 4, 5, 6, 7]
    count, offset = 37, 3
    out = oracle.planar_to_ints(m.sample_uniform(key, offset, count))
    bits = m.bit_size + 80
    W = (bits + 31) // 32
    r = max(m.limbs, W)
    for i in range(count):
        nonce_v = i + count * offset
        words = chacha20_block(key, 0, [nonce_v & 0xFFFFFFFF, nonce_v >> 32, 0])[:W]
        v = words_to_int(words, bits)
        assert out[i] == v * pow(1 << (32 * r), -1, q) % q


def test_uniform_is_statistically_uniform():
    m = oracle.Modulus(Q64)
    vals = np.array([v >> 60 for v in oracle.planar_to_ints(m.sample_uniform([9], 0, 1 << 14))])
    counts = np.bincount(vals, minlength=16)
    assert counts.min() > 850 and counts.max() < 1200


@pytest.mark.parametrize("bits", [2, 20, 40])
def test_cbd_explicit_key_closed_form(bits):
    q = Q128A
    m = oracle.Modulus(q)
    R = 1 << (32 * m.limbs)
    key = [42]
    count = 64
    out = oracle.planar_to_ints(m.sample_cbd(key, 2, bits, count))
    W = (bits + 31) // 32
    for i in range(count):
        nv = i + count * 2
        words = chacha20_block(key + [0] * 7, 0, [nv & 0xFFFFFFFF, nv >> 32, 0])[:W]
        val = bin(words_to_int(words, bits)).count("1") - bits // 2
        assert abs(val) <= bits // 2
        assert out[i] == (val % q) * R % q


def test_cbd_capability_layout():
    """Stage-0 fused sampling: one stream per butterfly, counter = i*N + tid, two draws."""
    q = Q64
    m = oracle.Modulus(q)
    R = 1 << (32 * m.limbs)
    key = [42, 0, 0, 0, 0, 0, 0, 1]
    nonce = [5, 0, 0]
    log_n, n, bits = 4, 3, 40
    N = 1 << log_n
    out = oracle.planar_to_ints(m.sample_cbd_capability(key, nonce, bits, log_n, n))
    for i in range(n):
        for tid in range(N // 2):
            words = stream_words(key, nonce, i * N + tid, 4)
            lo = bin(words_to_int(words[0:2], 40)).count("1") - 20
            hi = bin(words_to_int(words[2:4], 40)).count("1") - 20
            assert out[i * N + tid] == (lo % q) * R % q
            assert out[i * N + tid + N // 2] == (hi % q) * R % q


# ------------------------------------------------------------------ BGV
def _keygen(m, log_n, key, t_mont):
    """Project-defined key generation: B uniform, s and e centred binomial (variance 1/2, 10)."""
    N = 1 << log_n
    B = m.sample_uniform(key, 0, N)
    s = m.ntt(m.sample_cbd(key, 1, 2, N), log_n)
    e = m.ntt(m.sample_cbd(key, 2, 40, N), log_n)
    tcol = oracle.ints_to_planar([t_mont] * N, m.limbs)
    A = m.pointwise("add", m.pointwise("mul", B, s), m.pointwise("mul", e, tcol))
    return A, B, s


@pytest.mark.parametrize("q,log_n", [(Q128B, 10), (Q64, 8)])
def test_bgv_encrypt_decrypt_roundtrip(q, log_n):
    m = oracle.Modulus(q)
    R = 1 << (32 * m.limbs)
    key = [42]
    t_mont = m.from_unsigned(1 << 32) if q > (1 << 64) else m.from_unsigned(2)
    A, B, s = _keygen(m, log_n, key, t_mont)
    n = 3
    N = 1 << log_n
    if q > (1 << 64):
        msgs = oracle.sample_words(key, 3, n * N)
    else:
        msgs = (oracle.sample_words(key, 3, n * N) & 1).astype(np.uint32)
    enc_key = [42, 0, 0, 0, 0, 0, 0, 1]
    c0, c1, nonce = m.bgv_encrypt(log_n, A, B, msgs, t_mont, enc_key, [0, 0, 0], threads=4)
    assert nonce.tolist() == [3 * n * N, 0, 0]
    dec = m.bgv_decrypt(log_n, s, c0, c1)
    if q > (1 << 64):
        assert np.array_equal(dec, msgs)
    else:
        assert np.array_equal(dec & 1, msgs)
    # c1 = B*u + t*w recomputed from the sampled randomness
    u = m.ntt(m.sample_cbd_capability(enc_key, [0, 0, 0], 2, log_n, n), log_n)
    w = m.ntt(m.sample_cbd_capability(enc_key, [2 * n * N, 0, 0], 40, log_n, n), log_n)
    Bb = np.tile(B, (1, n))
    tcol = oracle.ints_to_planar([t_mont] * (n * N), m.limbs)
    want = m.pointwise("add", m.pointwise("mul", Bb, u), m.pointwise("mul", w, tcol))
    assert np.array_equal(c1, want)
    # c0 = A*u + t*v + NTT(m) with v drawn independently at nonce nN
    v = m.ntt(m.sample_cbd_capability(enc_key, [n * N, 0, 0], 40, log_n, n), log_n)
    msg = m.ntt(oracle.ints_to_planar([m.from_unsigned(int(w_)) for w_ in msgs], m.limbs), log_n)
    Ab = np.tile(A, (1, n))
    want0 = m.pointwise("add", m.pointwise("add", m.pointwise("mul", Ab, u), m.pointwise("mul", v, tcol)), msg)
    assert np.array_equal(c0, want0)


@pytest.mark.parametrize("bits", [(2, 40, 40), (0, 20, 20)])
@pytest.mark.parametrize("q", [Q128B, Q64])
def test_bgv_u32_messages_closed_form(q, bits):
    """The 32-bit-message encryption (examples/bgv.cpp:66-69 lift, t = 2^32 as a mod q constant) against
    python big ints: c0 = A U + t V + DFT(m), c1 = B U + t W with U, V, W from the textbook ChaCha
    words in the ntt layout (nonces 0, nN, 2nN), with hmpc's randomness (2, 40, 40 bits) and with a
    uniform ternary u and 20-bit (eta = 10) v, w."""
    m = oracle.Modulus(q)
    R = 1 << (32 * m.limbs)
    log_n, n = 3, 2
    N = 1 << log_n
    rng = random.Random(11)
    A = [rng.randrange(q) for _ in range(N)]
    B = [rng.randrange(q) for _ in range(N)]
    msgs = np.array([rng.getrandbits(32) for _ in range(n * N)], dtype=np.uint32)
    t = (1 << 32) % q
    key = [42, 0, 0, 0, 0, 0, 0, 1]
    c0, c1, _ = m.bgv_encrypt(log_n, oracle.ints_to_planar([a * R % q for a in A], m.limbs),
                              oracle.ints_to_planar([b * R % q for b in B], m.limbs), msgs, t * R % q, key, [0, 0, 0],
                              bits=bits)
    phi = python_root_search(q, log_n + 1)[0] * pow(R, -1, q) % q
    def dft_samples(k, bits):
        mont = capability_samples(key, [k * n * N, 0, 0], "centered_binomial", bits, "ntt", log_n, n, q)
        plain = [v_ * pow(R, -1, q) % q for v_ in mont]
        return [negacyclic_dft(plain[b * N:(b + 1) * N], phi, q) for b in range(n)]
    U, V, W = dft_samples(0, bits[0]), dft_samples(1, bits[1]), dft_samples(2, bits[2])
    Rinv = pow(R, -1, q)
    g0 = [v_ * Rinv % q for v_ in oracle.planar_to_ints(c0)]
    g1 = [v_ * Rinv % q for v_ in oracle.planar_to_ints(c1)]
    for b in range(n):
        M = negacyclic_dft([int(w_) % q for w_ in msgs[b * N:(b + 1) * N]], phi, q)
        for j in range(N):
            assert g0[b * N + j] == (A[j] * U[b][j] + t * V[b][j] + M[j]) % q
            assert g1[b * N + j] == (B[j] * U[b][j] + t * W[b][j]) % q


# ------------------------------------------------------------------ signed and cross-modulus constructors
def test_signed_constructor_kats():
    """tests/ints/mod.cpp:50-56: mod_p(-1) and mod_p(-(p * 2^128 + 42)) (a signed value wider than p,
    the 'larger signed' branch of hmpc/ints/mod.hpp:228-253)."""
    v = VEC["mod"]
    p = _int(v["p"])
    m = oracle.Modulus(p)
    big = _int(v["big_unsigned"])
    bits = big.bit_length() + 1
    assert m.to_unsigned(m.from_signed_bigint(-big, bits)) == _int(v["neg_big_unsigned_value"])
    assert m.to_unsigned(m.from_signed_bigint(-1, 2)) == _int(v["neg_one_value"])
    assert m.to_unsigned(m.from_signed_bigint(big, bits)) == v["big_unsigned_value"]


@pytest.mark.parametrize("q", [Q64, Q128A, P150, Q279])
@pytest.mark.parametrize("bits", [2, 7, 64, 65, 120, 150, 151, 200, 231])
def test_signed_bigint_closed_form(q, bits):
    """mod(sbigint<bits>) = (value mod q) R mod q for both branches (bits < and >= bit_size)."""
    m = oracle.Modulus(q)
    R = 1 << (32 * m.limbs)
    rng = random.Random(bits * 7 + q % 97)
    lo, hi = -(1 << (bits - 1)), (1 << (bits - 1)) - 1
    for value in [lo, hi, -1, 0, 1] + [rng.randint(lo, hi) for _ in range(20)]:
        assert m.from_signed_bigint(value, bits) == value % q * R % q, (bits, value)


def test_cross_modulus_kats():
    """tests/ints/mod.cpp:60-78: q = 99 against the 150-bit p in both directions, n = 0..49, with
    the negative values (-n) lifted centred."""
    v = VEC["mod"]
    p = _int(v["p"])
    c = v["cross"]
    q = c["q"]
    mp, mq = oracle.Modulus(p), oracle.Modulus(q)
    ns = list(range(c["n_max"] + 1))
    xq = oracle.ints_to_planar([mq.from_unsigned(n) for n in ns], mq.limbs)
    xp = oracle.ints_to_planar([mp.from_unsigned(n) for n in ns], mp.limbs)
    nq = oracle.ints_to_planar([mq.from_small_signed(-n) for n in ns], mq.limbs)
    np_ = oracle.ints_to_planar([mp.from_small_signed(-n) for n in ns], mp.limbs)
    got_pq = [mp.to_unsigned(x) for x in oracle.planar_to_ints(mq.cast_to(mp, xq))]
    got_qp = [mq.to_unsigned(x) for x in oracle.planar_to_ints(mp.cast_to(mq, xp))]
    got_pnq = [mp.to_unsigned(x) for x in oracle.planar_to_ints(mq.cast_to(mp, nq))]
    got_qnp = [mq.to_unsigned(x) for x in oracle.planar_to_ints(mp.cast_to(mq, np_))]
    assert got_pq == ns and got_qp == ns
    assert got_pnq == [p - n if n else 0 for n in ns]
    assert got_qnp == [q - n if n else 0 for n in ns]


CAST_PAIRS = [(P150, Q128A), (Q128A, P150), (Q279, P64_ENC), (P64_ENC, Q279), (Q64, Q128B), (Q128B, Q64),
              (P150, P64_EXAMPLE), (99, P150)]


@pytest.mark.parametrize("src,dst", CAST_PAIRS)
def test_cast_closed_form(src, dst):
    """mod<Q>(mod<P> x) = (centred(v) mod Q) R_Q mod Q with v = x R_P^-1 mod P, at the edges of the
    centring (0, (P-1)/2, (P+1)/2, P-1) and random values."""
    ms, md = oracle.Modulus(src), oracle.Modulus(dst)
    Rs, Rd = 1 << (32 * ms.limbs), 1 << (32 * md.limbs)
    rng = random.Random(src % 1009 + dst % 1013)
    vals = [0, 1, (src - 1) // 2, (src + 1) // 2, src - 1] + [rng.randrange(src) for _ in range(40)]
    x = oracle.ints_to_planar([v * Rs % src for v in vals], ms.limbs)
    got = oracle.planar_to_ints(ms.cast_to(md, x))
    assert got == [centred(v, src) % dst * Rd % dst for v in vals]


# ------------------------------------------------------------------ capability samplers, every layout
CAP_CASES = [("uniform", 0), ("centered_binomial", 2), ("centered_binomial", 40), ("centered_binomial", 0), ("binomial", 20),
             ("drown_signed", oracle.drown_bits(1 << 64, True)), ("drown_unsigned", oracle.drown_bits(12345, False)),
             ("drown_signed", 300)]


@pytest.mark.parametrize("q", [Q64, Q128A, P150, Q279])
@pytest.mark.parametrize("layout", ["element", "ntt", "intt"])
@pytest.mark.parametrize("dist,bits", CAP_CASES)
def test_capability_samplers_closed_form(q, layout, dist, bits):
    """Every (distribution, layout) of the capability path recomputed from the textbook ChaCha20 words
    (hmpc/comp/queue.hpp:185-218, :400-409; hmpc/expr/number_theoretic_transform.hpp:320-329, 480-489)."""
    m = oracle.Modulus(q)
    key, nonce = [42, 0, 0, 0, 0, 0, 0, 1], [3, 1, 0]
    log_n, n = 3, 2
    got = oracle.planar_to_ints(m.sample_capability(key, nonce, dist, bits, layout, log_n, n))
    assert got == capability_samples(key, nonce, dist, bits, layout, log_n, n, q)


def test_capability_ntt_layout_equals_the_fused_cbd_path():
    m = oracle.Modulus(Q128A)
    key, nonce = [42, 0, 0, 0, 0, 0, 0, 1], [9, 0, 0]
    a = m.sample_capability(key, nonce, "centered_binomial", 40, "ntt", 5, 3)
    assert np.array_equal(a, m.sample_cbd_capability(key, nonce, 40, 5, 3))


def test_drown_bits_closed_form():
    # bit_width((bound - 1) << 80) + signed (hmpc/random/uniform.hpp:34-47)
    assert oracle.drown_bits(1 << 64, True) == 64 + 80 + 1
    assert oracle.drown_bits(1 << 64, False) == 64 + 80
    assert oracle.drown_bits(2, True) == 82
    assert oracle.drown_bits(5, False, security=40) == 43


# ------------------------------------------------------------------ BGV with an Rp plaintext
def _python_bgv(q, p, log_n, A, B, x_mont_p, plaintext_ntt, key, nonce, bits_u, v_dist, v_bits, bits_w):
    """hmpc/expr/crypto/lhe/enc.hpp:19-69 in python big ints from the mathematics: plaintext values
    (Montgomery mod p) -> coefficients mod p (inverse negacyclic DFT for an NTT-form plaintext) ->
    centred lift to q -> DFT; u, v, w from the textbook ChaCha words (ntt layout, nonces nonce,
    nonce + nN, nonce + 2nN) -> DFT; c0 = A u + p v + M, c1 = B u + p w. Plain values in, plain out."""
    N = 1 << log_n
    n = len(x_mont_p) // N
    Lq, Lp = limbs_of(q), limbs_of(p)
    Rq, Rp = 1 << (32 * Lq), 1 << (32 * Lp)
    phi_q = python_root_search(q, log_n + 1)[0] * pow(Rq, -1, q) % q
    xs = [v * pow(Rp, -1, p) % p for v in x_mont_p]
    nonce_int = nonce[0] | nonce[1] << 32 | nonce[2] << 64
    def nonce_words(k):
        v = nonce_int + k * n * N
        return [v & 0xFFFFFFFF, (v >> 32) & 0xFFFFFFFF, (v >> 64) & 0xFFFFFFFF]
    def ntt_samples(k, dist, bits):
        mont = capability_samples(key, nonce_words(k), dist, bits, "ntt", log_n, n, q)
        plain = [v * pow(Rq, -1, q) % q for v in mont]
        return [negacyclic_dft(plain[b * N:(b + 1) * N], phi_q, q) for b in range(n)]
    U = ntt_samples(0, "centered_binomial", bits_u)
    V = ntt_samples(1, v_dist, v_bits)
    W = ntt_samples(2, "centered_binomial", bits_w)
    c0, c1 = [], []
    for b in range(n):
        coeffs = xs[b * N:(b + 1) * N]
        if plaintext_ntt:
            phi_p = python_root_search(p, log_n + 1)[0] * pow(Rp, -1, p) % p
            coeffs = inverse_negacyclic_dft(coeffs, phi_p, p)
        M = negacyclic_dft([centred(c, p) % q for c in coeffs], phi_q, q)
        for j in range(N):
            c0.append((A[j] * U[b][j] + p * V[b][j] + M[j]) % q)
            c1.append((B[j] * U[b][j] + p * W[b][j]) % q)
    return c0, c1


@pytest.mark.parametrize("q,p,plaintext_ntt,v_dist,v_bits", [
    (P150, P64_EXAMPLE, True, "centered_binomial", 40),
    (P150, P64_EXAMPLE, False, "centered_binomial", 40),
    (Q279, P64_ENC, True, "centered_binomial", 40),
    (Q128A, (1 << 64) - (1 << 32) + 1, True, "drown_signed", oracle.drown_bits(1 << 20, True)),
])
def test_bgv_plaintext_closed_form(q, p, plaintext_ntt, v_dist, v_bits):
    """The oracle's Rp-plaintext encryption against the python big-int statement: pins the message path
    (INTT_p, centred lift, NTT_q), the p v term, the capability layout and the nonce order."""
    mq, mp = oracle.Modulus(q), oracle.Modulus(p)
    log_n, n = 3, 2
    N = 1 << log_n
    Rq = 1 << (32 * mq.limbs)
    rng = random.Random(q % 997)
    A = [rng.randrange(q) for _ in range(N)]
    B = [rng.randrange(q) for _ in range(N)]
    x = [rng.randrange(p) for _ in range(n * N)]
    key, nonce = [42, 0, 0, 0, 0, 0, 0, 1], [7, 0, 0]
    pa = oracle.ints_to_planar([a * Rq % q for a in A], mq.limbs)
    pb = oracle.ints_to_planar([b * Rq % q for b in B], mq.limbs)
    px = oracle.ints_to_planar(x, mp.limbs)
    c0, c1, nz = mq.bgv_encrypt_plaintext(mp, log_n, pa, pb, px, plaintext_ntt, key, nonce, 2, v_dist, v_bits, 40)
    assert nz.tolist() == [7 + 3 * n * N, 0, 0]
    w0, w1 = _python_bgv(q, p, log_n, A, B, x, plaintext_ntt, key, nonce, 2, v_dist, v_bits, 40)
    Rinv = pow(Rq, -1, q)
    assert [v * Rinv % q for v in oracle.planar_to_ints(c0)] == w0
    assert [v * Rinv % q for v in oracle.planar_to_ints(c1)] == w1


def _iota_planar(count, limbs):
    out = np.zeros((limbs, count), np.uint32)
    out[0] = np.arange(count, dtype=np.uint32)
    return out


def test_reference_linear_bgv_case():
    """hmpc's own test, tests/expr/crypto/lhe/enc.cpp:15-72, at its size: q (279 bits, 9 limbs),
    p = 0x8822d80623320001, N = 2^16, shape {4, 2}, key a = b = raw words i, s = one, NTT-form
    plaintext x = raw words i; dec<plaintext>(s, enc(key, x, randomness)) == x."""
    t = VEC["lhe"]["enc_test"]
    q, p, log_n = int(t["q"], 16), int(t["p"], 16), t["log_n"]
    n = t["shape"][0] * t["shape"][1]
    N = 1 << log_n
    mq, mp = oracle.Modulus(q), oracle.Modulus(p)
    assert (mq.limbs, mp.limbs) == (9, 2)
    a = _iota_planar(N, mq.limbs)
    s = oracle.ints_to_planar([mq.from_unsigned(1)] * N, mq.limbs)
    x = _iota_planar(n * N, mp.limbs)
    c0, c1, nz = mq.bgv_encrypt_plaintext(mp, log_n, a, a, x, True, [1, 2, 3, 4, 5, 6, 7, 8], [0, 0, 0], threads=8)
    assert nz.tolist() == [3 * n * N, 0, 0]
    y = mq.bgv_decrypt_plaintext(mp, log_n, s, c0, c1, True, threads=8)
    assert np.array_equal(y, x)


@pytest.mark.parametrize("plaintext_ntt", [True, False])
def test_bgv_example_modulus_roundtrip(plaintext_ntt):
    """examples/bgv.cpp:25-29: q = the 150-bit prime, N = 2^14, plaintext modulus p =
    0x89d046e0ae640001; keys A = B s + p e (project keygen); decryption returns the plaintext."""
    e = VEC["lhe"]["bgv_example"]
    q, p, log_n = int(e["q"], 16), int(e["p"], 16), e["log_n"]
    mq, mp = oracle.Modulus(q), oracle.Modulus(p)
    N = 1 << log_n
    n = 2
    key = [42]
    p_mont = mq.from_unsigned(p)
    A, B, s = _keygen(mq, log_n, key, p_mont)
    x = mp.sample_uniform(key, 5, n * N)
    c0, c1, _ = mq.bgv_encrypt_plaintext(mp, log_n, A, B, x, plaintext_ntt, [42, 0, 0, 0, 0, 0, 0, 1], [0, 0, 0],
                                         threads=8)
    assert np.array_equal(mq.bgv_decrypt_plaintext(mp, log_n, s, c0, c1, plaintext_ntt, threads=8), x)


# ---------------------------------------------------------------- tensor cipher
def _words_of(byte_list):
    b = bytes(byte_list)
    return np.frombuffer(b[: len(b) // 4 * 4], dtype="<u4").astype(np.uint32)


def test_cipher_rfc8439_encryption_vector():
    """RFC 8439 section 2.4.2 (tests/crypto/cipher.cpp:8-36): the first 112 bytes = 28 words,
    as 7 elements of 4 limbs (element-major words are the byte stream), counter 1."""
    v = VEC["cipher"]["rfc8439_2_4_2"]
    plain, want = _words_of(v["plaintext"]), _words_of(v["ciphertext"])
    planar = plain.reshape(7, 4).T.copy()
    got = oracle.cipher_encrypt(planar, v["key"], v["nonce"], v["counter"])
    assert np.array_equal(got.reshape(-1), want)
    assert np.array_equal(oracle.cipher_decrypt(got, v["key"], v["nonce"], v["counter"]), planar)
    # 2 limbs per element: the same byte stream, another batch structure (lcm(16, 2) = 16)
    got2 = oracle.cipher_encrypt(plain.reshape(14, 2).T.copy(), v["key"], v["nonce"], v["counter"])
    assert np.array_equal(got2.reshape(-1), want)


@pytest.mark.parametrize("limbs", [1, 3, 4, 5, 7])
def test_cipher_is_one_keystream_across_batches(limbs):
    """hmpc's per-batch generators (lcm(16, L) words each) add up to one continuous keystream:
    word w of the element-major output is input ^ stream[w], for every limb count."""
    rng = np.random.default_rng(limbs)
    count = 37
    planar = rng.integers(0, 1 << 32, size=(limbs, count), dtype=np.uint64).astype(np.uint32)
    key, nonce = [7, 1, 2, 3, 4, 5, 6, 8], [9, 10, 11]
    got = oracle.cipher_encrypt(planar, key, nonce)
    stream = oracle.chacha20_stream(key, nonce, 0, count * limbs)
    assert np.array_equal(got.reshape(-1), planar.T.reshape(-1) ^ stream)


def test_cipher_tensor_roundtrip_reference_inputs():
    """tests/expr/crypto/cipher.cpp: ubigint<127> x_i = i over shape {4, 2}, key {42}, nonce {1}:
    decryption returns the inputs and the ciphertext has shape {4, 2, 4}."""
    t = VEC["cipher"]["tensor"]
    count = t["shape"][0] * t["shape"][1]
    planar = np.zeros((t["limbs"], count), np.uint32)
    planar[0] = np.arange(count, dtype=np.uint32)
    ct = oracle.cipher_encrypt(planar, t["key"], t["nonce"])
    assert ct.shape == (count, t["limbs"])
    assert not np.array_equal(ct.T, planar)
    assert np.array_equal(oracle.cipher_decrypt(ct, t["key"], t["nonce"]), planar)


def test_ciphertext_ops_under_decryption_oracle():
    """hmpc's ciphertext operator+ / operator- / operator* (expr/crypto/lhe/ciphertext.hpp:48-81)
    restated with the oracle's point-wise ops: decryption gives m_a + m_b, m_a - m_b and k m_a
    mod t = 2^32."""
    om = oracle.Modulus(Q128A)
    log_n, N = 10, 1 << 10
    t_mont = om.from_unsigned(1 << 32)
    key = [42, 0, 0, 0, 0, 0, 0, 0]
    B = om.sample_uniform(key, 10, N)
    s = om.ntt(om.sample_cbd(key, 11, 2, N), log_n)
    e = om.ntt(om.sample_cbd(key, 12, 40, N), log_n)
    A = om.pointwise("add", om.pointwise("mul", B, s), om.pointwise("mul", e, oracle.ints_to_planar([t_mont] * N, om.limbs)))
    ma, mb = oracle.sample_words(key, 13, N), oracle.sample_words(key, 14, N)
    enc_key = [42, 0, 0, 0, 0, 0, 0, 1]
    a0, a1, nonce = om.bgv_encrypt(log_n, A, B, ma, t_mont, enc_key, [0, 0, 0])
    b0, b1, _ = om.bgv_encrypt(log_n, A, B, mb, t_mont, enc_key, nonce)
    m64a, m64b = ma.astype(np.int64), mb.astype(np.int64)
    add = om.bgv_decrypt(log_n, s, om.pointwise("add", a0, b0), om.pointwise("add", a1, b1))
    assert np.array_equal(add, ((m64a + m64b) % (1 << 32)).astype(np.uint32))
    sub = om.bgv_decrypt(log_n, s, om.pointwise("sub", a0, b0), om.pointwise("sub", a1, b1))
    assert np.array_equal(sub, ((m64a - m64b) % (1 << 32)).astype(np.uint32))
    k = oracle.ints_to_planar([om.from_unsigned(7)], om.limbs)
    mul = om.bgv_decrypt(log_n, s, om.pointwise("mul", a0, k), om.pointwise("mul", a1, k))
    assert np.array_equal(mul, ((m64a * 7) % (1 << 32)).astype(np.uint32))
