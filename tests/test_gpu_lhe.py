"""GPU parity of hmpc's general plaintext path through the C ABI against the CPU oracle, bit for bit:
cross-modulus casts (centred lift), capability sampling in every layout and distribution, BGV with an
Rp plaintext (coefficient or NTT form) and dec<plaintext>, on hmpc's own moduli: the 279-bit q and
64-bit p of tests/expr/crypto/lhe/enc.cpp and the 150-bit q and 64-bit p of examples/bgv.cpp.
Every case prints its name before touching the GPU.
"""

import json
import os
import sys

import numpy as np
import pytest

import hmpc_amd
import oracle
from bigref import P150, Q64, Q128A, Q128B

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
ENC = VEC["lhe"]["enc_test"]
EXAMPLE = VEC["lhe"]["bgv_example"]
Q279, P64_ENC = int(ENC["q"], 16), int(ENC["p"], 16)
P64_EXAMPLE = int(EXAMPLE["p"], 16)
KEY = [42, 0, 0, 0, 0, 0, 0, 0]
ENC_KEY = [42, 0, 0, 0, 0, 0, 0, 1]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = hmpc_amd.Context(0, key=KEY)
    yield c
    c.wait()


_mods = {}


def mods(ctx, q):
    if q not in _mods:
        _mods[q] = (hmpc_amd.Modulus(ctx, q), oracle.Modulus(q))
    return _mods[q]


CAST_PAIRS = [(P150, Q128A), (Q128A, P150), (Q279, P64_ENC), (P64_ENC, Q279), (Q64, Q128B), (Q128B, Q64),
              (P150, P64_EXAMPLE), (P64_EXAMPLE, P150), (99, P150), (P150, 99), (Q279, Q279)]


@pytest.mark.parametrize("src,dst", CAST_PAIRS, ids=lambda v: hex(v)[:10])
def test_cast_matches_oracle(ctx, src, dst):
    gs, os_ = mods(ctx, src)
    gd, od = mods(ctx, dst)
    say(f"cast {hex(src)[:12]} (L={gs.limbs}) -> {hex(dst)[:12]} (L={gd.limbs})")
    Rs = 1 << (32 * gs.limbs)
    edges = [0, 1, (src - 1) // 2, (src + 1) // 2, src - 1]
    x = np.concatenate([oracle.ints_to_planar([v * Rs % src for v in edges], gs.limbs),
                        os_.sample_uniform(KEY, 3, 3000)], axis=1)
    count = x.shape[1]
    want = os_.cast_to(od, x)
    got = gd.download(gs.cast(gd, count, gs.upload(x)), count)
    assert np.array_equal(got, want)
    if gs.limbs >= gd.limbs:
        # in place: each element's limbs are read before its (fewer) output limbs are written
        buf = gs.upload(x)
        gs.cast(gd, count, buf, out=buf)
        assert np.array_equal(gd.download(buf, count), want)


SAMPLER_CASES = [("uniform", 0), ("centered_binomial", 2), ("centered_binomial", 40), ("centered_binomial", 64),
                 ("centered_binomial", hmpc_amd.BITS_TERNARY),
                 ("binomial", 20), ("binomial", 200), ("drown_signed", oracle.drown_bits(1 << 64, True)),
                 ("drown_unsigned", oracle.drown_bits(12345, False)), ("drown_signed", 300)]


@pytest.mark.parametrize("q", [Q64, Q128A, P150, Q279], ids=lambda v: hex(v)[:10])
@pytest.mark.parametrize("layout", ["element", "ntt", "intt"])
@pytest.mark.parametrize("dist,bits", SAMPLER_CASES)
def test_sample_capability_matches_oracle(ctx, q, layout, dist, bits):
    gm, om = mods(ctx, q)
    log_n, n = 6, 3
    say(f"sample_capability q={hex(q)[:12]} L={gm.limbs} {layout} {dist}({bits})")
    ctx.set_rng(ENC_KEY, [0xFFFFFF80, 7, 0])
    got = gm.download(gm.sample_capability(dist, bits, layout, log_n, n), n << log_n)
    want = om.sample_capability(ENC_KEY, [0xFFFFFF80, 7, 0], dist, bits, layout, log_n, n)
    assert np.array_equal(got, want)
    # the 96-bit nonce advanced by n N (carrying into the second word)
    assert ctx.rng[1].tolist() == [(0xFFFFFF80 + (n << log_n)) & 0xFFFFFFFF, 8, 0]


def test_sample_capability_element_tensor_without_polynomials(ctx):
    """log_n = 0: a plain tensor of elements, one stream per element"""
    gm, om = mods(ctx, Q128A)
    say("sample_capability element log_n=0")
    ctx.set_rng(ENC_KEY, [5, 0, 0])
    got = gm.download(gm.sample_capability("uniform", 0, "element", 0, 1000), 1000)
    assert np.array_equal(got, om.sample_capability(ENC_KEY, [5, 0, 0], "uniform", 0, "element", 0, 1000))


def test_sample_capability_rejects_bad_arguments(ctx):
    gm, _ = mods(ctx, Q128A)
    say("sample_capability argument checks")
    for dist, bits, layout, log_n in [("centered_binomial", 3, "ntt", 4), ("binomial", 0, "ntt", 4),
                                      ("drown_signed", 400, "ntt", 4), ("uniform", 0, "ntt", 0)]:
        with pytest.raises(hmpc_amd.HmpcError):
            gm.sample_capability(dist, bits, layout, log_n, 1)


def _iota_planar(count, limbs):
    out = np.zeros((limbs, count), np.uint32)
    out[0] = np.arange(count, dtype=np.uint32)
    return out


def test_reference_linear_bgv_case(ctx):
    """hmpc's own test, tests/expr/crypto/lhe/enc.cpp:15-72, at its size: q of 279 bits (9 limbs),
    p = 0x8822d80623320001, N = 2^16, shape {4, 2}, a = b = raw words i, s = one, NTT-form plaintext
    x = raw words i. Ciphertexts equal the oracle's; dec<plaintext>(s, c) == x on the GPU."""
    q, p, log_n = Q279, P64_ENC, ENC["log_n"]
    n = ENC["shape"][0] * ENC["shape"][1]
    N = 1 << log_n
    gq, oq = mods(ctx, q)
    gp, op = mods(ctx, p)
    say(f"reference linear BGV: q L={gq.limbs}, p L={gp.limbs}, N=2^{log_n}, n={n}")
    a = _iota_planar(N, gq.limbs)
    s = oracle.ints_to_planar([oq.from_unsigned(1)] * N, gq.limbs)
    x = _iota_planar(n * N, gp.limbs)
    want0, want1, nonce = oq.bgv_encrypt_plaintext(op, log_n, a, a, x, True, ENC_KEY, [0, 0, 0], threads=8)
    ctx.set_rng(ENC_KEY, [0, 0, 0])
    da = gq.upload(a)
    c0, c1 = gq.bgv_encrypt_plaintext(gp, log_n, n, da, da, gp.upload(x), True)
    assert np.array_equal(gq.download(c0, n * N), want0)
    assert np.array_equal(gq.download(c1, n * N), want1)
    assert ctx.rng[1].tolist() == nonce.tolist()
    y = gq.bgv_decrypt_plaintext(gp, log_n, n, gq.upload(s), c0, c1, True)
    assert np.array_equal(gp.download(y, n * N), x)


def _keygen(om, log_n, p_mont, offset=0):
    N = 1 << log_n
    B = om.sample_uniform(KEY, offset, N)
    s = om.ntt(om.sample_cbd(KEY, offset + 1, 2, N), log_n)
    e = om.ntt(om.sample_cbd(KEY, offset + 2, 40, N), log_n)
    col = oracle.ints_to_planar([p_mont] * N, om.limbs)
    A = om.pointwise("add", om.pointwise("mul", B, s), om.pointwise("mul", e, col))
    return A, B, s


@pytest.mark.parametrize("plaintext_ntt", [True, False])
@pytest.mark.parametrize("v_dist", ["centered_binomial", "drown_signed"])
def test_bgv_example_moduli(ctx, plaintext_ntt, v_dist):
    """examples/bgv.cpp:25-29: the 150-bit q (5 limbs), N = 2^14, plaintext modulus 0x89d046e0ae640001;
    project keygen A = B s + p e. Ciphertexts equal the oracle's (randomness and drowning randomness)
    and decrypt to the plaintext."""
    q, p, log_n = int(EXAMPLE["q"], 16), P64_EXAMPLE, EXAMPLE["log_n"]
    gq, oq = mods(ctx, q)
    gp, op = mods(ctx, p)
    n, N = 3, 1 << log_n
    say(f"bgv example: q L={gq.limbs} p L={gp.limbs} N=2^{log_n} n={n} ntt={plaintext_ntt} v={v_dist}")
    A, B, s = _keygen(oq, log_n, oq.from_unsigned(p))
    x = op.sample_uniform(KEY, 5, n * N)
    # drowning bound 2: v below 2^81, so p v stays below q / 2 and the ciphertext still decrypts
    bits_v = hmpc_amd.runtime.drown_bits(2, True) if v_dist == "drown_signed" else 40
    assert bits_v == (oracle.drown_bits(2, True) if v_dist == "drown_signed" else 40)
    want0, want1, nonce = oq.bgv_encrypt_plaintext(op, log_n, A, B, x, plaintext_ntt, ENC_KEY, [3, 0, 0], 2, v_dist,
                                                   bits_v, 40, threads=8)
    ctx.set_rng(ENC_KEY, [3, 0, 0])
    c0, c1 = gq.bgv_encrypt_plaintext(gp, log_n, n, gq.upload(A), gq.upload(B), gp.upload(x), plaintext_ntt,
                                      v_dist=v_dist, bits_v=bits_v)
    assert np.array_equal(gq.download(c0, n * N), want0)
    assert np.array_equal(gq.download(c1, n * N), want1)
    assert ctx.rng[1].tolist() == nonce.tolist()
    y = gp.download(gq.bgv_decrypt_plaintext(gp, log_n, n, gq.upload(s), c0, c1, plaintext_ntt), n * N)
    assert np.array_equal(y, x)
    assert np.array_equal(y, oq.bgv_decrypt_plaintext(op, log_n, s, want0, want1, plaintext_ntt, threads=8))


@pytest.mark.parametrize("q,p,log_n", [(Q128A, Q64, 14), (Q128B, P64_EXAMPLE, 12), (Q64, 12289, 10)],
                         ids=["q128a-goldilocks", "q128b-p64", "goldilocks-p12289"])
def test_bgv_plaintext_on_fast_moduli(ctx, q, p, log_n):
    """The plaintext path on the moduli of the fast kernels (their NTTs run there)."""
    gq, oq = mods(ctx, q)
    gp, op = mods(ctx, p)
    n, N = 2, 1 << log_n
    say(f"bgv plaintext fast moduli q={hex(q)[:12]} p={hex(p)} N=2^{log_n}")
    A, B, s = _keygen(oq, log_n, oq.from_unsigned(p), offset=20)
    x = op.sample_uniform(KEY, 6, n * N)
    want0, want1, _ = oq.bgv_encrypt_plaintext(op, log_n, A, B, x, False, ENC_KEY, [0, 0, 0], threads=8)
    ctx.set_rng(ENC_KEY, [0, 0, 0])
    c0, c1 = gq.bgv_encrypt_plaintext(gp, log_n, n, gq.upload(A), gq.upload(B), gp.upload(x), False)
    assert np.array_equal(gq.download(c0, n * N), want0)
    assert np.array_equal(gq.download(c1, n * N), want1)
    y = gp.download(gq.bgv_decrypt_plaintext(gp, log_n, n, gq.upload(s), c0, c1, False), n * N)
    assert np.array_equal(y, x)


@pytest.mark.parametrize("q,log_n,n", [(P150, 14, 2), (Q279, 12, 2), (P150, 11, 3)], ids=["L5-2^14", "L9-2^12", "L5-2^11"])
def test_bgv_u32_messages_on_hmpc_moduli(ctx, q, log_n, n):
    """The 32-bit-message encryption on hmpc's own moduli (L = 5: examples/bgv.cpp:25; L = 9:
    tests/expr/crypto/lhe/enc.cpp:17), batched and per ciphertext (plain), against the oracle."""
    gm, om = mods(ctx, q)
    N = 1 << log_n
    say(f"bgv u32 q L={gm.limbs} N=2^{log_n} n={n}")
    t_mont = om.from_unsigned(1 << 32)
    A, B, s = _keygen(om, log_n, t_mont)
    msgs = oracle.sample_words(KEY, 3, n * N)
    want0, want1, nonce = om.bgv_encrypt(log_n, A, B, msgs, t_mont, ENC_KEY, [0, 0, 0], threads=8)
    pm = oracle.int_to_limbs(t_mont, gm.limbs)
    dA, dB, dm = gm.upload(A), gm.upload(B), ctx.upload(msgs)
    for plain in (False, True):
        say(f"  plain={plain}")
        ctx.set_rng(ENC_KEY, [0, 0, 0])
        c0, c1 = gm.bgv_encrypt(log_n, n, dA, dB, dm, pm, plain=plain)
        assert np.array_equal(gm.download(c0, n * N), want0)
        assert np.array_equal(gm.download(c1, n * N), want1)
        assert ctx.rng[1].tolist() == nonce.tolist()
    assert np.array_equal(gm.bgv_decrypt(log_n, n, gm.upload(s), c0, c1).to_numpy(n * N), msgs)
