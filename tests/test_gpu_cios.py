"""GPU parity of the fast tile kernels on general moduli (the CIOS policy: 4 and 5 limbs with
4q < R): hmpc's examples/ntt.cpp prime (125 bits, 4 limbs) up to its N = 2^20 and the 150-bit
prime of examples/bgv.cpp (5 limbs), at both elements-per-thread schedules: transforms and
products against the oracle and the portable kernels, BGV encryption against the oracle, bit for
bit."""

import sys

import numpy as np
import pytest

import hmpc_amd
import oracle
from bigref import P150

pytestmark = pytest.mark.gpu

P125 = 0x18F25CD9A75CCBD9C146D4ABAEC00001  # examples/ntt.cpp:22
KEY = [42, 0, 0, 0, 0, 0, 0, 0]


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
        fast = hmpc_amd.Modulus(ctx, q)
        portable = hmpc_amd.Modulus(ctx, q)
        portable.set_kernels(1)
        _mods[q] = (fast, portable, oracle.Modulus(q))
    return _mods[q]


# P125 - 1 is divisible by 2^22 (N up to 2^21), P150 - 1 by 2^15 (N up to 2^14)
NTT_CASES = [(P125, 12, 3), (P125, 13, 2), (P125, 16, 1), (P125, 20, 1), (P150, 10, 5), (P150, 13, 2), (P150, 14, 2)]


@pytest.mark.parametrize("elems", [4, 8])
@pytest.mark.parametrize("q,log_n,n", NTT_CASES, ids=lambda v: hex(v)[:10] if isinstance(v, int) and v > 99 else str(v))
def test_ntt_general_moduli(ctx, q, log_n, n, elems):
    gm, pm, om = mods(ctx, q)
    gm.set_schedule(log_n, elems, -1)
    assert gm.schedule(log_n)["fast"], "the fast tile kernels serve this modulus"
    say(f"cios ntt q={hex(q)[:12]} L={gm.limbs} log_n={log_n} n={n} elems={elems}")
    count = n << log_n
    x = om.sample_uniform(KEY, 3, count, threads=8)
    want = om.ntt(x, log_n, threads=8)
    dx = gm.upload(x)
    f = gm.ntt(log_n, n, dx)
    assert np.array_equal(gm.download(f, count), want)
    assert np.array_equal(pm.download(pm.ntt(log_n, n, pm.upload(x)), count), want)
    back = gm.download(gm.ntt(log_n, n, f, inverse=True), count)
    assert np.array_equal(back, x)
    assert np.array_equal(gm.download(gm.ntt(log_n, n, dx, inverse=True), count), om.ntt(x, log_n, inverse=True, threads=8))


@pytest.mark.parametrize("elems", [4, 8])
@pytest.mark.parametrize("q,log_n,n", [(P125, 9, 6), (P125, 14, 2), (P125, 16, 1), (P150, 12, 2), (P150, 13, 2),
                                       (P150, 14, 1)],
                         ids=lambda v: hex(v)[:10] if isinstance(v, int) and v > 99 else str(v))
def test_polymul_general_moduli(ctx, q, log_n, n, elems):
    gm, pm, om = mods(ctx, q)
    gm.set_schedule(log_n, elems, -1)
    say(f"cios polymul q={hex(q)[:12]} L={gm.limbs} log_n={log_n} n={n} elems={elems}")
    count = n << log_n
    a = om.sample_uniform(KEY, 4, count, threads=8)
    b = om.sample_uniform(KEY, 5, count, threads=8)
    want, _ = om.polymul_timed(log_n, a, b, threads=8)
    got = gm.download(gm.polymul(log_n, n, gm.upload(a), gm.upload(b)), count)
    assert np.array_equal(got, want)
    assert np.array_equal(pm.download(pm.polymul(log_n, n, pm.upload(a), pm.upload(b)), count), want)


def test_edge_values_general_moduli(ctx):
    """q - 1, 0 and 1 everywhere: lazy values at the top of [0, 4q) still reduce exactly."""
    for q, log_n in ((P125, 13), (P150, 12)):
        gm, pm, om = mods(ctx, q)
        N = 1 << log_n
        vals = [q - 1] * N + [0] * N + [1, q - 1] * (N // 2)
        x = oracle.ints_to_planar(vals, gm.limbs)
        say(f"cios edge values q={hex(q)[:12]} log_n={log_n}")
        f = gm.download(gm.ntt(log_n, 3, gm.upload(x)), 3 * N)
        assert np.array_equal(f, om.ntt(x, log_n, threads=8))
        got = gm.download(gm.polymul(log_n, 3, gm.upload(x), gm.upload(x)), 3 * N)
        want, _ = om.polymul_timed(log_n, x, x, threads=8)
        assert np.array_equal(got, want)


P64_EXAMPLE = 0x89D046E0AE640001  # examples/bgv.cpp:29


@pytest.mark.parametrize("elems", [4, 8])
@pytest.mark.parametrize("q,log_n,n,p", [(P125, 13, 3, 1 << 32), (P125, 16, 1, P64_EXAMPLE), (P150, 13, 2, 1 << 32),
                                         (P150, 14, 3, P64_EXAMPLE)],
                         ids=lambda v: hex(v)[:10] if isinstance(v, int) and v > 99 else str(v))
def test_bgv_general_moduli(ctx, q, log_n, n, p, elems):
    """BGV encryption on the CIOS tile launches (N > 4096), with t = 2^32 and with the 64-bit
    plaintext modulus of examples/bgv.cpp (any p: taken as a Montgomery constant), against the
    oracle's ciphertexts; t = 2^32 ciphertexts also decrypt."""
    gm, pm, om = mods(ctx, q)
    gm.set_bgv_schedule(log_n, elems, -1)
    N = 1 << log_n
    say(f"cios bgv q={hex(q)[:12]} L={gm.limbs} log_n={log_n} n={n} p={hex(p)} elems={elems}")
    p_mont = om.from_unsigned(p)
    B = om.sample_uniform(KEY, 10, N)
    s = om.ntt(om.sample_cbd(KEY, 11, 2, N), log_n)
    e = om.ntt(om.sample_cbd(KEY, 12, 40, N), log_n)
    A = om.pointwise("add", om.pointwise("mul", B, s), om.pointwise("mul", e, oracle.ints_to_planar([p_mont] * N, gm.limbs)))
    msgs = oracle.sample_words(KEY, 13, n * N)
    enc_key = [42, 0, 0, 0, 0, 0, 0, 1]
    want0, want1, nonce = om.bgv_encrypt(log_n, A, B, msgs, p_mont, enc_key, [5, 0, 0], threads=8)
    pm_words = oracle.int_to_limbs(p_mont, gm.limbs)
    ctx.set_rng(enc_key, [5, 0, 0])
    c0, c1 = gm.bgv_encrypt(log_n, n, gm.upload(A), gm.upload(B), ctx.upload(msgs), pm_words)
    assert np.array_equal(gm.download(c0, n * N), want0)
    assert np.array_equal(gm.download(c1, n * N), want1)
    assert ctx.rng[1].tolist() == nonce.tolist()
    if p == 1 << 32:
        assert np.array_equal(gm.bgv_decrypt(log_n, n, gm.upload(s), c0, c1).to_numpy(n * N), msgs)
    gm.set_bgv_schedule(log_n, 0, -1)
