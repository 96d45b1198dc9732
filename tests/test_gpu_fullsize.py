"""Full-size GPU parity for the measured workloads: every coefficient of the batched product (512
pairs, N = 2^16, q = 2^124 + 2^20 + 2^18 + 1) and of the batched BGV encryption (1024 ciphertexts,
N = 2^14, q = 2^124 + 0x78001, t = 2^32) equals the oracle's, on each schedule the library can
choose, and every ciphertext decrypts to its message. (Small-batch parity is in test_gpu_parity.py;
these cover what only a full batch exercises: every tile of every polynomial, job and tile tables
at their full length, workspace offsets of the last polynomials.)"""

import os
import sys

import numpy as np
import pytest

import hmpc_amd
import oracle
from bigref import Q128A, Q128B

pytestmark = pytest.mark.gpu

KEY = [42, 0, 0, 0, 0, 0, 0, 0]
ENC_KEY = [42, 0, 0, 0, 0, 0, 0, 1]
THREADS = max(1, min(len(os.sched_getaffinity(0)), 16))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = hmpc_amd.Context(0, key=KEY)
    yield c
    c.wait()


def test_c2_full_batch(ctx):
    log_n, n = 16, 512
    count = n << log_n
    gm, om = hmpc_amd.Modulus(ctx, Q128A), oracle.Modulus(Q128A)
    say(f"full C2: {n} pairs N=2^{log_n} q128a, oracle on {THREADS} threads")
    a = om.sample_uniform(KEY, 0, count, threads=THREADS)
    b = om.sample_uniform(KEY, 1, count, threads=THREADS)
    want, _ = om.polymul_timed(log_n, a, b, threads=THREADS)
    da, db = gm.upload(a), gm.upload(b)
    out = gm.alloc(count)
    for elems, persistent in ((0, -2), (4, 0), (4, 1), (8, 1)):
        say(f"  schedule elems={elems} persistent={persistent}")
        gm.set_schedule(log_n, elems, persistent)
        gm.polymul(log_n, n, da, db, out=out)
        got = gm.download(out, count)
        bad = np.flatnonzero((got != want).any(axis=0))
        assert bad.size == 0, f"{bad.size} coefficients differ, first at pair {bad[0] >> log_n}"
    gm.set_schedule(log_n, 0, -2)


def test_c3_full_batch(ctx):
    log_n, n = 14, 1024
    N, count = 1 << log_n, n << log_n
    gm, om = hmpc_amd.Modulus(ctx, Q128B), oracle.Modulus(Q128B)
    say(f"full C3: {n} ciphertexts N=2^{log_n} q128b, oracle on {THREADS} threads")
    t_mont = om.from_unsigned(1 << 32)
    B = om.sample_uniform(KEY, 0, N)
    s = om.ntt(om.sample_cbd(KEY, 1, 2, N), log_n)
    e = om.ntt(om.sample_cbd(KEY, 2, 40, N), log_n)
    A = om.pointwise("add", om.pointwise("mul", B, s), om.pointwise("mul", e, oracle.ints_to_planar([t_mont] * N, 4)))
    msgs = oracle.sample_words(KEY, 3, count)
    want0, want1, nonce = om.bgv_encrypt(log_n, A, B, msgs, t_mont, ENC_KEY, [0, 0, 0], threads=THREADS)
    dA, dB, dm = gm.upload(A), gm.upload(B), ctx.upload(msgs)
    pm = oracle.int_to_limbs(t_mont, 4)
    c0, c1 = gm.alloc(count), gm.alloc(count)
    for elems, tiles in ((0, -1), (8, 1), (4, 0)):
        say(f"  bgv schedule elems={elems} tiles={tiles}")
        gm.set_bgv_schedule(log_n, elems, tiles)
        ctx.set_rng(ENC_KEY, [0, 0, 0])
        gm.bgv_encrypt(log_n, n, dA, dB, dm, pm, c0=c0, c1=c1)
        for got, want, name in ((gm.download(c0, count), want0, "c0"), (gm.download(c1, count), want1, "c1")):
            bad = np.flatnonzero((got != want).any(axis=0))
            assert bad.size == 0, f"{name}: {bad.size} coefficients differ, first in ciphertext {bad[0] >> log_n}"
        assert ctx.rng[1].tolist() == nonce.tolist()
    gm.set_bgv_schedule(log_n, 0, -1)
    dec = gm.bgv_decrypt(log_n, n, gm.upload(s), c0, c1).to_numpy(count)
    assert np.array_equal(dec, msgs)
