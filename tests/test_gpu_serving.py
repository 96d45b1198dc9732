"""GPU tests of the serving loop (hmpc_server_*): continuous batching must not change results.

A request's ciphertexts depend only on its inputs, key and nonce, so the same request served
alone and in the middle of a mix of other requests (other degrees, sizes and kinds, overlapping
launches on several streams) must give identical bytes, equal to the oracle's. Admission by
memory budget: requests wait while the budget is full and run once memory is released.
"""

import sys

import numpy as np
import pytest

import hmpc_amd
import oracle
from bigref import Q128A

pytestmark = pytest.mark.gpu

KEY = [42, 0, 0, 0, 0, 0, 0, 0]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    ctx = hmpc_amd.Context(0, key=KEY)
    gm = hmpc_amd.Modulus(ctx, Q128A)
    om = oracle.Modulus(Q128A)
    t_mont = om.from_unsigned(1 << 32)
    keys = {}
    for log_n in (12, 13, 14):
        N = 1 << log_n
        B = om.sample_uniform(KEY, 0, N)
        s = om.ntt(om.sample_cbd(KEY, 1, 2, N), log_n)
        e = om.ntt(om.sample_cbd(KEY, 2, 40, N), log_n)
        tcol = oracle.ints_to_planar([t_mont] * N, om.limbs)
        A = om.pointwise("add", om.pointwise("mul", B, s), om.pointwise("mul", e, tcol))
        keys[log_n] = (A, B, gm.upload(A), gm.upload(B))
    yield ctx, gm, om, t_mont, keys
    ctx.wait()


def _encrypt_expected(om, log_n, keys, msgs, t_mont, nonce):
    A, B, _, _ = keys[log_n]
    c0, c1, _ = om.bgv_encrypt(log_n, A, B, msgs, t_mont, KEY, nonce, threads=8)
    return c0, c1


def test_result_same_alone_and_in_mix(env):
    ctx, gm, om, t_mont, keys = env
    pm = oracle.int_to_limbs(t_mont, gm.limbs)
    srv = hmpc_amd.Server(ctx, budget_bytes=8 << 30, slots=3)
    kid = {ln: srv.add_key(gm, ln, keys[ln][2], keys[ln][3], pm) for ln in keys}
    # the request under test: 5 messages at N = 2^14
    log_n, k, nonce = 14, 5, [1000, 0, 0]
    N = 1 << log_n
    msgs = oracle.sample_words(KEY, 77, k * N)
    dmsgs = ctx.upload(msgs)
    want0, want1 = _encrypt_expected(om, log_n, keys, msgs, t_mont, nonce)

    say("serving: request alone")
    r = srv.submit_encrypt(kid[log_n], k, dmsgs.ptr, nonce)
    srv.wait(r)
    c0, c1 = srv.result(r)
    alone0, alone1 = c0.to_numpy((gm.limbs, k * N)), c1.to_numpy((gm.limbs, k * N))
    srv.release(r)
    assert np.array_equal(alone0, want0) and np.array_equal(alone1, want1)

    say("serving: same request in a mix of 40")
    rng = np.random.default_rng(3)
    pool = ctx.upload(oracle.sample_words(KEY, 78, 16 * (1 << 14)))
    pa = gm.upload(om.sample_uniform(KEY, 90, 4 << 13))
    pb = gm.upload(om.sample_uniform(KEY, 91, 4 << 13))
    others = []
    for i in range(40):
        if i == 20:
            target = srv.submit_encrypt(kid[log_n], k, dmsgs.ptr, nonce)
        if i % 5 == 4:
            others.append(("mul", srv.submit_polymul(gm, 13, 4, pa.ptr, pb.ptr)))
        else:
            ln = int(rng.choice([12, 13, 14]))
            kk = int(rng.integers(1, 16))
            others.append(("enc", srv.submit_encrypt(kid[ln], kk, pool.ptr, [5000 + 97 * i, 0, 0])))
    srv.wait(target)
    for _, rid in others:
        srv.wait(rid)
    c0, c1 = srv.result(target)
    assert np.array_equal(c0.to_numpy((gm.limbs, k * N)), alone0), "c0 differs in the mix"
    assert np.array_equal(c1.to_numpy((gm.limbs, k * N)), alone1), "c1 differs in the mix"
    # a product request of the mix against the direct batched product
    want_p = gm.download(gm.polymul(13, 4, pa, pb), 4 << 13)
    prod = [rid for kind, rid in others if kind == "mul"][0]
    p0, p1 = srv.result(prod)
    assert p1 is None and np.array_equal(p0.to_numpy((gm.limbs, 4 << 13)), want_p)
    done = dict(srv.completed())
    assert target in done and all(rid in done for _, rid in others)
    assert all(lat > 0 for lat in done.values())
    st = srv.stats()
    assert st["queued"] == 0 and st["running"] == 0 and st["launches"] >= 2
    for _, rid in others:
        srv.release(rid)
    srv.release(target)
    assert srv.stats()["bytes_in_use"] == 0
    srv.close()


def test_budget_admission_and_release(env):
    ctx, gm, om, t_mont, keys = env
    pm = oracle.int_to_limbs(t_mont, gm.limbs)
    # one request of 2 messages at N = 2^14 holds 2 x 2 x 2^14 x 16 B = 1 MiB of outputs
    srv = hmpc_amd.Server(ctx, budget_bytes=2 << 20, slots=2)
    kid = srv.add_key(gm, 14, keys[14][2], keys[14][3], pm)
    msgs = ctx.upload(oracle.sample_words(KEY, 79, 2 << 14))
    say("serving: budget admission")
    with pytest.raises(hmpc_amd.HmpcError) as err:
        srv.submit_encrypt(kid, 8, msgs.ptr, [0, 0, 0])  # 4 MiB > budget
    assert err.value.code == 5
    r1 = srv.submit_encrypt(kid, 2, msgs.ptr, [1, 0, 0])
    r2 = srv.submit_encrypt(kid, 2, msgs.ptr, [2, 0, 0])
    r3 = srv.submit_encrypt(kid, 2, msgs.ptr, [3, 0, 0])
    srv.wait(r1)
    srv.wait(r2)
    assert srv.poll(r3)[0] == srv.QUEUED  # budget full until a result is released
    assert srv.stats()["bytes_in_use"] == 2 << 20
    srv.release(r1)
    srv.wait(r3)
    state, lat = srv.poll(r3)
    assert state == srv.DONE and lat > 0
    # r3 equals the direct encryption with its nonce
    c0, _ = srv.result(r3)
    ctx.set_rng(KEY, [3, 0, 0])
    d0, _ = gm.bgv_encrypt(14, 2, keys[14][2], keys[14][3], msgs, pm)
    assert np.array_equal(c0.to_numpy((gm.limbs, 2 << 14)), gm.download(d0, 2 << 14))
    srv.release(r2)
    srv.release(r3)
    srv.close()


def test_large_request_spans_tile_chunks(env):
    """700 ciphertexts at N = 2^14 (q128a) hold 700 x 3 x 4 x 2^14 words of column results, more
    than one tile launch's 2^27-word workspace: the request runs as two chunks in sequence and
    equals the direct batched encryption with its nonce (the persistent kernel at this count)."""
    ctx, gm, om, t_mont, keys = env
    pm = oracle.int_to_limbs(t_mont, gm.limbs)
    log_n, k, nonce = 14, 700, [77, 1, 0]
    N = 1 << log_n
    srv = hmpc_amd.Server(ctx, budget_bytes=4 << 30, slots=1)
    kid = srv.add_key(gm, log_n, keys[log_n][2], keys[log_n][3], pm)
    msgs = ctx.upload(oracle.sample_words(KEY, 80, k * N))
    say("serving: 700 ciphertexts in two tile chunks")
    r = srv.submit_encrypt(kid, k, msgs.ptr, nonce)
    srv.wait(r)
    c0, c1 = srv.result(r)
    ctx.set_rng(KEY, nonce)
    d0, d1 = gm.bgv_encrypt(log_n, k, keys[log_n][2], keys[log_n][3], msgs, pm)
    assert gm.schedule(log_n)["fast"]
    assert np.array_equal(c0.to_numpy((gm.limbs, k * N)), gm.download(d0, k * N)), "c0"
    assert np.array_equal(c1.to_numpy((gm.limbs, k * N)), gm.download(d1, k * N)), "c1"
    srv.release(r)
    srv.close()


def test_mixed_degrees_share_launches(env):
    """Encryptions at N = 2^15 and 2^16 and products at N = 2^15, 2^16 and 2^17 submitted together:
    the server splits its tile launches by degree (row kernels for rows of 2^8, 2^9 and 2^10
    with the pass lengths at compile time); every result equals the direct batched call with the
    same inputs (and, for encryptions, the same nonce)."""
    ctx, gm, om, t_mont, keys = env
    pm = oracle.int_to_limbs(t_mont, gm.limbs)
    srv = hmpc_amd.Server(ctx, budget_bytes=8 << 30, slots=2)
    rng_key = {}
    for log_n in (15, 16):
        N = 1 << log_n
        A = gm.upload(om.sample_uniform(KEY, 60 + log_n, N))
        B = gm.upload(om.sample_uniform(KEY, 70 + log_n, N))
        rng_key[log_n] = (A, B, srv.add_key(gm, log_n, A, B, pm))
    say("serving: mixed degrees 2^15..2^17")
    enc = []
    for i, (log_n, k) in enumerate([(15, 3), (16, 2), (15, 5), (16, 1)]):
        N = 1 << log_n
        msgs = ctx.upload(oracle.sample_words(KEY, 100 + i, k * N))
        nonce = [9000 + 1000 * i, 0, 0]
        enc.append((log_n, k, msgs, nonce, srv.submit_encrypt(rng_key[log_n][2], k, msgs.ptr, nonce)))
    mul = []
    for i, (log_n, n) in enumerate([(15, 3), (16, 2), (17, 1)]):
        N = 1 << log_n
        a = gm.upload(om.sample_uniform(KEY, 120 + i, n * N))
        b = gm.upload(om.sample_uniform(KEY, 130 + i, n * N))
        mul.append((log_n, n, a, b, srv.submit_polymul(gm, log_n, n, a.ptr, b.ptr)))
    for *_, rid in enc + mul:
        srv.wait(rid)
    for log_n, k, msgs, nonce, rid in enc:
        N = 1 << log_n
        c0, c1 = srv.result(rid)
        ctx.set_rng(KEY, nonce)
        d0, d1 = gm.bgv_encrypt(log_n, k, rng_key[log_n][0], rng_key[log_n][1], msgs, pm)
        assert np.array_equal(c0.to_numpy((gm.limbs, k * N)), gm.download(d0, k * N)), ("c0", log_n, k)
        assert np.array_equal(c1.to_numpy((gm.limbs, k * N)), gm.download(d1, k * N)), ("c1", log_n, k)
        srv.release(rid)
    for log_n, n, a, b, rid in mul:
        N = 1 << log_n
        p0, _ = srv.result(rid)
        assert np.array_equal(p0.to_numpy((gm.limbs, n * N)), gm.download(gm.polymul(log_n, n, a, b), n * N)), log_n
        srv.release(rid)
    srv.close()
