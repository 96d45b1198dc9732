"""GPU parity of the element-wise operations around the hot path, through the C ABI:

- the tensor cipher (hmpc::expr::crypto::enc / dec with a ChaCha20 cipher) against the oracle,
  the RFC 8439 section 2.4.2 vector and hmpc's own round-trip test inputs;
- homomorphic operations on BGV ciphertexts (sum, difference, scalar and plaintext products)
  against the oracle's point-wise ops, and the properties decryption fixes for them;
- pinned host staging with asynchronous copies (the wire-format bridge);
- fused multiply-add, sum / product reductions and matrix-vector products over mod q.
"""

import json
import os
import sys

import numpy as np
import pytest

import hmpc_amd
import oracle
from bigref import Q64, Q128A

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
KEY = [42, 0, 0, 0, 0, 0, 0, 0]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = hmpc_amd.Context(0, key=KEY)
    yield c
    c.wait()


def _words_of(byte_list):
    b = bytes(byte_list)
    return np.frombuffer(b[: len(b) // 4 * 4], dtype="<u4").astype(np.uint32)


def test_cipher_rfc8439_vector(ctx):
    v = VEC["cipher"]["rfc8439_2_4_2"]
    say("cipher RFC 8439 2.4.2")
    plain, want = _words_of(v["plaintext"]), _words_of(v["ciphertext"])
    planar = plain.reshape(7, 4).T.copy()
    ct = ctx.cipher_encrypt(4, 7, v["key"], v["nonce"], ctx.upload(planar), counter=v["counter"])
    assert np.array_equal(ct.to_numpy(28), want)
    back = ctx.cipher_decrypt(4, 7, v["key"], v["nonce"], ct, counter=v["counter"])
    assert np.array_equal(back.to_numpy((4, 7)), planar)


@pytest.mark.parametrize("limbs,count", [(1, 1), (1, 4096), (2, 37), (2, 1 << 16), (4, 4099), (4, 1 << 18), (5, 1000),
                                         (8, 1000), (9, 333)])
def test_cipher_matches_oracle(ctx, limbs, count):
    say(f"cipher limbs={limbs} count={count}")
    rng = np.random.default_rng(limbs * 1000 + count)
    planar = rng.integers(0, 1 << 32, size=(limbs, count), dtype=np.uint64).astype(np.uint32)
    key, nonce = [3, 1, 4, 1, 5, 9, 2, 6], [5, 3, 5]
    want = oracle.cipher_encrypt(planar, key, nonce)
    ct = ctx.cipher_encrypt(limbs, count, key, nonce, ctx.upload(planar))
    assert np.array_equal(ct.to_numpy((count, limbs)), want)
    back = ctx.cipher_decrypt(limbs, count, key, nonce, ct)
    assert np.array_equal(back.to_numpy((limbs, count)), planar)


class _At:
    """a device address inside a buffer (to pass unaligned pointers through the C ABI)"""

    def __init__(self, buf, byte_offset):
        self.ptr = buf.ptr + byte_offset


@pytest.mark.parametrize("limbs", [1, 2, 4])
def test_cipher_unaligned_buffers(ctx, limbs):
    """the 16-byte fast path is taken only for aligned buffers; offset pointers give the same words"""
    say(f"cipher unaligned limbs={limbs}")
    count = 4096
    rng = np.random.default_rng(77 + limbs)
    planar = rng.integers(0, 1 << 32, size=(limbs, count), dtype=np.uint64).astype(np.uint32)
    key, nonce = [1, 2, 3, 4, 5, 6, 7, 8], [0, 0, 9]
    want = oracle.cipher_encrypt(planar, key, nonce)
    src = ctx.upload(np.concatenate([np.zeros(1, np.uint32), planar.reshape(-1)]))
    dst = ctx.buffer(4 * (limbs * count + 1))
    ctx.cipher_encrypt(limbs, count, key, nonce, _At(src, 4), out=_At(dst, 4))
    got = dst.to_numpy(limbs * count + 1)[1:]
    assert np.array_equal(got, want.reshape(-1))
    back = ctx.buffer(4 * (limbs * count + 1))
    ctx.cipher_decrypt(limbs, count, key, nonce, _At(dst, 4), out=_At(back, 4))
    assert np.array_equal(back.to_numpy(limbs * count + 1)[1:], planar.reshape(-1))


def test_cipher_reference_tensor_roundtrip(ctx):
    """tests/expr/crypto/cipher.cpp inputs: ubigint<127> x_i = i, shape {4, 2}, key {42}, nonce {1}"""
    t = VEC["cipher"]["tensor"]
    say("cipher hmpc tensor round trip")
    count = t["shape"][0] * t["shape"][1]
    planar = np.zeros((t["limbs"], count), np.uint32)
    planar[0] = np.arange(count, dtype=np.uint32)
    ct = ctx.cipher_encrypt(t["limbs"], count, t["key"], t["nonce"], ctx.upload(planar))
    assert np.array_equal(ct.to_numpy((count, t["limbs"])), oracle.cipher_encrypt(planar, t["key"], t["nonce"]))
    back = ctx.cipher_decrypt(t["limbs"], count, t["key"], t["nonce"], ct)
    assert np.array_equal(back.to_numpy((t["limbs"], count)), planar)


def test_cipher_rejects_aliasing_and_bad_limbs(ctx):
    buf = ctx.upload(np.zeros(16, np.uint32))
    with pytest.raises(hmpc_amd.HmpcError):
        ctx.cipher_encrypt(4, 4, KEY, [0, 0, 0], buf, out=buf)
    with pytest.raises(hmpc_amd.HmpcError):
        ctx.cipher_encrypt(10, 1, KEY, [0, 0, 0], buf)


@pytest.mark.parametrize("q", [Q64, Q128A])
def test_ciphertext_ops_match_oracle(ctx, q):
    gm, om = hmpc_amd.Modulus(ctx, q), oracle.Modulus(q)
    say(f"ciphertext ops q={hex(q)[:12]}")
    N, n = 1 << 10, 3
    count = n * N
    a0, a1 = om.sample_uniform(KEY, 1, count), om.sample_uniform(KEY, 2, count)
    b0, b1 = om.sample_uniform(KEY, 3, count), om.sample_uniform(KEY, 4, count)
    p = om.sample_uniform(KEY, 5, N)
    da0, da1, db0, db1, dp = (gm.upload(x) for x in (a0, a1, b0, b1, p))
    for op, name in ((gm.ADD, "add"), (gm.SUB, "sub"), (gm.MUL, "mul")):
        o0, o1 = gm.ciphertext_op(op, count, da0, da1, db0, db1)
        assert np.array_equal(gm.download(o0, count), om.pointwise(name, a0, b0)), name
        assert np.array_equal(gm.download(o1, count), om.pointwise(name, a1, b1)), name
    # plaintext polynomial broadcast over the batch
    o0, o1 = gm.ciphertext_op(gm.MUL, count, da0, da1, dp, dp, b_count=N)
    assert np.array_equal(gm.download(o0, count), om.pointwise("mul", a0, p))
    assert np.array_equal(gm.download(o1, count), om.pointwise("mul", a1, p))
    # scalar, in place
    c = om.from_unsigned(12345)
    gm.ciphertext_scalar(gm.MUL, count, da0, da1, oracle.int_to_limbs(c, gm.limbs), out0=da0, out1=da1)
    cc = oracle.ints_to_planar([c], om.limbs)
    assert np.array_equal(gm.download(da0, count), om.pointwise("mul", a0, cc))
    assert np.array_equal(gm.download(da1, count), om.pointwise("mul", a1, cc))


def test_ciphertext_ops_decrypt_to_plaintext_ops(ctx):
    """dec(ct_a + ct_b) = m_a + m_b, dec(ct_a - ct_b) = m_a - m_b, dec(k ct_a) = k m_a and
    dec(ct_a o NTT(x)) = x m_a (negacyclic shift) mod t = 2^32 (small noise)"""
    gm, om = hmpc_amd.Modulus(ctx, Q128A), oracle.Modulus(Q128A)
    say("ciphertext ops under decryption")
    log_n, n = 12, 2
    N, count = 1 << log_n, n << log_n
    t_mont = om.from_unsigned(1 << 32)
    B = om.sample_uniform(KEY, 10, N)
    s = om.ntt(om.sample_cbd(KEY, 11, 2, N), log_n)
    e = om.ntt(om.sample_cbd(KEY, 12, 40, N), log_n)
    A = om.pointwise("add", om.pointwise("mul", B, s), om.pointwise("mul", e, oracle.ints_to_planar([t_mont] * N, om.limbs)))
    dA, dB, ds = gm.upload(A), gm.upload(B), gm.upload(s)
    pm = oracle.int_to_limbs(t_mont, gm.limbs)
    ma, mb = oracle.sample_words(KEY, 13, count), oracle.sample_words(KEY, 14, count)
    ctx.set_rng([42, 0, 0, 0, 0, 0, 0, 1], [0, 0, 0])
    a0, a1 = gm.bgv_encrypt(log_n, n, dA, dB, ctx.upload(ma), pm)
    b0, b1 = gm.bgv_encrypt(log_n, n, dA, dB, ctx.upload(mb), pm)

    def dec(c0, c1):
        return gm.bgv_decrypt(log_n, n, ds, c0, c1).to_numpy(count)

    m64a, m64b = ma.astype(np.uint64), mb.astype(np.uint64)
    mask = np.uint64(0xFFFFFFFF)
    assert np.array_equal(dec(*gm.ciphertext_op(gm.ADD, count, a0, a1, b0, b1)), ((m64a + m64b) & mask).astype(np.uint32))
    assert np.array_equal(dec(*gm.ciphertext_op(gm.SUB, count, a0, a1, b0, b1)),
                          ((m64a - m64b + (np.uint64(1) << np.uint64(32))) & mask).astype(np.uint32))
    k = gm.mont(7)
    assert np.array_equal(dec(*gm.ciphertext_scalar(gm.MUL, count, a0, a1, k)), ((m64a * np.uint64(7)) & mask).astype(np.uint32))
    # x (the monomial) in NTT form: x * m shifts each polynomial by one, negating the wrapped coefficient
    x = np.zeros(N, np.uint32)
    x[1] = 1
    dx = gm.ntt(log_n, 1, gm.from_u32(N, ctx.upload(x)))
    got = dec(*gm.ciphertext_op(gm.MUL, count, a0, a1, dx, dx, b_count=N))
    want = np.empty(count, np.uint32)
    for j in range(n):
        poly = ma[j * N:(j + 1) * N].astype(np.int64)
        shifted = np.concatenate([[-poly[-1]], poly[:-1]])
        want[j * N:(j + 1) * N] = (shifted % (1 << 32)).astype(np.uint32)
    assert np.array_equal(got, want)


def test_pinned_host_staging_roundtrip(ctx):
    say("pinned host staging")
    nbytes = 4 << 20
    h = ctx.host_buffer(nbytes)
    data = np.random.default_rng(9).integers(0, 1 << 32, size=nbytes // 4, dtype=np.uint64).astype(np.uint32)
    h.array()[:] = data
    st = ctx.stream()
    dev = ctx.buffer(nbytes)
    h.to_device(dev, stream=st)
    st.synchronize()
    assert np.array_equal(dev.to_numpy(nbytes // 4), data)
    h2 = ctx.host_buffer(nbytes)
    h2.from_device(dev, stream=st)
    st.synchronize()
    assert np.array_equal(h2.array(), data)
    h.free()
    h2.free()


@pytest.mark.parametrize("q", [Q64, Q128A])
def test_fma_reduce_matvec_match_oracle(ctx, q):
    """fused a b + c against the oracle's point-wise ops; sum and product reductions and the
    matrix-vector product against Python big-integer arithmetic on the Montgomery values
    (a R + b R = (a + b) R; mont(a R, b R) = a b R)"""
    gm, om = hmpc_amd.Modulus(ctx, q), oracle.Modulus(q)
    say(f"fma / reduce / matvec q={hex(q)[:12]}")
    R = 1 << (32 * gm.limbs)
    Rinv = pow(R, -1, q)
    count = 100003
    a, b, c = (om.sample_uniform(KEY, 20 + i, count) for i in range(3))
    got = gm.download(gm.fma(count, gm.upload(a), gm.upload(b), gm.upload(c)), count)
    assert np.array_equal(got, om.pointwise("add", om.pointwise("mul", a, b), c))
    xs = oracle.planar_to_ints(a)
    for n in (1, 255, 256, 257, count):
        s = oracle.planar_to_ints(gm.download(gm.reduce(gm.ADD, n, gm.upload(a[:, :n].copy())), 1))[0]
        assert s == sum(xs[:n]) % q, n
    p_want = R % q
    for x in xs[:5000]:
        p_want = p_want * x * Rinv % q
    p = oracle.planar_to_ints(gm.download(gm.reduce(gm.MUL, 5000, gm.upload(a[:, :5000].copy())), 1))[0]
    assert p == p_want
    z = oracle.planar_to_ints(gm.download(gm.reduce(gm.ADD, 0, gm.upload(a[:, :1].copy())), 1))[0]
    assert z == 0
    # matvec: batch 3 of 17 x 70 matrices, then one matrix broadcast over 3 vectors; the same
    # with rows of 1500 (one workgroup per row)
    for batch, rows, cols in ((3, 17, 70), (2, 5, 1500)):
        _check_matvec(gm, om, batch, rows, cols, Rinv, q)


def _check_matvec(gm, om, batch, rows, cols, Rinv, q):
    M = om.sample_uniform(KEY, 30, batch * rows * cols)
    v = om.sample_uniform(KEY, 31, batch * cols)
    mi, vi = oracle.planar_to_ints(M), oracle.planar_to_ints(v)
    out = oracle.planar_to_ints(gm.download(gm.matvec(batch, rows, cols, gm.upload(M), gm.upload(v)), batch * rows))
    for bb in range(batch):
        for r in range(rows):
            want = sum(mi[(bb * rows + r) * cols + k] * vi[bb * cols + k] for k in range(cols)) * Rinv % q
            assert out[bb * rows + r] == want, (bb, r)
    out = oracle.planar_to_ints(gm.download(gm.matvec(batch, rows, cols, gm.upload(M[:, :rows * cols].copy()), gm.upload(v),
                                                      broadcast_matrix=True), batch * rows))
    for bb in range(batch):
        for r in range(rows):
            want = sum(mi[r * cols + k] * vi[bb * cols + k] for k in range(cols)) * Rinv % q
            assert out[bb * rows + r] == want, (bb, r)


@pytest.mark.parametrize("variant", [0, 1])
def test_matvec_long_rows_and_largest_values(ctx, variant):
    """q = 2^124 + c: the radix-2^26 product sums (variant 0: reductions only every 16 products per
    accumulator) and the portable Montgomery kernels (variant 1) give the same canonical results:
    rows long enough for many settles in both kernels (one wave and one workgroup per row), and
    matrices of q - 1 only (the largest canonical value in every product and partial sum)"""
    q = Q128A
    gm, om = hmpc_amd.Modulus(ctx, q), oracle.Modulus(q)
    gm.set_kernels(variant)
    Rinv = pow(1 << 128, -1, q)
    say(f"matvec long rows variant={variant}")
    for batch, rows, cols in ((2, 3, 1000), (1, 2, 40000)):
        _check_matvec(gm, om, batch, rows, cols, Rinv, q)
    for rows, cols in ((4, 1000), (2, 40000)):
        top = oracle.ints_to_planar([q - 1] * (rows * cols), 4)
        vec = oracle.ints_to_planar([q - 1] * cols, 4)
        out = oracle.planar_to_ints(gm.download(gm.matvec(1, rows, cols, gm.upload(top), gm.upload(vec)), rows))
        assert out == [cols * (q - 1) * (q - 1) * Rinv % q] * rows, (rows, cols)
