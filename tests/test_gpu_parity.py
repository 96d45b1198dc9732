"""GPU parity: the HIP library through its C ABI against the CPU oracle, bit for bit.

Inputs are seeded explicit-key ChaCha20 streams (key 42), sizes small enough for the
oracle to finish in seconds. Every case prints its name before touching the GPU.
"""

import sys

import numpy as np
import pytest

import hmpc_amd
import oracle
from bigref import P150, Q64, Q128A, Q128B

pytestmark = pytest.mark.gpu

Q279 = 0x591F5B834C0D961F67343BCC8902BDEDA2771F54306FF151162FF8D2B40F4194DC0001  # tests/expr/crypto/lhe/enc.cpp:15
# primes = 1 mod 2^14 that select each arithmetic form of the fast kernels
Q_GENERIC4 = 0x80CD613ED8F16ADF91B7584A226B8001  # 128-bit: no 4q < R headroom
Q_LAZY4 = 0x201E2FEB414C343C1027C4D1C3924001     # 126-bit, dense limbs
Q_LAZY2 = 0x20030B397ED50001                     # 62-bit
Q_GENERIC2 = 0xC78E5106731EC001                  # 64-bit, not Goldilocks
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


def mods(ctx, q, variant=0):
    """(GPU modulus, oracle modulus); variant 0 = fast kernels, 1 = portable kernels."""
    if (q, variant) not in _mods:
        gm = hmpc_amd.Modulus(ctx, q)
        gm.set_kernels(variant)
        _mods[(q, variant)] = (gm, oracle.Modulus(q))
    return _mods[(q, variant)]


NTT_CASES = [
    (Q64, 2, 3), (Q64, 3, 2), (Q64, 5, 4), (Q64, 9, 13), (Q64, 12, 3), (Q64, 13, 2), (Q64, 14, 2), (Q64, 15, 1),
    (Q64, 16, 1), (Q64, 17, 1),
    (Q128A, 2, 1), (Q128A, 4, 3), (Q128A, 7, 33), (Q128A, 12, 2), (Q128A, 13, 2), (Q128A, 15, 1), (Q128A, 16, 2),
    (Q128A, 17, 1), (Q128B, 14, 3),
    (Q_GENERIC4, 12, 2), (Q_GENERIC4, 14, 2), (Q_LAZY4, 12, 2), (Q_LAZY4, 13, 3), (Q_LAZY2, 11, 3), (Q_LAZY2, 14, 1),
    (Q_GENERIC2, 12, 2), (Q_GENERIC2, 13, 1),
    (P150, 11, 2), (Q279, 10, 2),
]


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("q,log_n,n", NTT_CASES, ids=lambda v: hex(v)[:10] if isinstance(v, int) and v > 99 else str(v))
def test_ntt_matches_oracle(ctx, q, log_n, n, variant):
    gm, om = mods(ctx, q, variant)
    say(f"ntt q={hex(q)[:12]} L={gm.limbs} log_n={log_n} n={n} variant={variant}")
    count = n << log_n
    x = om.sample_uniform(KEY, 0, count, threads=8)
    want_f = om.ntt(x, log_n, threads=8)
    want_i = om.ntt(x, log_n, inverse=True, threads=8)
    dx = gm.upload(x)
    for plain in (False, True):
        f = gm.download(gm.ntt(log_n, n, dx, plain=plain), count)
        assert np.array_equal(f, want_f), f"forward plain={plain}"
        i = gm.download(gm.ntt(log_n, n, dx, inverse=True, plain=plain), count)
        assert np.array_equal(i, want_i), f"inverse plain={plain}"
    # in place round trip
    buf = gm.upload(x)
    gm.ntt(log_n, n, buf, out=buf)
    gm.ntt(log_n, n, buf, out=buf, inverse=True)
    assert np.array_equal(gm.download(buf, count), x)


@pytest.mark.parametrize("q,log_n", [(Q64, 12), (Q128A, 16), (Q128B, 14), (P150, 14)])
def test_twiddles_and_root_match_oracle(ctx, q, log_n):
    gm, om = mods(ctx, q)
    say(f"twiddles q={hex(q)[:12]} log_n={log_n}")
    assert gm.root(log_n) == om.root_of_unity(log_n + 1)
    assert gm.root(log_n, inverse=True) == om.root_of_unity(log_n + 1, inverse=True)
    assert np.array_equal(gm.twiddles(log_n), om.roots(log_n))
    assert np.array_equal(gm.twiddles(log_n, inverse=True), om.roots(log_n, inverse=True))
    npr, r1, r2 = gm.constants()
    assert (npr, r1, r2) == om.constants()


def test_poly_mod_kat_on_gpu(ctx):
    """tests/ints/poly_mod.cpp:60-103 on the GPU: x_i = i, two polynomials, N = 2^14."""
    gm, om = mods(ctx, P150)
    say("poly_mod KAT p150 N=2^14")
    log_n, N = 14, 1 << 14
    R = 1 << (32 * gm.limbs)
    x = oracle.ints_to_planar([(i % N) * R % P150 for i in range(2 * N)], gm.limbs)
    y = gm.download(gm.ntt(log_n, 2, gm.upload(x)), 2 * N)
    vals = oracle.planar_to_ints(y)
    Rinv = pow(R, -1, P150)
    kat = {0: 0x7bf0fe81e23706642244f0cf8fb65af112666, 1: 0xb8129497504c45e2fab6f665863b8d0cdb431,
           2: 0x25b2015ae56339a59389c056ae782f1e58b697, 10: 0x2433793de67b5574558745981a36837082b73e,
           19: 0x1cdb8c2cce3758d6acd87bf78549948e08faba}
    for j in range(2):
        for k, want in kat.items():
            assert vals[j * N + k] * Rinv % P150 == want


POLYMUL_CASES = [(Q64, 4, 3), (Q64, 10, 4), (Q64, 12, 3), (Q64, 13, 2), (Q128A, 12, 2), (Q128A, 15, 2), (Q128A, 16, 2),
                 (Q128A, 17, 1),
                 # more column tiles than the streaming column passes' 512 resident workgroups, not a
                 # multiple of them (1280 and 264 forward tiles: workgroups walk 2 or 3 tiles, or 1)
                 (Q128A, 16, 40), (Q128B, 14, 33),
                 (Q64, 18, 1),
                 (Q128A, 5, 7), (Q128B, 14, 2), (Q128B, 13, 3), (Q_GENERIC4, 12, 2), (Q_GENERIC4, 14, 1),
                 (Q_LAZY4, 13, 2), (Q_LAZY2, 12, 2), (Q_GENERIC2, 13, 2), (P150, 9, 2)]


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("q,log_n,n", POLYMUL_CASES, ids=lambda v: hex(v)[:10] if isinstance(v, int) and v > 99 else str(v))
def test_polymul_matches_oracle(ctx, q, log_n, n, variant):
    gm, om = mods(ctx, q, variant)
    say(f"polymul q={hex(q)[:12]} log_n={log_n} n={n} variant={variant}")
    count = n << log_n
    a = om.sample_uniform(KEY, 0, count, threads=8)
    b = om.sample_uniform(KEY, 1, count, threads=8)
    want, _ = om.polymul_timed(log_n, a, b, threads=8)
    got = gm.download(gm.polymul(log_n, n, gm.upload(a), gm.upload(b)), count)
    assert np.array_equal(got, want)


def test_polymul_equals_schoolbook_c1(ctx):
    """N=4096, q64: INTT(NTT(x) o NTT(y)) == schoolbook negacyclic product."""
    gm, om = mods(ctx, Q64)
    say("polymul vs schoolbook q64 N=4096")
    x = om.sample_uniform(KEY, 0, 4096)
    y = om.sample_uniform(KEY, 1, 4096)
    got = gm.download(gm.polymul(12, 1, gm.upload(x), gm.upload(y)), 4096)
    assert np.array_equal(got, om.schoolbook_negacyclic(12, x, y))


@pytest.mark.parametrize("q", [Q64, Q128A, P150, Q279])
def test_samplers_match_oracle(ctx, q):
    gm, om = mods(ctx, q)
    say(f"samplers q={hex(q)[:12]}")
    count = 5000
    for offset in (0, 3):
        got = gm.download(gm.sample_uniform(KEY, offset, count), count)
        assert np.array_equal(got, om.sample_uniform(KEY, offset, count))
    for bits in (2, 20, 40, 64):
        got = gm.download(gm.sample_cbd(KEY, 2, bits, count), count)
        assert np.array_equal(got, om.sample_cbd(KEY, 2, bits, count)), bits
    words = ctx.sample_words(KEY, 3, count).to_numpy(count)
    assert np.array_equal(words, oracle.sample_words(KEY, 3, count))


def test_from_u32_and_to_unsigned(ctx):
    gm, om = mods(ctx, Q128A)
    say("from_u32 / to_unsigned")
    count = 4096
    words = oracle.sample_words(KEY, 5, count)
    dw = ctx.upload(words)
    m = gm.download(gm.from_u32(count, dw), count)
    R = 1 << 128
    assert oracle.planar_to_ints(m) == [int(w) * R % Q128A for w in words]
    back = gm.download(gm.to_unsigned(count, gm.upload(m)), count)
    assert oracle.planar_to_ints(back) == [int(w) for w in words]


@pytest.mark.parametrize("q,log_n,n,bits", [(Q64, 12, 3, 40), (Q128B, 14, 2, 2), (Q128A, 6, 3, 64), (P150, 10, 2, 40),
                                            (Q128B, 12, 2, hmpc_amd.BITS_TERNARY)])
def test_capability_cbd_ntt(ctx, q, log_n, n, bits):
    gm, om = mods(ctx, q)
    say(f"cbd_ntt q={hex(q)[:12]} log_n={log_n} n={n} bits={bits}")
    key = [42, 0, 0, 0, 0, 0, 0, 1]
    ctx.set_rng(key, [7, 0, 0])
    got = gm.download(gm.sample_cbd_ntt(bits, log_n, n), n << log_n)
    want = om.ntt(om.sample_cbd_capability(key, [7, 0, 0], bits, log_n, n), log_n, threads=8)
    assert np.array_equal(got, want)
    assert ctx.rng[1].tolist() == [7 + (n << log_n), 0, 0]


def _keygen(om, log_n, t_mont):
    N = 1 << log_n
    B = om.sample_uniform(KEY, 0, N)
    s = om.ntt(om.sample_cbd(KEY, 1, 2, N), log_n)
    e = om.ntt(om.sample_cbd(KEY, 2, 40, N), log_n)
    tcol = oracle.ints_to_planar([t_mont] * N, om.limbs)
    A = om.pointwise("add", om.pointwise("mul", B, s), om.pointwise("mul", e, tcol))
    return A, B, s


# (Q128A, 17): the largest degree of any q = 2^124 + c with c < 2^24 (v2(q - 1) <= 18), so the longest
# lazy chain of the radix-2^26 transforms
@pytest.mark.parametrize("q,log_n,n", [(Q128B, 14, 3), (Q128A, 12, 2), (Q64, 10, 3), (Q128A, 15, 2), (Q128A, 16, 1),
                                       (Q64, 15, 2), (Q128A, 13, 4), (Q128A, 17, 1)])
def test_bgv_encrypt_matches_oracle_and_decrypts(ctx, q, log_n, n):
    gm, om = mods(ctx, q)
    say(f"bgv q={hex(q)[:12]} log_n={log_n} n={n}")
    N = 1 << log_n
    t_mont = om.from_unsigned(1 << 32)
    A, B, s = _keygen(om, log_n, t_mont)
    msgs = oracle.sample_words(KEY, 3, n * N)
    enc_key = [42, 0, 0, 0, 0, 0, 0, 1]
    want0, want1, nonce = om.bgv_encrypt(log_n, A, B, msgs, t_mont, enc_key, [0, 0, 0], threads=8)
    dA, dB, dm = gm.upload(A), gm.upload(B), ctx.upload(msgs)
    pm = oracle.int_to_limbs(t_mont, gm.limbs)
    gp, _ = mods(ctx, q, 1)
    fast_elems = (4, 8) if gm.limbs == 4 else (8, 16)
    modes = [("jobs", e) for e in fast_elems] + ([("tiles", e) for e in fast_elems] if log_n > 12 else [])
    for mode, elems in modes + [("portable", 0), ("plain", 0)]:
        say(f"  bgv mode={mode} elems={elems}")
        if mode in ("jobs", "tiles"):
            gm.set_bgv_schedule(log_n, elems, 1 if mode == "tiles" else 0)
        ctx.set_rng(enc_key, [0, 0, 0])
        g = gp if mode == "portable" else gm
        c0, c1 = g.bgv_encrypt(log_n, n, dA, dB, dm, pm, plain=mode == "plain")
        assert np.array_equal(gm.download(c0, n * N), want0), f"c0 mode={mode} elems={elems}"
        assert np.array_equal(gm.download(c1, n * N), want1), f"c1 mode={mode} elems={elems}"
        assert ctx.rng[1].tolist() == nonce.tolist()
    gm.set_bgv_schedule(log_n, 0, -1)
    dec = gm.bgv_decrypt(log_n, n, gm.upload(s), c0, c1).to_numpy(n * N)
    assert np.array_equal(dec, msgs)
    assert np.array_equal(dec, om.bgv_decrypt(log_n, s, want0, want1))


# The common BGV parameter set: s and u uniform ternary, errors centred binomial with eta = 10 (20
# bits): on the tile launches, the job kernel, the portable kernels and the plain path.
@pytest.mark.parametrize("q,log_n,n", [(Q128B, 14, 3), (Q64, 13, 2), (Q128A, 12, 2), (P150, 13, 2)])
def test_bgv_ternary_eta10_randomness(ctx, q, log_n, n):
    gm, om = mods(ctx, q)
    say(f"bgv ternary/eta10 q={hex(q)[:12]} log_n={log_n} n={n}")
    N = 1 << log_n
    bits = (hmpc_amd.BITS_TERNARY, 20, 20)
    t_mont = om.from_unsigned(1 << 32)
    B = om.sample_uniform(KEY, 0, N)
    s = om.ntt(om.sample_capability(KEY, [1, 0, 0], "centered_binomial", hmpc_amd.BITS_TERNARY, "element", log_n, 1), log_n)
    e = om.ntt(om.sample_cbd(KEY, 2, 20, N), log_n)
    A = om.pointwise("add", om.pointwise("mul", B, s), om.pointwise("mul", e, oracle.ints_to_planar([t_mont] * N, om.limbs)))
    msgs = oracle.sample_words(KEY, 3, n * N)
    enc_key = [42, 0, 0, 0, 0, 0, 0, 1]
    want0, want1, nonce = om.bgv_encrypt(log_n, A, B, msgs, t_mont, enc_key, [0, 0, 0], bits=bits, threads=8)
    dA, dB, dm = gm.upload(A), gm.upload(B), ctx.upload(msgs)
    pm = oracle.int_to_limbs(t_mont, gm.limbs)
    gp, _ = mods(ctx, q, 1)
    fast_r26_gold = q in (Q64, Q128A, Q128B)
    modes = ([("jobs", 8)] if fast_r26_gold else []) + [("tiles", 8), ("portable", 0), ("plain", 0)]
    for mode, elems in modes:
        say(f"  bgv mode={mode} elems={elems}")
        if mode in ("jobs", "tiles"):
            gm.set_bgv_schedule(log_n, elems, 1 if mode == "tiles" else 0)
        ctx.set_rng(enc_key, [0, 0, 0])
        g = gp if mode == "portable" else gm
        c0, c1 = g.bgv_encrypt(log_n, n, dA, dB, dm, pm, bits=bits, plain=mode == "plain")
        assert np.array_equal(gm.download(c0, n * N), want0), f"c0 mode={mode}"
        assert np.array_equal(gm.download(c1, n * N), want1), f"c1 mode={mode}"
        assert ctx.rng[1].tolist() == nonce.tolist()
    gm.set_bgv_schedule(log_n, 0, -1)
    dec = gm.bgv_decrypt(log_n, n, gm.upload(s), c0, c1).to_numpy(n * N)
    assert np.array_equal(dec, msgs)


def test_empty_inputs_are_no_ops(ctx):
    """n = 0 / count = 0 through every batched entry point: no launch, no error, outputs untouched"""
    gm, om = mods(ctx, Q128A)
    say("empty inputs")
    buf = gm.upload(om.sample_uniform(KEY, 0, 16))
    before = gm.download(buf, 16)
    gm.ntt(12, 0, buf, out=buf)
    gm.polymul(14, 0, buf, buf, out=buf)
    gm.pointwise(gm.ADD, 0, buf, buf, out=buf)
    gm.ciphertext_op(gm.ADD, 0, buf, buf, buf, buf, b_count=1, out0=buf, out1=buf)
    gm.fma(0, buf, buf, buf, out=buf)
    gm.matvec(0, 4, 4, buf, buf, out=buf)
    ctx.cipher_encrypt(4, 0, KEY, [0, 0, 0], buf, out=ctx.buffer(64))
    pm = oracle.int_to_limbs(om.from_unsigned(1 << 32), gm.limbs)
    nonce = ctx.rng[1].tolist()
    gm.bgv_encrypt(14, 0, buf, buf, buf, pm, c0=buf, c1=buf)
    assert ctx.rng[1].tolist() == nonce
    ctx.wait()
    assert np.array_equal(gm.download(buf, 16), before)


def test_errors_are_reported(ctx):
    say("error paths")
    with pytest.raises(hmpc_amd.HmpcError):
        hmpc_amd.Modulus(ctx, 1 << 70)  # even
    gm, _ = mods(ctx, Q128B)
    with pytest.raises(hmpc_amd.HmpcError):
        gm.prepare(16)  # 2N does not divide q - 1
    with pytest.raises(hmpc_amd.HmpcError):
        gm.prepare(1)


def _edge_values(q, limbs, count, seed):
    rng = np.random.default_rng(seed)
    R = 1 << (32 * limbs)
    edges = [0, 1, 2, q - 1, q - 2, (q - 1) // 2, (q + 1) // 2, R % q, (R - 1) % q, (1 << (32 * limbs - 1)) % q]
    edges += [((1 << (32 * k)) - 1) % q for k in range(1, limbs + 1)]
    vals = [int.from_bytes(rng.bytes(4 * limbs), "little") % q for _ in range(count - len(edges))]
    return edges + vals


@pytest.mark.parametrize("q", [Q64, Q128A, Q128B, P150, Q279, 1601, 0x8822D80623320001, Q_GENERIC4, Q_LAZY2])
def test_pointwise_ops_edge_values(ctx, q):
    gm, om = mods(ctx, q)
    say(f"pointwise edges q={hex(q)[:12]}")
    count = 20000
    a_vals = _edge_values(q, gm.limbs, count, 1)
    b_vals = _edge_values(q, gm.limbs, count, 2)[::-1]
    # every edge against every edge in the first block
    k = 20
    a = oracle.ints_to_planar(a_vals, gm.limbs)
    b = oracle.ints_to_planar(b_vals, gm.limbs)
    da, db = gm.upload(a), gm.upload(b)
    for op, name in ((gm.ADD, "add"), (gm.SUB, "sub"), (gm.MUL, "mul")):
        got = gm.download(gm.pointwise(op, count, da, db), count)
        assert np.array_equal(got, om.pointwise(name, a, b)), name
    # broadcast of a short operand (unsqueeze) and scalar constants
    bk = np.ascontiguousarray(b[:, :k])
    got = gm.download(gm.pointwise(gm.MUL, count, da, gm.upload(bk), b_count=k), count)
    assert np.array_equal(got, om.pointwise("mul", a, bk)), "broadcast"
    for c in (0, 1, q - 1, a_vals[5000]):
        cw = oracle.int_to_limbs(c, gm.limbs)
        got = gm.download(gm.pointwise_scalar(gm.MUL, count, da, cw), count)
        want = om.pointwise("mul", a, oracle.ints_to_planar([c], gm.limbs))
        assert np.array_equal(got, want), hex(c)


def _structured_pairs(q, limbs, N):
    """pairs of inputs at the value extremes: constants q - 1 and (q - 1) / 2, alternating q - 1 / 0
    and q - 1 / 1, an impulse, the edge values of _edge_values cycled and reversed, and random
    values against q - 1"""
    def const(v):
        return [v] * N

    def alt(v, w):
        return [v if i % 2 == 0 else w for i in range(N)]

    rng = np.random.default_rng(11)
    edges = _edge_values(q, limbs, 64, 3)
    cyc = [edges[i % len(edges)] for i in range(N)]
    rnd = [int.from_bytes(rng.bytes(4 * limbs), "little") % q for _ in range(N)]
    impulse = [1] + [0] * (N - 1)
    return [(const(q - 1), const(q - 1)), (alt(q - 1, 0), const(1)), (const((q - 1) // 2), alt(q - 1, 1)),
            (cyc, cyc[::-1]), (impulse, const(q - 1)), (rnd, const(q - 1))]


# the radix-2^26 lengths whose passes take expanded twiddles (streaming columns at N1 = 2^7, the
# one-wave rows of 2^9 at N = 2^16) and their neighbours
@pytest.mark.parametrize("q,log_n", [(Q128A, 16), (Q128A, 17), (Q128A, 15), (Q128B, 14)],
                         ids=lambda v: hex(v)[:10] if isinstance(v, int) and v > 99 else str(v))
def test_polymul_structured_inputs(ctx, q, log_n):
    """Products of inputs at the value extremes through the default fast schedule, bit for bit
    against the oracle, and the forward transform of the same inputs."""
    gm, om = mods(ctx, q)
    N = 1 << log_n
    pairs = _structured_pairs(q, gm.limbs, N)
    n = len(pairs)
    say(f"polymul structured q={hex(q)[:12]} log_n={log_n} n={n}")
    a = oracle.ints_to_planar([v for x, _ in pairs for v in x], gm.limbs)
    b = oracle.ints_to_planar([v for _, y in pairs for v in y], gm.limbs)
    want, _ = om.polymul_timed(log_n, a, b, threads=8)
    got = gm.download(gm.polymul(log_n, n, gm.upload(a), gm.upload(b)), n * N)
    assert np.array_equal(got, want)
    fwd = gm.download(gm.ntt(log_n, n, gm.upload(a)), n * N)
    assert np.array_equal(fwd, om.ntt(a, log_n, threads=8))


FAST_CASES = [(Q128A, 4, 5), (Q128A, 12, 3), (Q128A, 13, 2), (Q128A, 15, 2), (Q128A, 16, 2), (Q128A, 17, 1), (Q128B, 14, 2),
              (Q64, 3, 9), (Q64, 12, 3), (Q64, 14, 2), (Q64, 15, 2), (Q64, 16, 1)]


@pytest.mark.parametrize("q,log_n,n", FAST_CASES, ids=lambda v: hex(v)[:10] if isinstance(v, int) and v > 99 else str(v))
def test_fast_kernel_schedules(ctx, q, log_n, n):
    """Every elements-per-thread schedule of the fast kernels gives hmpc's results."""
    gm, om = mods(ctx, q)
    assert gm.schedule(log_n)["fast"]
    count = n << log_n
    a = om.sample_uniform(KEY, 0, count, threads=8)
    b = om.sample_uniform(KEY, 1, count, threads=8)
    want_f = om.ntt(a, log_n, threads=8)
    want_i = om.ntt(a, log_n, inverse=True, threads=8)
    want_p, _ = om.polymul_timed(log_n, a, b, threads=8)
    da, db = gm.upload(a), gm.upload(b)
    for elems in ((4, 8) if gm.limbs == 4 else (8, 16)):
        for persistent in (1, 0):
            say(f"fast schedule q={hex(q)[:12]} log_n={log_n} n={n} elems={elems} persistent={persistent}")
            gm.set_schedule(log_n, elems, persistent)
            assert gm.schedule(log_n)["persistent"] == bool(persistent)
            assert np.array_equal(gm.download(gm.ntt(log_n, n, da), count), want_f), f"forward elems={elems}"
            assert np.array_equal(gm.download(gm.ntt(log_n, n, da, inverse=True), count), want_i), f"inverse elems={elems}"
            got = gm.download(gm.polymul(log_n, n, da, db), count)
            assert np.array_equal(got, want_p), f"polymul elems={elems} persistent={persistent}"
    gm.set_schedule(log_n, 0, -2)


def test_autotune_bgv_keeps_results_and_rng(ctx):
    gm, om = mods(ctx, Q128B)
    log_n, n = 14, 4
    say(f"autotune bgv q={hex(Q128B)[:12]} log_n={log_n} n={n}")
    enc_key = [42, 0, 0, 0, 0, 0, 0, 1]
    ctx.set_rng(enc_key, [5, 0, 0])
    best = gm.autotune_bgv(log_n, n)
    assert best["elems_per_thread"] in (4, 8) and best["ms"] > 0 and best["tiles"] in (True, False)
    assert ctx.rng[1].tolist() == [5, 0, 0]
    N = 1 << log_n
    t_mont = om.from_unsigned(1 << 32)
    A, B, _ = _keygen(om, log_n, t_mont)
    msgs = oracle.sample_words(KEY, 3, n * N)
    want0, want1, _ = om.bgv_encrypt(log_n, A, B, msgs, t_mont, enc_key, [5, 0, 0], threads=8)
    c0, c1 = gm.bgv_encrypt(log_n, n, gm.upload(A), gm.upload(B), ctx.upload(msgs), oracle.int_to_limbs(t_mont, gm.limbs))
    assert np.array_equal(gm.download(c0, n * N), want0) and np.array_equal(gm.download(c1, n * N), want1)


def test_default_schedule_is_one_launch_per_pass(ctx):
    """Untuned products run one launch per pass (the faster structure at every measured degree) at 8
    elements per thread for 4 limbs and 16 for 2."""
    for q, elems in ((Q128A, 8), (Q64, 16)):
        gm, _ = mods(ctx, q)
        say(f"default schedule q={hex(q)[:12]}")
        gm.set_schedule(16, 0, -2)
        sched = gm.schedule(16)
        assert sched["fast"] and not sched["persistent"] and sched["elems_per_thread"] == elems


def test_autotune_installs_a_correct_schedule(ctx):
    gm, om = mods(ctx, Q128A)
    log_n, n = 14, 8
    say(f"autotune q={hex(Q128A)[:12]} log_n={log_n} pairs={n}")
    best = gm.autotune(log_n, n)
    sched = gm.schedule(log_n)
    assert sched["elems_per_thread"] == best["elems_per_thread"] and sched["persistent"] == best["persistent"]
    assert best["ms"] > 0
    count = n << log_n
    a = om.sample_uniform(KEY, 0, count, threads=8)
    b = om.sample_uniform(KEY, 1, count, threads=8)
    want, _ = om.polymul_timed(log_n, a, b, threads=8)
    assert np.array_equal(gm.download(gm.polymul(log_n, n, gm.upload(a), gm.upload(b)), count), want)
    gm.set_schedule(log_n, 0, -2)


# uneven job list: products and encryptions, N from 2^12 to 2^16, both fast moduli
BATCH_JOBS = [("mul", Q64, 12), ("bgv", Q128A, 13), ("mul", Q128A, 16), ("bgv", Q64, 14), ("mul", Q128A, 12),
              ("bgv", Q128A, 16), ("mul", Q64, 15), ("bgv", Q64, 12), ("mul", Q128A, 14), ("mul", Q64, 16),
              ("bgv", Q128A, 12), ("mul", Q128A, 13), ("bgv", Q128A, 14), ("bgv", Q64, 16), ("bgv", Q128A, 13),
              ("mul", Q128A, 15), ("bgv", Q128A, 15), ("mul", Q128A, 16), ("mul", Q128A, 17), ("bgv", Q128A, 17)]


# the same without the jobs of 2^17: the q128a chunks' largest degree is then 2^16, so their row
# launches run one wave per workgroup over mixed row lengths (with 2^17 jobs they run eight)
BATCH_JOBS_16 = [j for j in BATCH_JOBS if j[2] <= 16]


@pytest.mark.parametrize("structure,jobs", [("tiles", "to17"), ("jobs", "to17"), ("tiles", "to16")])
def test_batch_of_uneven_jobs_matches_oracle(ctx, structure, jobs):
    """Persistent job kernels (one per modulus class) and, for jobs at N > 4096, tile launches
    over job tables (or, when so scheduled, the persistent kernel) over an uneven job list, eager
    and replayed from a hipGraph three times with host work between replays; every job's output
    equals the oracle's, and each encryption consumes the nonces a queue call would."""
    job_list = BATCH_JOBS if jobs == "to17" else BATCH_JOBS_16
    two_pass = sorted({(kind, q, log_n) for kind, q, log_n in job_list if log_n > 12})
    if structure == "jobs":
        for kind, q, log_n in two_pass:
            gm = mods(ctx, q)[0]
            if kind == "bgv":
                gm.set_bgv_schedule(log_n, 0, 0)
            else:
                gm.set_schedule(log_n, 0, 1)
    try:
        info = _run_uneven_batch(ctx, job_list)
        tiled = len({(kind, q) for kind, q, _ in two_pass})
        assert info["classes"] >= (tiled if structure == "tiles" else 2)
    finally:
        for kind, q, log_n in two_pass:
            gm = mods(ctx, q)[0]
            if kind == "bgv":
                gm.set_bgv_schedule(log_n, 0, -1)
            else:
                gm.set_schedule(log_n, 0, -2)


def _run_uneven_batch(ctx, job_list):
    enc_key = [42, 0, 0, 0, 0, 0, 0, 1]
    ctx.set_rng(enc_key, [0, 0, 0])
    batch = hmpc_amd.Batch(ctx)
    checks = []
    nonce = np.zeros(3, np.uint32)
    for i, (kind, q, log_n) in enumerate(job_list):
        gm, om = mods(ctx, q)
        N = 1 << log_n
        if kind == "mul":
            a = om.sample_uniform(KEY, 10 + 2 * i, N)
            b = om.sample_uniform(KEY, 11 + 2 * i, N)
            out = gm.alloc(N)
            batch.add_polymul(gm, log_n, gm.upload(a), gm.upload(b), out)
            want, _ = om.polymul_timed(log_n, a, b, threads=8)
            checks.append((gm, N, [(out, want)]))
        else:
            t_mont = om.from_unsigned(1 << 32)
            A, B, _ = _keygen(om, log_n, t_mont)
            msgs = oracle.sample_words(KEY, 100 + i, N)
            c0, c1 = gm.alloc(N), gm.alloc(N)
            batch.add_bgv_encrypt(gm, log_n, gm.upload(A), gm.upload(B), ctx.upload(msgs),
                                  oracle.int_to_limbs(t_mont, gm.limbs), c0, c1)
            want0, want1, nonce = om.bgv_encrypt(log_n, A, B, msgs, t_mont, enc_key, nonce, threads=8)
            checks.append((gm, N, [(c0, want0), (c1, want1)]))
    assert ctx.rng[1].tolist() == list(nonce)
    for run, graph in enumerate((False, True, True, True)):
        say(f"batch run={run} graph={graph} jobs={len(job_list)}")
        for gm, N, outs in checks:
            for buf, _ in outs:  # eager work between replays: clear every output
                gm.pointwise_scalar(gm.MUL, N, buf, np.zeros(gm.limbs, np.uint32), out=buf)
        ctx.wait()
        batch.run(graph=graph)
        ctx.wait()
        for k, (gm, N, outs) in enumerate(checks):
            for buf, want in outs:
                assert np.array_equal(gm.download(buf, N), want), f"job {k} {job_list[k]} run {run}"
    info = batch.info()
    assert info["jobs"] == len(job_list) and info["classes"] >= 2
    return info
