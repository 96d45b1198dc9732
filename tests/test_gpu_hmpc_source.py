"""GPU test of hmpc's own spelling of the path (include/hmpc_amd/hmpc.hpp): the program
tests/cpp/hmpc_source_test, built by hmpc_amd.build, runs three workloads written with hmpc's names
and types (compile-time moduli, poly_mod, make_tensor, host_accessor, planned expressions with a
shared capability sampler, lhe enc / ciphertext operators / dec<plaintext>) with the parameters of
hmpc's own code; every tensor it writes is recomputed here with the oracle, bit for bit:

  linear BGV    tests/expr/crypto/lhe/enc.cpp:15-72 (q of 279 bits, p = 0x8822d80623320001,
                N = 2^16, shape {4, 2}, a = b = raw words i, s = one, x = raw words i)
  bgv example   examples/bgv.cpp:25-90 with n = 4 (150-bit q, p = 0x89d046e0ae640001 as a constant,
                u shared by c0 and c1)
  ntt example   examples/ntt.cpp:22-55 with n = 2 (125-bit p, N = 2^20)

The program also asserts, on the planner, the execution-cache sizes of hmpc's cache test
(tests/expr/cache.cpp:13-57) and that INTT(NTT(x)) and dec(enc(x)) return their inputs; it exits
non-zero when any of those fails.
"""

import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from bigref import P150

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "hmpc_source_test")
KEY = [42, 0, 0, 0, 0, 0, 0, 1]  # the program's queue key
Q279 = 0x591F5B834C0D961F67343BCC8902BDEDA2771F54306FF151162FF8D2B40F4194DC0001
P64_ENC = 0x8822D80623320001
P64_EXAMPLE = 0x89D046E0AE640001
P125 = 0x18F25CD9A75CCBD9C146D4ABAEC00001


def _read(path):
    out = []
    with open(path, "rb") as f:
        while True:
            h = f.read(8)
            if not h:
                return out
            n = int(np.frombuffer(h, np.uint64)[0])
            out.append(np.frombuffer(f.read(4 * n), np.uint32))


def _iota_planar(count, limbs):
    out = np.zeros((limbs, count), np.uint32)
    out[0] = np.arange(count, dtype=np.uint32)
    return out


def _unsigned_planar(om, q, values):
    R = 1 << (32 * om.limbs)
    return oracle.ints_to_planar([v * R % q for v in values], om.limbs)


def test_hmpc_spelling_matches_oracle(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/hmpc_source_test is missing: run hmpc_amd.build first")
    out = tmp_path / "hmpc_source.bin"
    print("hmpc source program", flush=True)
    sys.stdout.flush()
    r = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    (c0, c1, s0, s1, t0, t1, y, na, nb, e0, e1, ny) = _read(out)

    # linear BGV: enc, ct + ct, 2 * ct, dec<plaintext>
    log_n, n = 16, 8
    N = 1 << log_n
    oq, op = oracle.Modulus(Q279), oracle.Modulus(P64_ENC)
    a = _iota_planar(N, oq.limbs)
    x = _iota_planar(n * N, op.limbs)
    want0, want1, _ = oq.bgv_encrypt_plaintext(op, log_n, a, a, x, True, KEY, [0, 0, 0], threads=8)
    assert np.array_equal(c0, want0.reshape(-1)) and np.array_equal(c1, want1.reshape(-1))
    assert np.array_equal(s0, oq.pointwise("add", want0, want0).reshape(-1))
    assert np.array_equal(s1, oq.pointwise("add", want1, want1).reshape(-1))
    two = oracle.ints_to_planar([oq.from_unsigned(2)] * (n * N), oq.limbs)
    assert np.array_equal(t0, oq.pointwise("mul", want0, two).reshape(-1))
    assert np.array_equal(t1, oq.pointwise("mul", want1, two).reshape(-1))
    assert np.array_equal(y, x.reshape(-1))

    # bgv example: c0 = b u + v p + m, c1 = a u + w p with one u on the queue's first nonce
    log_n, n, argc = 14, 4, 1
    N = 1 << log_n
    om = oracle.Modulus(P150)
    A_in = _unsigned_planar(om, P150, [argc + n + i for i in range(N)])
    B_in = _unsigned_planar(om, P150, [argc + N - i for i in range(N)])
    msgs = np.array([(argc + n - i) % (1 << 32) for i in range(n * N)], dtype=np.uint32)
    ntt_a, ntt_b = om.ntt(A_in, log_n), om.ntt(B_in, log_n)
    assert np.array_equal(na, ntt_a.reshape(-1)) and np.array_equal(nb, ntt_b.reshape(-1))
    want0, want1, _ = om.bgv_encrypt(log_n, ntt_b, ntt_a, msgs, om.from_unsigned(P64_EXAMPLE), KEY, [0, 0, 0], threads=8)
    assert np.array_equal(e0, want0.reshape(-1)) and np.array_equal(e1, want1.reshape(-1))

    # ntt example: NTT of x_i = argc + N n - i
    log_n, n = 20, 2
    N = 1 << log_n
    op = oracle.Modulus(P125)
    xs = _unsigned_planar(op, P125, [argc + N * n - i for i in range(N * n)])
    assert np.array_equal(ny, op.ntt(xs, log_n, threads=8).reshape(-1))
