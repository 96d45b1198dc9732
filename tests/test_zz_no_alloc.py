"""GPU test of the allocation-free entry point hmpc_polymul_noalloc through the C ABI: the same results
as hmpc_polymul on two calls in a row (the second with fewer pairs, over the workspace the first left),
for the radix-2^26, Goldilocks and portable kernels; oversized workspaces and unprepared degrees are
refused with an error."""

import sys

import numpy as np
import pytest

import hmpc_amd
import oracle

pytestmark = pytest.mark.gpu

KEY = [42, 0, 0, 0, 0, 0, 0, 0]
Q_GENERIC4 = 0x80CD613ED8F16ADF91B7584A226B8001  # portable kernels (no special form)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = hmpc_amd.Context(0, key=KEY)
    yield c
    c.wait()


@pytest.mark.parametrize("q,log_n,n_first,n_second", [(hmpc_amd.Q128A, 16, 6, 2), (hmpc_amd.Q64, 15, 20, 3),
                                                      (Q_GENERIC4, 14, 8, 1), (hmpc_amd.Q128A, 12, 9, 4)],
                         ids=["q128a-2^16", "q64-2^15", "generic-2^14", "q128a-4096"])
def test_polymul_noalloc_matches_polymul(ctx, q, log_n, n_first, n_second):
    m = hmpc_amd.Modulus(ctx, q)
    m.prepare(log_n)
    for n in (n_first, n_second):
        say(f"polymul_noalloc q={hex(q)[:12]} log_n={log_n} n={n}")
        count = n << log_n
        a = m.sample_uniform(KEY, 2 * n, count)
        b = m.sample_uniform(KEY, 2 * n + 1, count)
        want = m.download(m.polymul(log_n, n, a, b), count)
        out = m.alloc(count)
        m.polymul_noalloc(log_n, n, a, b, out)
        got = m.download(out, count)
        assert np.array_equal(got, want)
    if log_n == 16:
        # the two calls match the oracle too
        om = oracle.Modulus(q)
        pa = om.sample_uniform(KEY, 2 * n_second, n_second << log_n, threads=8)
        pb = om.sample_uniform(KEY, 2 * n_second + 1, n_second << log_n, threads=8)
        assert np.array_equal(got, om.polymul_timed(log_n, pa, pb, threads=8)[0])


def test_polymul_noalloc_refuses_what_it_cannot_do(ctx):
    m = hmpc_amd.Modulus(ctx, hmpc_amd.Q128A)
    say("polymul_noalloc refusals")
    m.prepare(16)
    big = 64  # 2 x 5 x 64 x 2^16 words = 160 MiB of workspace
    buf = m.alloc(big << 16)
    with pytest.raises(hmpc_amd.HmpcError):
        m.polymul_noalloc(16, big, buf, buf, buf)
    small = m.alloc(1 << 13)
    with pytest.raises(hmpc_amd.HmpcError):
        m.polymul_noalloc(13, 1, small, small, small)  # tables of N = 2^13 not prepared
