"""CPU checks of the radix-2^26 device arithmetic (hmpc_amd/csrc/r26.hpp) built for the host.

The same header is compiled into the gfx950 kernels; here it runs on the CPU against python
big integers, including the lazy bounds the fast transform relies on: forward stages work on
signed digits and add less than 2q (1 + 2^-18) in magnitude and at most one digit bound per limb,
so 29 stages from inputs in [0, q) keep every mont input within 64 q with limbs below 2^31 in
magnitude; inverse stages stay below 2q.
"""

import os
import random
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "r26_check.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "r26_check")

PRIMES = [(1 << 124) + (1 << 20) + (1 << 18) + 1, 0x10000000000000000000000000078001]
M26 = (1 << 26) - 1


@pytest.fixture(scope="module")
def tool():
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < os.path.getmtime(SRC) or \
            os.path.getmtime(BIN) < os.path.getmtime(os.path.join(ROOT, "hmpc_amd", "csrc", "r26.hpp")):
        subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", SRC, "-o", BIN],
                       check=True)
    p = subprocess.Popen([BIN], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
    yield p
    p.stdin.close()
    p.wait()


def run(p, line):
    p.stdin.write(line + "\n")
    p.stdin.flush()
    return [int(v) for v in p.stdout.readline().split()]


def limbs(v, top_bits=32):
    out = [(v >> (26 * j)) & M26 for j in range(4)]
    out.append(v >> 104)
    return out


def value(l):
    return sum(int(x) << (26 * j) for j, x in enumerate(l))


def redundant(v, rng, max_bits):
    """a random redundant representation of v with limbs < 2^max_bits (v small enough)"""
    l = limbs(v)
    # move random multiples of 2^26 down from higher limbs
    for j in range(4, 0, -1):
        room = (1 << max_bits) - 1 - l[j - 1]
        k = min(l[j], room >> 26)
        k = rng.randrange(k + 1) if k else 0
        l[j] -= k
        l[j - 1] += k << 26
    assert value(l) == v and all(0 <= x < (1 << 32) for x in l)
    return l


def signed_redundant(v, rng, max_abs):
    """a random signed-digit representation of v (any sign): limbs |x_j| < max_abs"""
    l = [(v >> (26 * j)) & M26 for j in range(4)] + [v >> 104]
    for j in range(1, 5):
        # borrow k from limb j into limb j - 1 (k < 0 lends the other way), keeping the bound
        for _ in range(8):
            k = rng.randrange(-15, 16)
            a, b = l[j - 1] + (k << 26), l[j] - k
            if abs(a) < max_abs and abs(b) < max_abs:
                l[j - 1], l[j] = a, b
                break
    assert value(l) == v
    return l


@pytest.mark.parametrize("q", PRIMES)
def test_mont_fold_canonical(tool, q):
    c = q - (1 << 124)
    rng = random.Random(q & 0xFFFF)
    R = 1 << 130
    for it in range(600):
        b = rng.randrange(q)
        lb = limbs(b)
        if it % 2 == 0:
            a = rng.randrange(50 * q)
            la = redundant(a, rng, 31)
        else:  # a signed-digit operand, as the forward transform's values
            a = rng.randrange(-50 * q, 50 * q)
            la = signed_redundant(a, rng, 1 << 31)
        r = run(tool, f"M {c} " + " ".join(map(str, la + lb)))
        v = value(r)
        assert v % q == a * b * pow(R, -1, q) % q
        if min(la) >= 0:  # the unsigned product (inverse transforms) gives the same limbs
            assert run(tool, f"U {c} " + " ".join(map(str, la + lb))) == r
        # quotient digits m_0..m_3 are not masked: bound 2q (1 + 2^-19); digits 0..3 in [0, 2^26)
        assert all(0 <= x < (1 << 26) for x in r[:4]) and abs(r[4]) < (1 << 22)
        if a >= 0:
            assert 0 <= v < 2 * q + (q >> 18)
        else:
            assert -q < v < 2 * q + (q >> 18)
        # the signed fold of a signed-digit value
        y = rng.randrange(-63 * q, 63 * q)
        ly = signed_redundant(y, rng, (1 << 31) - (1 << 26))
        if abs(ly[4]) < (1 << 29):
            g = run(tool, f"G {c} " + " ".join(map(str, ly)))
            assert value(g) % q == y % q and 0 <= value(g) < 2 * q and all(0 <= v < (1 << 28) for v in g)
            w = run(tool, f"C {c} " + " ".join(map(str, g)))
            assert sum(v << (32 * i) for i, v in enumerate(w)) == y % q
        x = rng.randrange(1 << 135)
        lx = redundant(x, rng, 31) if x < (1 << 133) else limbs(x)
        if max(lx) >= (1 << 31):
            continue
        f = run(tool, f"F {c} " + " ".join(map(str, lx)))
        assert value(f) % q == x % q and value(f) < 2 * q and all(v < (1 << 28) for v in f)
        w = run(tool, f"C {c} " + " ".join(map(str, f)))
        assert sum(v << (32 * i) for i, v in enumerate(w)) == x % q


@pytest.mark.parametrize("q", PRIMES)
def test_from_signed(tool, q):
    """z -> z R mod q (R = 2^128) without a product: exact, value < q, limbs within the forward
    transform's input bounds; edge values of the signed 64-bit range included."""
    c = q - (1 << 124)
    rng = random.Random(5)
    zs = [0, 1, -1, 2, -2, (1 << 63) - 1, -(1 << 63), (1 << 45) + 12345, -(1 << 45) - 7]
    zs += [rng.randrange(-(1 << 63), 1 << 63) for _ in range(300)] + [rng.randrange(-(1 << 46), 1 << 46) for _ in range(300)]
    for z in zs:
        x = run(tool, f"S {c} {z}")
        v = value(x)
        assert v % q == z * (1 << 128) % q, z
        assert v <= q and x[0] < (1 << 27) and all(l < (1 << 26) for l in x[1:4]) and x[4] < (1 << 20)


LIMB_IN = (1 << 27, 1 << 26, 1 << 26, 1 << 26, 1 << 22)    # forward input limbs (and from_signed's)
LIMB_STEP = (1 << 26, 1 << 26, 1 << 26, 1 << 26, 1 << 22)  # per-stage limb growth bound (digits of t)


def _forward_input(q, rng):
    """a forward-transform input: canonical limbs of v < q, or from_signed's shape q - t (limbs
    qb - t, t < 2^92)"""
    if rng.random() < 0.5:
        return limbs(rng.randrange(q))
    c = q - (1 << 124)
    qb = [c + (1 << 26), (1 << 26) - 1, (1 << 26) - 1, (1 << 26) - 1, (1 << 20) - 1]
    t = limbs(rng.randrange(1 << 92))
    return [k - x for k, x in zip(qb, t)]


@pytest.mark.parametrize("q", PRIMES)
@pytest.mark.parametrize("stages", [12, 17, 29])
def test_forward_chain_bounds(tool, q, stages):
    """Lazy forward stages on signed digits from inputs in [0, q) (canonical or from_signed-shaped),
    each stage keeping either output pair or a doubled output: after s stages every value is
    within q + 2q (1 + 2^-18) s in magnitude with limbs below LIMB_IN + s x LIMB_STEP in
    magnitude, so every mont input is within 64 q with limbs below 2^31; the results are exact
    mod q, and the last ones fold (signed) to [0, 2q) with limbs < 2^28."""
    c = q - (1 << 124)
    rng = random.Random(7 + stages)
    R = 1 << 130
    for it in range(150):
        lx, ly = _forward_input(q, rng), _forward_input(q, rng)
        ws = [rng.randrange(q) if rng.random() < 0.8 else rng.randrange(4) for _ in range(stages)]
        sels = [1 if it % 3 == 0 else rng.randrange(3) for _ in range(stages)]
        line = f"B {c} {stages} " + " ".join(map(str, lx + ly)) + " " + \
               " ".join(" ".join(map(str, limbs(w))) + f" {sel}" for w, sel in zip(ws, sels))
        out = run(tool, line)
        assert len(out) == 10 * stages
        ex, ey = value(lx) % q, value(ly) % q
        prev_y = ly
        for s, (w, sel) in enumerate(zip(ws, sels)):
            assert abs(value(prev_y)) < 64 * q and max(abs(v) for v in prev_y) < (1 << 31), "mont input"
            t = ey * w * pow(R, -1, q) % q
            ex, ey = (ex + t) % q, (ex - t) % q
            if sel == 1:
                ex = ey
            elif sel == 2:
                ey = ex
            st = out[10 * s:10 * s + 10]
            X, Y = value(st[:5]), value(st[5:])
            assert X % q == ex and Y % q == ey, f"stage {s}"
            for v in (X, Y):
                assert abs(v) < q + (2 * q + (q >> 17)) * (s + 1)
            for half in (st[:5], st[5:]):
                assert all(abs(half[j]) < LIMB_IN[j] + (s + 1) * LIMB_STEP[j] for j in range(5)), f"limbs at stage {s}"
            prev_y = st[5:]
        # a forward output folds (point-wise products, stores) to [0, 2q) with limbs < 2^28
        for half in (out[-10:-5], out[-5:]):
            f = run(tool, f"G {c} " + " ".join(map(str, half)))
            assert value(f) % q == value(half) % q and 0 <= value(f) < 2 * q and all(0 <= v < (1 << 28) for v in f)


def _forward_output(q, rng, extreme):
    """a forward-transform output after 17 stages: |value| < 35.01 q, |limbs| < 19 2^26 (limb 4:
    < 2^22 + 17 2^22); extreme: magnitudes near those bounds"""
    bound = q + (2 * q + (q >> 17)) * 17
    v = (bound - 1 - rng.randrange(q >> 8)) * rng.choice((-1, 1)) if extreme else rng.randrange(-bound + 1, bound)
    for _ in range(200):
        l = signed_redundant(v, rng, 19 << 26)
        if all(abs(l[j]) < LIMB_IN[j] + 17 * LIMB_STEP[j] for j in range(5)):
            return l
    return None


def _product(tool, q, rng, extreme=False):
    """a point-wise product of two forward outputs through r26_point, with its exact value checked"""
    c = q - (1 << 124)
    while True:
        la, lb = _forward_output(q, rng, extreme), _forward_output(q, rng, extreme)
        if la is not None and lb is not None:
            break
    r = run(tool, f"P {c} " + " ".join(map(str, la + lb)))
    assert value(r) % q == value(la) * value(lb) * pow(1 << 128, -1, q) % q
    return r


@pytest.mark.parametrize("q", PRIMES)
def test_point_product_bounds(tool, q):
    """r26_point on forward outputs of up to 17 stages (b folded, a the signed operand): exact
    a b R^-1 mod q, within (-4.4q, 8.4q), digits 0..3 in [0, 2^28) and a signed top limb below
    2^24 in magnitude; operands near the magnitude bounds included."""
    rng = random.Random(q & 0xFFF)
    for it in range(400):
        r = _product(tool, q, rng, extreme=it % 2 == 0)
        assert -44 * q < 10 * value(r) < 84 * q
        assert all(0 <= v < (1 << 28) for v in r[:4]) and abs(r[4]) < (1 << 24)


@pytest.mark.parametrize("q", PRIMES)
@pytest.mark.parametrize("stages", [12, 16, 17])
def test_inverse_chain_bounds(tool, q, stages):
    """Lazy inverse stages on signed digits, x + y folded at stages 1, 4, 7, ... only (the kernels'
    schedule, r26_inverse_fold_stage), from point-wise products (r26_point of forward outputs:
    within (-4.4q, 8.4q), digits < 2^28), each stage keeping either output pair or a doubled output:
    every mont input x - y stays within 33.6q with limbs below 2^30, fold-stage inputs within 8.4q
    (1 + 2^-16) (16.8q at stage 1) with limbs below 2^29.2 (2^30 at stage 1), the folded outputs in
    [0, 2q) with limbs < 2^28, the mont outputs within (-q, 2q (1 + 2^-19)), and the results are
    exact mod q."""
    c = q - (1 << 124)
    rng = random.Random(11 + stages)
    R = 1 << 130
    for it in range(150):
        lx, ly = _product(tool, q, rng, extreme=it % 2 == 0), _product(tool, q, rng, extreme=it % 3 == 0)
        x, y = value(lx), value(ly)
        ws = [rng.randrange(q) for _ in range(stages)]
        sels = [1 if it % 3 == 0 else rng.randrange(3) for _ in range(stages)]
        folds = [int(s % 3 == 1) for s in range(stages)]
        line = f"I {c} {stages} " + " ".join(map(str, lx + ly)) + " " + \
               " ".join(" ".join(map(str, limbs(w))) + f" {f} {sel}" for w, f, sel in zip(ws, folds, sels))
        out = run(tool, line)
        assert len(out) == 10 * stages
        ex, ey = x % q, y % q
        px, py = lx, ly
        for s, (w, f, sel) in enumerate(zip(ws, folds, sels)):
            d = [a - b for a, b in zip(px, py)]
            assert 10 * abs(value(d)) < 336 * q and max(abs(v) for v in d) < (1 << 30), f"mont input at stage {s}"
            if f:  # fold stage: inputs are sums of two stages since the last fold (of products at stage 1)
                bound = 168 * q // 10 if s == 1 else 84 * q // 10 + (q >> 16)
                limb = (1 << 30) if s == 1 else int(2 ** 29.2)
                assert abs(value(px)) < bound and abs(value(py)) < bound and max(abs(v) for v in px + py) < limb
            ex, ey = (ex + ey) % q, (ex - ey) * w * pow(R, -1, q) % q
            if sel == 1:
                ey = ex
            elif sel == 2:
                ex = ey
            st = out[10 * s:10 * s + 10]
            X, Y = value(st[:5]), value(st[5:])
            assert X % q == ex and Y % q == ey, f"stage {s}"
            if f and sel != 2:
                assert 0 <= X < 2 * q and all(0 <= v < (1 << 28) for v in st[:5])
            if sel != 1:
                assert -q < Y < 2 * q + (q >> 18) and all(0 <= v < (1 << 26) for v in st[5:9])
            px, py = st[:5], st[5:]
