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
        # built under a name of this process and renamed, so that parallel workers do not race
        tmp = f"{BIN}.{os.getpid()}"
        subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", SRC, "-o", tmp],
                       check=True)
        os.replace(tmp, BIN)
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


# ---- products by a twiddle in expanded form (r26_mont_tw) and the streaming column schedule

def _expanded(q, w):
    """the expanded twiddle of w from its closed form: V_i = 2^(26 (i + 1)) w mod q, canonical limbs"""
    return [limbs(w * (1 << (26 * (i + 1))) % q) for i in range(5)]


@pytest.mark.parametrize("q", PRIMES)
def test_expanded_twiddle_and_mont_tw(tool, q):
    """r26_expand_twiddle from a stored entry W = w 2^130 mod q against the closed form
    V_i = 2^(26 (i + 1)) w mod q, and r26_mont_tw against a w mod q for signed-digit a with limbs
    up to |a_j| < 2^31 (extreme limbs included): digits 0..3 in [0, 2^26) and
    |r_4| < 5 A / 2^6 + 2^20 + 2^8 (A = max |a_j|)."""
    c = q - (1 << 124)
    rng = random.Random(q & 0xFFFFF)
    for it in range(800):
        w = rng.randrange(q) if it % 8 else (0, 1, 2, q - 1)[(it // 8) % 4]
        W = limbs(w * (1 << 130) % q)
        V = run(tool, f"X {c} " + " ".join(map(str, W)))
        assert [V[5 * i:5 * i + 5] for i in range(5)] == _expanded(q, w)
        A = (1 << 26, 1 << 29, 1 << 30, (1 << 31) - 1)[it % 4]
        a = [rng.choice((-A + 1, A - 1)) if rng.random() < 0.3 else rng.randrange(-A + 1, A) for _ in range(5)]
        r = run(tool, f"T {c} " + " ".join(map(str, a + W)))
        assert value(r) % q == value(a) * w % q
        assert all(0 <= v < (1 << 26) for v in r[:4])
        assert abs(r[4]) < 5 * max(abs(v) for v in a) / 64 + (1 << 20) + (1 << 8)


@pytest.mark.parametrize("q", PRIMES)
def test_fold_signed_whole_range(tool, q):
    """r26_fold_signed over its whole input range, |x_j| < 2^31 - 2^26 for j < 4 and
    |x_4| < 2^30 - 2^6 (values within 1024 q of zero), edges included: a congruent value in [0, 2q)
    with limbs < 2^28."""
    c = q - (1 << 124)
    rng = random.Random(17)
    lo_max, top_max = (1 << 31) - (1 << 26) - 1, (1 << 30) - (1 << 6) - 1
    for _ in range(1500):
        x = [rng.choice((-lo_max, lo_max)) if rng.random() < 0.3 else rng.randrange(-lo_max, lo_max + 1) for _ in range(4)]
        x.append(rng.choice((-top_max, top_max)) if rng.random() < 0.3 else rng.randrange(-top_max, top_max + 1))
        g = run(tool, f"G {c} " + " ".join(map(str, x)))
        assert value(g) % q == value(x) % q and 0 <= value(g) < 2 * q and all(0 <= v < (1 << 28) for v in g)


# per-operation limb bounds, as stated in r26.hpp: an element is (max |limb_j| for j < 4, |limb_4|)
def _mont_tw_bound(x):
    return (1 << 26, 5 * max(x) / 64 + (1 << 20) + (1 << 8))


def _mont_bound(x):
    return (1 << 26, (x[1] + 64) / 64 + (1 << 20) + 4)


def _stream_kinds(rows):
    """the product of each stage in the product's schedule for N = 2^(7 + rows): forward columns
    (stages 0-5 on mont_tw, stage 6 on mont), then the row stages: mont, except at rows of 2^9 (one
    row per wave) the first round (row stages 0-2, wave-uniform twiddles) on mont_tw; the inverse
    rows likewise (stages 6-8 on mont_tw at rows of 2^9), then the columns' stage 0 (mont) and
    stages 1-5 (mont_tw); the scaled last stage (mont) is applied separately"""
    if rows == 9:
        fwd_rows = ["tw"] * 3 + ["mont"] * 6
        inv_rows = ["mont"] * 6 + ["tw"] * 3
    else:
        fwd_rows = ["mont"] * rows
        inv_rows = ["mont"] * rows
    return ["tw"] * 6 + ["mont"] + fwd_rows, inv_rows + ["mont"] + ["tw"] * 5


@pytest.mark.parametrize("rows", [7, 8, 9, 10])
def test_stream_schedule_worst_case(rows):
    """Worst-case limb propagation over the streaming schedule with the bounds above, from
    canonical inputs: every product input keeps |limbs| < 2^30.2 (mont and mont_tw need < 2^31),
    the point-wise product's operands meet its bounds, and every folded sum (stages 1, 4, 7, ...)
    has |x_4| < 2^29.6 and |x_j| < 2^31 - 2^26 (the signed fold's range)."""
    fwd, inv = _stream_kinds(rows)
    x = (1 << 26, 1 << 21)
    for k in fwd:
        assert max(x) < 2 ** 30.2
        t = (_mont_tw_bound if k == "tw" else _mont_bound)(x)
        x = (x[0] + t[0], x[1] + t[1])
    assert max(x) < 2 ** 30.2
    # r26_point: 4 mont(a, folded b): digits < 2^28, |top| < 4 ((|a_4| + 64) / 16 + 2^20 + 2)
    y = (1 << 28, (x[1] + 64) / 4 + (1 << 22) + 8)
    for i, k in enumerate(inv):
        s = (2 * y[0], 2 * y[1])  # x + y and x - y of two elements within y's bounds
        assert max(s) < 2 ** 30.2
        if i % 3 == 1:
            assert s[1] < 2 ** 29.6 and s[0] < (1 << 31) - (1 << 26)
            s = ((1 << 27) + (1 << 24), 1 << 21)  # folded: limb 0 < 2^26 + qb_0 < 2^27 + c, others < 2^27
        t = (_mont_tw_bound if k == "tw" else _mont_bound)(s)
        y = (max(s[0], t[0]), max(s[1], t[1]))
    assert max(2 * y[0], 2 * y[1]) < 2 ** 30.2  # the scaled last stage's mont inputs


def _chain(tool, q, cmd, x, y, stages, rng, rows):
    """run a BX / IX chain on one pair; stages: list of (kind, fold); returns the per-stage outputs
    and checks each against exact arithmetic mod q"""
    c = q - (1 << 124)
    R = 1 << 130
    ws = [rng.randrange(q) for _ in stages]
    sels = [rng.randrange(3) for _ in stages]
    parts = []
    for (kind, fold), w, sel in zip(stages, ws, sels):
        extra = f" {fold}" if cmd == "IX" else ""
        parts.append(" ".join(map(str, limbs(w))) + f" {int(kind == 'tw')}{extra} {sel}")
    out = run(tool, f"{cmd} {c} {len(stages)} " + " ".join(map(str, x + y)) + " " + " ".join(parts))
    assert len(out) == 10 * len(stages)
    ex, ey = value(x) % q, value(y) % q
    px, py = x, y
    for s, ((kind, fold), w, sel) in enumerate(zip(stages, ws, sels)):
        if cmd == "BX":
            assert max(abs(v) for v in py) < 2 ** 30.2, f"product input at stage {s}"
            t = ey * w * pow(R, -1, q) % q
            ex, ey = (ex + t) % q, (ex - t) % q
            if sel == 1:
                ex = ey
            elif sel == 2:
                ey = ex
        else:
            d = [a - b for a, b in zip(px, py)]
            assert max(abs(v) for v in d) < 2 ** 30.2, f"product input at stage {s}"
            if fold:
                top = abs(px[4] + py[4])
                assert top < 2 ** 29.6 and max(abs(a + b) for a, b in zip(px[:4], py[:4])) < (1 << 31) - (1 << 26)
            ex, ey = (ex + ey) % q, (ex - ey) * w * pow(R, -1, q) % q
            if sel == 1:
                ey = ex
            elif sel == 2:
                ex = ey
        st = out[10 * s:10 * s + 10]
        assert value(st[:5]) % q == ex and value(st[5:]) % q == ey, f"{cmd} stage {s}"
        px, py = st[:5], st[5:]
    return px, py


@pytest.mark.parametrize("q", PRIMES)
@pytest.mark.parametrize("rows", [7, 9, 10])
def test_stream_schedule_chains(tool, q, rows):
    """The streaming schedule on pairs through the host build: forward chains (mont_tw and mont
    stages as in the columns and rows), a point-wise product of two forward outputs, then inverse
    chains (x + y folded at stages 1, 4, 7, ...), each stage keeping either output pair or a
    doubled one: exact mod q at every stage, every product input below 2^30.2 per limb, every
    folded sum within the signed fold's range, and the results fold (signed) to [0, 2q)."""
    c = q - (1 << 124)
    rng = random.Random(rows * 31 + (q & 0xFF))
    fwd, inv = _stream_kinds(rows)
    for _ in range(40):
        outs = []
        for _ in range(2):
            x, y = _forward_input(q, rng), _forward_input(q, rng)
            outs.append(_chain(tool, q, "BX", x, y, [(k, 0) for k in fwd], rng, rows))
        # forward outputs fold (stores) and multiply point-wise
        for half in outs[0]:
            f = run(tool, f"G {c} " + " ".join(map(str, half)))
            assert value(f) % q == value(half) % q and 0 <= value(f) < 2 * q
        a, b = outs[0][0], outs[1][1]
        pa = run(tool, f"P {c} " + " ".join(map(str, a + b)))
        pb = run(tool, f"P {c} " + " ".join(map(str, outs[0][1] + outs[1][0])))
        assert value(pa) % q == value(a) * value(b) * pow(1 << 128, -1, q) % q
        x, y = _chain(tool, q, "IX", pa, pb, [(k, int(i % 3 == 1)) for i, k in enumerate(inv)], rng, rows)
        for half in (x, y):
            f = run(tool, f"G {c} " + " ".join(map(str, half)))
            assert value(f) % q == value(half) % q and 0 <= value(f) < 2 * q
