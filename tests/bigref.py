"""Independent python big-int statements of the mathematics the oracle must satisfy.

These are textbook definitions (RFC 8439 ChaCha20, Montgomery representation
a*R mod q, the negacyclic DFT with bit-reversed output order), written from the
mathematics rather than from the oracle's code, for small inputs only.
"""

Q64 = (1 << 64) - (1 << 32) + 1
Q128A = (1 << 124) + (1 << 20) + (1 << 18) + 1
Q128B = 0x10000000000000000000000000078001
P150 = 0x2FAEADBE7A0195C011AC195AD10269830E8001


def is_probable_prime(n: int) -> bool:
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]
    for p in small:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def limbs_of(q: int) -> int:
    return ((q - 1).bit_length() + 31) // 32


def rotl(x, r):
    return ((x << r) | (x >> (32 - r))) & 0xFFFFFFFF


def quarter_round(a, b, c, d):
    M = 0xFFFFFFFF
    a = (a + b) & M; d ^= a; d = rotl(d, 16)
    c = (c + d) & M; b ^= c; b = rotl(b, 12)
    a = (a + b) & M; d ^= a; d = rotl(d, 8)
    c = (c + d) & M; b ^= c; b = rotl(b, 7)
    return a, b, c, d


def chacha20_block(key, counter, nonce):
    """RFC 8439 section 2.3."""
    s = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + list(key) + [counter & 0xFFFFFFFF] + list(nonce)
    x = list(s)
    for _ in range(10):
        for (a, b, c, d) in ((0, 4, 8, 12), (1, 5, 9, 13), (2, 6, 10, 14), (3, 7, 11, 15),
                             (0, 5, 10, 15), (1, 6, 11, 12), (2, 7, 8, 13), (3, 4, 9, 14)):
            x[a], x[b], x[c], x[d] = quarter_round(x[a], x[b], x[c], x[d])
    return [(x[i] + s[i]) & 0xFFFFFFFF for i in range(16)]


def stream_words(key, nonce, counter, count):
    out = []
    while len(out) < count:
        out += chacha20_block(key, counter, nonce)
        counter += 1
    return out[:count]


def brv(v, bits):
    return int(format(v, f"0{bits}b")[::-1], 2) if bits else 0


def words_to_int(words, bits):
    v = sum(w << (32 * i) for i, w in enumerate(words))
    return v & ((1 << bits) - 1)


def negacyclic_dft(a, phi, q):
    """A[j] = sum_i a_i phi^((2 brv(j) + 1) i) mod q, plain (non-Montgomery) values."""
    N = len(a)
    bits = N.bit_length() - 1
    out = []
    for j in range(N):
        e = 2 * brv(j, bits) + 1
        w = pow(phi, e, q)
        acc, p = 0, 1
        for i in range(N):
            acc += a[i] * p
            p = p * w % q
        out.append(acc % q)
    return out


def schoolbook_negacyclic(a, b, q):
    N = len(a)
    c = [0] * N
    for i in range(N):
        for j in range(N):
            k = i + j
            if k < N:
                c[k] += a[i] * b[j]
            else:
                c[k - N] -= a[i] * b[j]
    return [x % q for x in c]


class Stream:
    """hmpc's number_generator: 32-bit words of the ChaCha20 stream (key, nonce, counter), one block
    after another (the counter advances per block)."""

    def __init__(self, key, nonce, counter):
        self.key, self.nonce, self.counter = list(key) + [0] * (8 - len(key)), list(nonce), counter
        self.buf = []

    def next(self):
        if not self.buf:
            self.buf = chacha20_block(self.key, self.counter & 0xFFFFFFFF, self.nonce)
            self.counter += 1
        return self.buf.pop(0)

    def bits(self, bits):
        words = (bits + 31) // 32
        v = sum(self.next() << (32 * i) for i in range(words))
        return v & ((1 << bits) - 1)


def sample_mont(st, dist, bits, q):
    """One capability draw of a distribution, as the stored Montgomery word value (R = 2^(32 L))."""
    L = limbs_of(q)
    R = 1 << (32 * L)
    if dist == "uniform":
        b = (q - 1).bit_length() + 80
        W = (b + 31) // 32
        v = st.bits(b)
        return v * pow(1 << (32 * max(L, W)), -1, q) % q
    if dist == "centered_binomial" and bits == 0:
        # uniform ternary (not in hmpc): floor(3 w / 2^64) - 1 for the next 64 bits w
        val = (3 * st.bits(64) >> 64) - 1
    elif dist == "centered_binomial":
        val = bin(st.bits(bits)).count("1") - bits // 2
    elif dist == "binomial":
        val = bin(st.bits(bits)).count("1")
    elif dist == "drown_unsigned":
        words = (bits + 31) // 32
        val = sum(st.next() << (32 * i) for i in range(words)) & ((1 << bits) - 1)
    elif dist == "drown_signed":
        words = (bits + 31) // 32
        raw = sum(st.next() << (32 * i) for i in range(words)) & ((1 << bits) - 1)
        val = raw - (1 << bits) if raw >> (bits - 1) else raw
    else:
        raise ValueError(dist)
    return val % q * R % q


def capability_samples(key, nonce, dist, bits, layout, log_n, n, q):
    """Capability-path samples (stored Montgomery values) of n polynomials of N coefficients:
    element: coefficient e from the stream with counter e; ntt: the stream of b N + t gives t, then
    t + N/2; intt: the stream of b N + 2t gives 2t, then 2t + 1."""
    N = 1 << log_n
    out = [0] * (n * N)
    if layout == "element":
        for e in range(n * N):
            out[e] = sample_mont(Stream(key, nonce, e), dist, bits, q)
        return out
    for b in range(n):
        for t in range(N // 2):
            first = b * N + t if layout == "ntt" else b * N + 2 * t
            second = first + N // 2 if layout == "ntt" else first + 1
            st = Stream(key, nonce, first)
            out[first] = sample_mont(st, dist, bits, q)
            out[second] = sample_mont(st, dist, bits, q)
    return out


def inverse_negacyclic_dft(A, phi, q):
    """a_i = N^-1 sum_j A_j phi^(-(2 brv(j) + 1) i): the inverse of negacyclic_dft."""
    N = len(A)
    bits = N.bit_length() - 1
    ninv = pow(N, -1, q)
    out = []
    for i in range(N):
        acc = 0
        for j in range(N):
            acc += A[j] * pow(phi, -(2 * brv(j, bits) + 1) * i, q)
        out.append(acc * ninv % q)
    return out


def centred(v, p):
    """The centred representative of v in [0, p): values above floor(p / 2) are v - p."""
    return v - p if v > p // 2 else v
