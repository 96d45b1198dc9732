"""Extract the known-answer vectors that hmpc's own tests hold for this path into
reference_vectors.json (data only: inputs and expected outputs).

Sources (read as text, nothing is compiled or run):
  tests/crypto/chacha.cpp:9-196        RFC 8439 quarter round, block function, 5 extra vectors
  tests/ints/mod.cpp:12-40             Montgomery constants for the 150-bit prime p
  tests/ints/poly_mod.cpp:78-96        root / inverse root and 5 NTT outputs for x_i = i, N = 2^14
  tests/ints/num/theory/root_of_unity.cpp:8-9   p = 1601, n = 64
  tests/crypto/cipher.cpp:8-36         RFC 8439 section 2.4.2 encryption (key, nonce, counter 1, 114 bytes)
  tests/expr/crypto/cipher.cpp:8-30    tensor cipher inputs: ubigint<127> x_i = i, shape {4, 2}, key {42}, nonce {1}
  tests/ints/mod.cpp:50-78             signed constructors and the cross-modulus constructors (q = 99)
  tests/expr/crypto/lhe/enc.cpp:15-72  linear BGV: p, q, N, shape, a = b = iota (raw words), s = one, x = iota
  examples/bgv.cpp:25-29               the BGV example's q, N and plaintext modulus p

Usage: python tests/golden/make_reference_vectors.py [/root/reference]
"""
import json
import os
import re
import sys

ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def hexes(text):
    return [int(h.replace("'", ""), 16) for h in re.findall(r"0x([0-9a-fA-F']+)", text)]


def section(text, start, end=None):
    i = text.index(start)
    j = text.index(end, i) if end else len(text)
    return text[i:j]


def chacha_vectors():
    src = open(os.path.join(ROOT, "tests/crypto/chacha.cpp")).read()
    qr = hexes(section(src, 'SECTION("Quarter round")', 'SECTION("Block function")'))
    assert len(qr) == 8
    blk = section(src, 'SECTION("Block function")', "auto param")
    nums = hexes(blk)
    key, nonce, counter = nums[0:8], nums[8:11], nums[11]
    state_init = nums[12:28]
    out = nums[28:44]
    assert len(nums) == 44
    extra = section(src, 'SECTION("Additional test vectors for the block function")')
    # key / nonce / counter evolve exactly as in the test body
    vectors = []
    k, n, c = [0] * 8, [0] * 3, 0
    parts = re.split(r"// Test Vector #\d+", extra)[1:]
    assert len(parts) == 5
    for idx, part in enumerate(parts):
        for m in re.finditer(r"counter\[0\] = 0x([0-9a-fA-F']+)", part):
            c = int(m.group(1).replace("'", ""), 16)
        for m in re.finditer(r"key\[(\d)\] = 0x([0-9a-fA-F']+)", part):
            k[int(m.group(1))] = int(m.group(2).replace("'", ""), 16)
        for m in re.finditer(r"nonce\[(\d)\] = 0x([0-9a-fA-F']+)", part):
            n[int(m.group(1))] = int(m.group(2).replace("'", ""), 16)
        expected = [int(v.replace("'", ""), 16) for v in re.findall(r"state\[\d+\] == 0x([0-9a-fA-F']+)", part)]
        assert len(expected) == 16
        # vector #2 is the second block of the same stream (counter advanced by the engine)
        vectors.append({"key": list(k), "nonce": list(n), "counter": c + (1 if idx == 1 else 0), "block": expected})
    return {
        "quarter_round": {"in": qr[0:4], "out": qr[4:8]},
        "block": {"key": key, "nonce": nonce, "counter": counter, "initial_state": state_init, "block": out},
        "extra": vectors,
    }


def cipher_vectors():
    src = open(os.path.join(ROOT, "tests/crypto/cipher.cpp")).read()
    key = hexes(section(src, "key[] =", ";"))
    nonce = hexes(section(src, "nonce[] =", ";"))
    counter = hexes(section(src, "counter[] =", ";"))
    plain = hexes(section(src, "plaintext_bytes[] =", "};"))
    cipher = hexes(section(src, "expected_ciphertext_bytes[] =", "};"))
    assert len(key) == 8 and len(nonce) == 3 and len(counter) == 1 and len(plain) == 114 and len(cipher) == 114
    tsrc = open(os.path.join(ROOT, "tests/expr/crypto/cipher.cpp")).read()
    bits = int(re.search(r"ubigint<(\d+)>", tsrc).group(1))
    shape = [int(v) for v in re.search(r"hmpc::shape\{(\d+), (\d+)\}", tsrc).groups()]
    tkey = int(re.search(r"key = \{(\d+)\}", tsrc).group(1))
    tnonce = int(re.search(r"nonce = \{(\d+)\}", tsrc).group(1))
    return {
        "rfc8439_2_4_2": {"key": key, "nonce": nonce, "counter": counter[0], "plaintext": plain, "ciphertext": cipher},
        "tensor": {"bits": bits, "limbs": (bits + 31) // 32, "shape": shape, "key": [tkey] + [0] * 7,
                   "nonce": [tnonce, 0, 0], "values": "x_i = i"},
    }


def mod_vectors():
    src = open(os.path.join(ROOT, "tests/ints/mod.cpp")).read()
    p = int(re.search(r"constexpr auto p = 0x([0-9a-f']+)_int", src).group(1).replace("'", ""), 16)
    r_mod_p = int(re.search(r"reduced_auxiliary_modulus == 0x([0-9a-f']+)_int", src).group(1).replace("'", ""), 16)
    m512 = int(re.search(r"fivetwelve_montgomery = 0x([0-9a-f']+)_int", src).group(1).replace("'", ""), 16)
    big = int(re.search(r"mod_p\(0x(2faeadbe7a0195c011ac195ad10269830e80010+2a)_int\)\)", src).group(1), 16)
    neg_big = int(re.search(r"mod_p\(-0x(2faeadbe7a0195c011ac195ad10269830e80010+2a)_int\)\) == 0x([0-9a-f']+)_int",
                            src).group(2).replace("'", ""), 16)
    modulo = section(src, 'SECTION("Modulo")')
    small_q = int(re.search(r"constexpr auto q = (\d+)_int", src).group(1))
    hi = int(re.search(r"i <= (\d+)", modulo).group(1))
    return {"p": p, "R_mod_p": r_mod_p, "mont_512": m512, "big_unsigned": big, "big_unsigned_value": 42,
            "neg_one_value": p - 1, "neg_big_unsigned": -big, "neg_big_unsigned_value": neg_big,
            "cross": {"q": small_q, "n_max": hi,
                      "rules": "mod_p(mod_q(n)) == n, mod_q(mod_p(n)) == n, mod_p(mod_q(-n)) == p - n and "
                               "mod_q(mod_p(-n)) == q - n for 0 < n <= n_max (0 for n = 0)"}}


def lhe_vectors():
    src = open(os.path.join(ROOT, "tests/expr/crypto/lhe/enc.cpp")).read()
    p = int(re.search(r"constexpr auto p = 0x([0-9a-f']+)_int", src).group(1).replace("'", ""), 16)
    q = int(re.search(r"constexpr auto q = 0x([0-9a-f']+)_int", src).group(1).replace("'", ""), 16)
    log_n = int(re.search(r"N = hmpc::size\{1\} << (\d+)", src).group(1))
    shape = [int(v) for v in re.search(r"auto shape = hmpc::shape\{(\d+), (\d+)\}", src).groups()]
    ex = open(os.path.join(ROOT, "examples/bgv.cpp")).read()
    eq = int(re.search(r"constexpr auto q = 0x([0-9a-f']+)_int", ex).group(1).replace("'", ""), 16)
    elog = int(re.search(r"N = 1 << (\d+)", ex).group(1))
    ep = int(re.search(r"mod_q\{0x([0-9a-f']+)_int\}", ex).group(1).replace("'", ""), 16)
    return {
        "enc_test": {"p": p, "q": q, "log_n": log_n, "shape": shape, "plaintext": "ntt",
                     "a": "raw word i (bit_copy)", "b": "a", "s": "one (R mod q)", "x": "raw word i (bit_copy)",
                     "check": "dec<plaintext>(s, enc(key{a, b}, x, randomness(shape))) == x"},
        "bgv_example": {"q": eq, "log_n": elog, "p": ep},
    }


def poly_vectors():
    src = open(os.path.join(ROOT, "tests/ints/poly_mod.cpp")).read()
    p = int(re.search(r"constexpr auto p = 0x([0-9a-f']+)_int", src).group(1).replace("'", ""), 16)
    root = int(re.search(r"<integer>\(y\.root\) == 0x([0-9a-f']+)_int", src).group(1).replace("'", ""), 16)
    iroot = int(re.search(r"<integer>\(ntt_y\.root\) == 0x([0-9a-f']+)_int", src).group(1).replace("'", ""), 16)
    outs = {int(i): int(v.replace("'", ""), 16) for i, v in
            re.findall(r"ntt_x_elements\[index_type\{j, (\d+)\}\]\)\) == 0x([0-9a-f']+)_int", src)}
    assert len(outs) == 5
    return {"p": p, "log_n": 14, "root": root, "inverse_root": iroot, "ntt_of_iota": {str(k): v for k, v in outs.items()}}


def root_vectors():
    src = open(os.path.join(ROOT, "tests/ints/num/theory/root_of_unity.cpp")).read()
    p = int(re.search(r"constexpr auto p = 0b([01']+)_int", src).group(1).replace("'", ""), 2)
    n = 1 << int(re.search(r"n = 1 << (\d+)", src).group(1))
    return {"p": p, "n": n}


def main():
    data = {
        "source": "iko4/hmpc tests (see make_reference_vectors.py for file:line)",
        "chacha20": chacha_vectors(),
        "mod": mod_vectors(),
        "poly_mod": poly_vectors(),
        "root_of_unity": root_vectors(),
        "cipher": cipher_vectors(),
        "lhe": lhe_vectors(),
    }

    def hexify(o):
        if isinstance(o, dict):
            return {k: hexify(v) for k, v in o.items()}
        if isinstance(o, list):
            return [hexify(v) for v in o]
        if isinstance(o, int) and abs(o) >= 1 << 32:
            return hex(o)
        return o

    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(hexify(data), f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
