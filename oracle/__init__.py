"""CPU oracle for the NTT / BGV / ChaCha20 path — TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``liboracle.so``, a plain C++ restatement of hmpc's algorithms
(``hmpc_oracle.cpp``, which cites the reference file:line for every function).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import this package, and only as the checker / CPU baseline: the product
(``hmpc_amd``) never routes through it.

All polynomial buffers are limb-planar ``uint32[L][count]`` numpy arrays (the
layout of hmpc's tensors, hmpc/comp/accessor.hpp:191-197).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)


def build() -> str:
    """Compile liboracle.so from hmpc_oracle.cpp (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        sz = ctypes.c_size_t
        i = ctypes.c_int
        u64 = ctypes.c_uint64
        L.oracle_modulus_new.restype = vp
        L.oracle_modulus_new.argtypes = [_u32p, i]
        L.oracle_modulus_free.argtypes = [vp]
        L.oracle_modulus_limbs.argtypes = [vp]
        L.oracle_modulus_bit_size.argtypes = [vp]
        L.oracle_modulus_constants.argtypes = [vp, _u32p, _u32p, _u32p]
        L.oracle_from_unsigned.argtypes = [vp, _u32p, i, _u32p]
        L.oracle_from_small_signed.argtypes = [vp, ctypes.c_longlong, _u32p]
        L.oracle_to_unsigned.argtypes = [vp, _u32p, _u32p]
        for f in (L.oracle_mul, L.oracle_add, L.oracle_sub):
            f.argtypes = [vp, _u32p, _u32p, _u32p]
        L.oracle_pow.argtypes = [vp, _u32p, _u32p, i, _u32p]
        L.oracle_invert.argtypes = [vp, _u32p, _u32p]
        L.oracle_montgomery_reduce.argtypes = [vp, _u32p, i, i, _u32p]
        L.oracle_root_of_unity.argtypes = [vp, i, i, _u32p]
        L.oracle_roots.argtypes = [vp, i, i, _u32p]
        L.oracle_chacha20_block.argtypes = [_u32p, ctypes.c_uint32, _u32p, _u32p]
        L.oracle_chacha20_stream.argtypes = [_u32p, _u32p, ctypes.c_uint32, sz, _u32p]
        L.oracle_cipher_encrypt.argtypes = [i, sz, _u32p, _u32p, ctypes.c_uint32, _u32p, _u32p]
        L.oracle_cipher_decrypt.argtypes = [i, sz, _u32p, _u32p, ctypes.c_uint32, _u32p, _u32p]
        L.oracle_ntt.argtypes = [vp, i, sz, i, _u32p, _u32p, i]
        L.oracle_pointwise.argtypes = [vp, i, sz, _u32p, _u32p, sz, _u32p, i]
        L.oracle_sample_uniform.argtypes = [vp, _u32p, u64, sz, _u32p, i]
        L.oracle_sample_cbd.argtypes = [vp, _u32p, u64, i, sz, _u32p, i]
        L.oracle_sample_words.argtypes = [_u32p, u64, sz, _u32p]
        L.oracle_sample_cbd_capability.argtypes = [vp, _u32p, _u32p, i, i, sz, _u32p]
        L.oracle_bgv_encrypt.argtypes = [vp, i, sz, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p,
                                         i, i, i, _u32p, _u32p, i]
        L.oracle_bgv_decrypt.argtypes = [vp, i, sz, _u32p, _u32p, _u32p, _u32p, i]
        L.oracle_schoolbook_negacyclic.argtypes = [vp, i, _u32p, _u32p, _u32p]
        L.oracle_cast.argtypes = [vp, vp, sz, _u32p, _u32p]
        L.oracle_from_signed_bigint.argtypes = [vp, _u32p, i, i, _u32p]
        L.oracle_sample_capability.argtypes = [vp, _u32p, _u32p, i, i, i, i, sz, _u32p]
        L.oracle_bgv_encrypt_plaintext.argtypes = [vp, vp, i, sz, _u32p, _u32p, _u32p, i, _u32p, _u32p, i, i, i, i,
                                                   _u32p, _u32p, i]
        L.oracle_bgv_decrypt_plaintext.argtypes = [vp, vp, i, sz, _u32p, _u32p, _u32p, i, _u32p, i]
        L.oracle_polymul.restype = ctypes.c_double
        L.oracle_polymul.argtypes = [vp, i, sz, _u32p, _u32p, _u32p, i]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_u32p)


def int_to_limbs(value: int, limbs: int) -> np.ndarray:
    return np.array([(value >> (32 * i)) & 0xFFFFFFFF for i in range(limbs)], dtype=np.uint32)


def limbs_to_int(limbs) -> int:
    return sum(int(x) << (32 * i) for i, x in enumerate(limbs))


def planar_to_ints(planes: np.ndarray) -> list[int]:
    """[L][count] -> list of python ints."""
    L = planes.shape[0]
    acc = np.zeros(planes.shape[1], dtype=object)
    for l in range(L - 1, -1, -1):
        acc = acc * (1 << 32) + planes[l].astype(object)
    return [int(x) for x in acc]


def ints_to_planar(values, limbs: int) -> np.ndarray:
    values = [int(v) for v in values]
    out = np.empty((limbs, len(values)), dtype=np.uint32)
    for l in range(limbs):
        out[l] = [(v >> (32 * l)) & 0xFFFFFFFF for v in values]
    return out


def key_array(key) -> np.ndarray:
    k = np.zeros(8, dtype=np.uint32)
    k[: len(key)] = key
    return k


class Modulus:
    """Oracle view of hmpc::ints::mod<q> (runtime modulus)."""

    def __init__(self, q: int):
        self.q = int(q)
        words = max(1, (self.q.bit_length() + 31) // 32)
        qa = int_to_limbs(self.q, words)
        self._h = lib().oracle_modulus_new(_p(qa), words)
        if not self._h:
            raise ValueError("invalid modulus")
        self.limbs = lib().oracle_modulus_limbs(self._h)
        self.bit_size = lib().oracle_modulus_bit_size(self._h)
        self.R = 1 << (32 * self.limbs)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.oracle_modulus_free(h)
            self._h = None

    # -- constants
    def constants(self):
        npr = np.zeros(1, np.uint32)
        r1 = np.zeros(self.limbs, np.uint32)
        r2 = np.zeros(self.limbs, np.uint32)
        lib().oracle_modulus_constants(self._h, _p(npr), _p(r1), _p(r2))
        return int(npr[0]), limbs_to_int(r1), limbs_to_int(r2)

    # -- scalars (python ints in Montgomery representation unless noted)
    def from_unsigned(self, value: int) -> int:
        words = max(1, (int(value).bit_length() + 31) // 32)
        v = int_to_limbs(value, words)
        out = np.zeros(self.limbs, np.uint32)
        lib().oracle_from_unsigned(self._h, _p(v), words, _p(out))
        return limbs_to_int(out)

    def from_small_signed(self, value: int) -> int:
        out = np.zeros(self.limbs, np.uint32)
        lib().oracle_from_small_signed(self._h, int(value), _p(out))
        return limbs_to_int(out)

    def to_unsigned(self, x: int) -> int:
        a = int_to_limbs(x, self.limbs)
        out = np.zeros(self.limbs, np.uint32)
        lib().oracle_to_unsigned(self._h, _p(a), _p(out))
        return limbs_to_int(out)

    def _bin(self, f, a: int, b: int) -> int:
        x = int_to_limbs(a, self.limbs)
        y = int_to_limbs(b, self.limbs)
        out = np.zeros(self.limbs, np.uint32)
        f(self._h, _p(x), _p(y), _p(out))
        return limbs_to_int(out)

    def mul(self, a, b):
        return self._bin(lib().oracle_mul, a, b)

    def add(self, a, b):
        return self._bin(lib().oracle_add, a, b)

    def sub(self, a, b):
        return self._bin(lib().oracle_sub, a, b)

    def pow(self, base: int, exponent: int) -> int:
        words = max(1, (int(exponent).bit_length() + 31) // 32)
        e = int_to_limbs(exponent, words)
        b = int_to_limbs(base, self.limbs)
        out = np.zeros(self.limbs, np.uint32)
        lib().oracle_pow(self._h, _p(b), _p(e), words, _p(out))
        return limbs_to_int(out)

    def invert(self, x: int) -> int:
        a = int_to_limbs(x, self.limbs)
        out = np.zeros(self.limbs, np.uint32)
        lib().oracle_invert(self._h, _p(a), _p(out))
        return limbs_to_int(out)

    def montgomery_reduce(self, value: int, value_limbs: int, rounds: int) -> int:
        v = int_to_limbs(value, value_limbs)
        out = np.zeros(self.limbs, np.uint32)
        lib().oracle_montgomery_reduce(self._h, _p(v), value_limbs, rounds, _p(out))
        return limbs_to_int(out)

    def root_of_unity(self, log_degree: int, inverse: bool = False) -> int:
        out = np.zeros(self.limbs, np.uint32)
        rc = lib().oracle_root_of_unity(self._h, log_degree, int(inverse), _p(out))
        if rc:
            raise ValueError("no root of unity of that order")
        return limbs_to_int(out)

    def roots(self, log_n: int, inverse: bool = False) -> np.ndarray:
        out = np.zeros((self.limbs, 1 << log_n), np.uint32)
        if lib().oracle_roots(self._h, log_n, int(inverse), _p(out)):
            raise ValueError("no root of unity of that order")
        return out

    # -- tensors ([L][count] limb-planar numpy arrays)
    def ntt(self, x: np.ndarray, log_n: int, inverse: bool = False, threads: int = 1) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.uint32)
        assert x.shape[0] == self.limbs and x.shape[1] % (1 << log_n) == 0
        n = x.shape[1] >> log_n
        out = np.empty_like(x)
        rc = lib().oracle_ntt(self._h, log_n, n, int(inverse), _p(x), _p(out), threads)
        if rc:
            raise ValueError(f"oracle_ntt failed ({rc})")
        return out

    def pointwise(self, op: str, a: np.ndarray, b: np.ndarray, threads: int = 1) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        out = np.empty_like(a)
        code = {"add": 0, "sub": 1, "mul": 2}[op]
        lib().oracle_pointwise(self._h, code, a.shape[1], _p(a), _p(b), b.shape[1], _p(out), threads)
        return out

    def sample_uniform(self, key, offset: int, count: int, threads: int = 1) -> np.ndarray:
        k = key_array(key)
        out = np.empty((self.limbs, count), np.uint32)
        lib().oracle_sample_uniform(self._h, _p(k), offset, count, _p(out), threads)
        return out

    def sample_cbd(self, key, offset: int, bits: int, count: int, threads: int = 1) -> np.ndarray:
        k = key_array(key)
        out = np.empty((self.limbs, count), np.uint32)
        lib().oracle_sample_cbd(self._h, _p(k), offset, bits, count, _p(out), threads)
        return out

    def sample_cbd_capability(self, key, nonce, bits: int, log_n: int, n: int) -> np.ndarray:
        k = key_array(key)
        nz = np.array(nonce, dtype=np.uint32)
        out = np.empty((self.limbs, n << log_n), np.uint32)
        lib().oracle_sample_cbd_capability(self._h, _p(k), _p(nz), bits, log_n, n, _p(out))
        return out

    def bgv_encrypt(self, log_n, pk_a, pk_b, messages, p_mont: int, key, nonce,
                    bits=(2, 40, 40), threads: int = 1):
        n = messages.size >> log_n
        k = key_array(key)
        nz = np.array(nonce, dtype=np.uint32).copy()
        pm = int_to_limbs(p_mont, self.limbs)
        msg = np.ascontiguousarray(messages.reshape(-1), dtype=np.uint32)
        c0 = np.empty((self.limbs, n << log_n), np.uint32)
        c1 = np.empty_like(c0)
        rc = lib().oracle_bgv_encrypt(self._h, log_n, n, _p(np.ascontiguousarray(pk_a)), _p(np.ascontiguousarray(pk_b)),
                                      _p(msg), _p(pm), _p(k), _p(nz), bits[0], bits[1], bits[2],
                                      _p(c0), _p(c1), threads)
        if rc:
            raise ValueError("oracle_bgv_encrypt failed")
        return c0, c1, nz

    def bgv_decrypt(self, log_n, s_ntt, c0, c1, threads: int = 1) -> np.ndarray:
        n = c0.shape[1] >> log_n
        out = np.empty(n << log_n, np.uint32)
        rc = lib().oracle_bgv_decrypt(self._h, log_n, n, _p(np.ascontiguousarray(s_ntt)), _p(np.ascontiguousarray(c0)),
                                      _p(np.ascontiguousarray(c1)), _p(out), threads)
        if rc:
            raise ValueError("oracle_bgv_decrypt failed")
        return out

    # -- cross-modulus cast, capability samplers, BGV with an Rp plaintext
    def cast_to(self, other: "Modulus", x: np.ndarray) -> np.ndarray:
        """mod<other>(mod<self>) element-wise (centred lift, hmpc/ints/mod.hpp:259-307)."""
        x = np.ascontiguousarray(x, dtype=np.uint32)
        assert x.shape[0] == self.limbs
        out = np.empty((other.limbs, x.shape[1]), np.uint32)
        lib().oracle_cast(self._h, other._h, x.shape[1], _p(x), _p(out))
        return out

    def from_signed_bigint(self, value: int, bits: int) -> int:
        """mod(sbigint<bits>) of a python int in [-2^(bits-1), 2^(bits-1))."""
        words = (bits + 31) // 32
        w = int_to_limbs(value % (1 << (32 * words)), words)
        out = np.zeros(self.limbs, np.uint32)
        lib().oracle_from_signed_bigint(self._h, _p(w), words, bits, _p(out))
        return limbs_to_int(out)

    DISTRIBUTIONS = {"uniform": 0, "centered_binomial": 1, "binomial": 2, "drown_signed": 3, "drown_unsigned": 4}
    LAYOUTS = {"element": 0, "ntt": 1, "intt": 2}

    def sample_capability(self, key, nonce, dist: str, bits: int, layout: str, log_n: int, n: int) -> np.ndarray:
        k = key_array(key)
        nz = np.array(nonce, dtype=np.uint32)
        out = np.empty((self.limbs, n << log_n), np.uint32)
        rc = lib().oracle_sample_capability(self._h, _p(k), _p(nz), self.DISTRIBUTIONS[dist], bits, self.LAYOUTS[layout],
                                            log_n, n, _p(out))
        if rc:
            raise ValueError("oracle_sample_capability failed")
        return out

    def bgv_encrypt_plaintext(self, pmod: "Modulus", log_n, pk_a, pk_b, plaintext, plaintext_ntt: bool, key, nonce,
                              bits_u=2, v_dist="centered_binomial", v_bits=40, bits_w=40, threads: int = 1):
        n = plaintext.shape[1] >> log_n
        k = key_array(key)
        nz = np.array(nonce, dtype=np.uint32).copy()
        x = np.ascontiguousarray(plaintext, dtype=np.uint32)
        c0 = np.empty((self.limbs, n << log_n), np.uint32)
        c1 = np.empty_like(c0)
        rc = lib().oracle_bgv_encrypt_plaintext(self._h, pmod._h, log_n, n, _p(np.ascontiguousarray(pk_a)),
                                                _p(np.ascontiguousarray(pk_b)), _p(x), int(plaintext_ntt), _p(k), _p(nz),
                                                bits_u, self.DISTRIBUTIONS[v_dist], v_bits, bits_w, _p(c0), _p(c1), threads)
        if rc:
            raise ValueError("oracle_bgv_encrypt_plaintext failed")
        return c0, c1, nz

    def bgv_decrypt_plaintext(self, pmod: "Modulus", log_n, s_ntt, c0, c1, plaintext_ntt: bool, threads: int = 1):
        n = c0.shape[1] >> log_n
        out = np.empty((pmod.limbs, n << log_n), np.uint32)
        rc = lib().oracle_bgv_decrypt_plaintext(self._h, pmod._h, log_n, n, _p(np.ascontiguousarray(s_ntt)),
                                                _p(np.ascontiguousarray(c0)), _p(np.ascontiguousarray(c1)),
                                                int(plaintext_ntt), _p(out), threads)
        if rc:
            raise ValueError("oracle_bgv_decrypt_plaintext failed")
        return out

    def schoolbook_negacyclic(self, log_n, a, b) -> np.ndarray:
        out = np.empty((self.limbs, 1 << log_n), np.uint32)
        lib().oracle_schoolbook_negacyclic(self._h, log_n, _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)), _p(out))
        return out

    def polymul_timed(self, log_n, a, b, threads: int):
        out = np.empty_like(a)
        n = a.shape[1] >> log_n
        secs = lib().oracle_polymul(self._h, log_n, n, _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)), _p(out), threads)
        return out, secs


def chacha20_block(key, counter: int, nonce) -> np.ndarray:
    k = key_array(key)
    nz = np.array(nonce, dtype=np.uint32)
    out = np.empty(16, np.uint32)
    lib().oracle_chacha20_block(_p(k), counter, _p(nz), _p(out))
    return out


def chacha20_stream(key, nonce, counter: int, words: int) -> np.ndarray:
    k = key_array(key)
    nz = np.array(nonce, dtype=np.uint32)
    out = np.empty(words, np.uint32)
    lib().oracle_chacha20_stream(_p(k), _p(nz), counter, words, _p(out))
    return out


def cipher_encrypt(planar: np.ndarray, key, nonce, counter: int = 0) -> np.ndarray:
    """hmpc's tensor cipher: limb-planar u32[L][count] -> element-major u32[count][L]"""
    planar = np.ascontiguousarray(planar, dtype=np.uint32)
    limbs, count = planar.shape
    k, nz = key_array(key), np.array(list(nonce) + [0] * (3 - len(nonce)), dtype=np.uint32)
    out = np.empty((count, limbs), np.uint32)
    lib().oracle_cipher_encrypt(limbs, count, _p(k), _p(nz), counter, _p(planar), _p(out))
    return out


def cipher_decrypt(words: np.ndarray, key, nonce, counter: int = 0) -> np.ndarray:
    """element-major u32[count][L] -> limb-planar u32[L][count]"""
    words = np.ascontiguousarray(words, dtype=np.uint32)
    count, limbs = words.shape
    k, nz = key_array(key), np.array(list(nonce) + [0] * (3 - len(nonce)), dtype=np.uint32)
    out = np.empty((limbs, count), np.uint32)
    lib().oracle_cipher_decrypt(limbs, count, _p(k), _p(nz), counter, _p(words), _p(out))
    return out


def drown_bits(bound: int, signed: bool, security: int = 80) -> int:
    """Width of drown_uniform's draw: bit_width((bound - 1) << security) + signed
    (hmpc/random/uniform.hpp:34-47)."""
    return ((bound - 1) << security).bit_length() + int(bool(signed))


def sample_words(key, offset: int, count: int) -> np.ndarray:
    k = key_array(key)
    out = np.empty(count, np.uint32)
    lib().oracle_sample_words(_p(k), offset, count, _p(out))
    return out
