"""ctypes bindings to libhmpc_hip.so (the C ABI in include/hmpc_hip.h).

This is the Python host side used by the tests and the benchmark. It is plumbing only:
every computation runs in the HIP library, and loading fails loudly if the library is
missing (there is no CPU fallback).
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhmpc_hip.so")
# capability-stream width that draws uniform ternary values (HMPC_BITS_TERNARY)
BITS_TERNARY = 0

_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_i = ctypes.c_int
_u64 = ctypes.c_uint64

# exported symbol -> (restype, argtypes); every symbol declared in include/hmpc_hip.h
SIGNATURES = {
    "hmpc_last_error": (ctypes.c_char_p, []),
    "hmpc_context_create": (_i, [_i, _u32p, ctypes.POINTER(_vp)]),
    "hmpc_context_destroy": (_i, [_vp]),
    "hmpc_context_get_rng": (_i, [_vp, _u32p, _u32p]),
    "hmpc_context_set_rng": (_i, [_vp, _u32p, _u32p]),
    "hmpc_context_wait": (_i, [_vp]),
    "hmpc_context_info": (_i, [_vp, _vp]),
    "hmpc_malloc": (_i, [_vp, _sz, ctypes.POINTER(_vp)]),
    "hmpc_free": (_i, [_vp, _vp]),
    "hmpc_memcpy_h2d": (_i, [_vp, _vp, _vp, _sz]),
    "hmpc_memcpy_d2h": (_i, [_vp, _vp, _vp, _sz]),
    "hmpc_memcpy_d2d": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "hmpc_host_alloc": (_i, [_vp, _sz, ctypes.POINTER(_vp)]),
    "hmpc_host_free": (_i, [_vp, _vp]),
    "hmpc_memcpy_h2d_async": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "hmpc_memcpy_d2h_async": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "hmpc_stream_create": (_i, [_vp, ctypes.POINTER(_vp)]),
    "hmpc_stream_destroy": (_i, [_vp, _vp]),
    "hmpc_stream_synchronize": (_i, [_vp, _vp]),
    "hmpc_modulus_create": (_i, [_vp, _u32p, _i, ctypes.POINTER(_vp)]),
    "hmpc_modulus_destroy": (_i, [_vp]),
    "hmpc_modulus_info": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _u32p, _u32p, _u32p]),
    "hmpc_modulus_prepare": (_i, [_vp, _i]),
    "hmpc_modulus_set_kernels": (_i, [_vp, _i]),
    "hmpc_modulus_set_schedule": (_i, [_vp, _i, _i, _i]),
    "hmpc_modulus_set_bgv_schedule": (_i, [_vp, _i, _i, _i]),
    "hmpc_modulus_get_schedule": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "hmpc_autotune": (_i, [_vp, _i, _sz, _vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_float)]),
    "hmpc_autotune_bgv": (_i, [_vp, _vp, _i, _sz, _vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_float)]),
    "hmpc_modulus_root": (_i, [_vp, _i, _i, _u32p]),
    "hmpc_modulus_twiddles": (_i, [_vp, _i, _i, ctypes.POINTER(_vp)]),
    "hmpc_ntt": (_i, [_vp, _i, _sz, _i, _vp, _vp, _vp]),
    "hmpc_ntt_plain": (_i, [_vp, _i, _sz, _i, _vp, _vp, _vp, _vp]),
    "hmpc_pointwise": (_i, [_vp, _i, _sz, _vp, _vp, _sz, _vp, _vp]),
    "hmpc_pointwise_scalar": (_i, [_vp, _i, _sz, _vp, _u32p, _vp, _vp]),
    "hmpc_ciphertext_op": (_i, [_vp, _i, _sz, _vp, _vp, _vp, _vp, _sz, _vp, _vp, _vp]),
    "hmpc_ciphertext_scalar": (_i, [_vp, _i, _sz, _vp, _vp, _u32p, _vp, _vp, _vp]),
    "hmpc_fma": (_i, [_vp, _sz, _vp, _vp, _vp, _vp, _vp]),
    "hmpc_reduce": (_i, [_vp, _i, _sz, _vp, _vp, _vp]),
    "hmpc_matvec": (_i, [_vp, _sz, _sz, _sz, _vp, _i, _vp, _i, _vp, _vp]),
    "hmpc_cipher_encrypt": (_i, [_vp, _i, _sz, _u32p, _u32p, ctypes.c_uint32, _vp, _vp, _vp]),
    "hmpc_cipher_decrypt": (_i, [_vp, _i, _sz, _u32p, _u32p, ctypes.c_uint32, _vp, _vp, _vp]),
    "hmpc_from_u32": (_i, [_vp, _sz, _vp, _vp, _vp]),
    "hmpc_to_unsigned": (_i, [_vp, _sz, _vp, _vp, _vp]),
    "hmpc_sample_uniform": (_i, [_vp, _u32p, _u64, _sz, _vp, _vp]),
    "hmpc_sample_cbd": (_i, [_vp, _u32p, _u64, _i, _sz, _vp, _vp]),
    "hmpc_sample_words": (_i, [_vp, _u32p, _u64, _sz, _vp, _vp]),
    "hmpc_sample_cbd_ntt": (_i, [_vp, _vp, _i, _i, _sz, _vp, _vp]),
    "hmpc_polymul": (_i, [_vp, _i, _sz, _vp, _vp, _vp, _vp]),
    "hmpc_polymul_noalloc": (_i, [_vp, _i, _sz, _vp, _vp, _vp, _vp]),
    "hmpc_bgv_encrypt": (_i, [_vp, _vp, _i, _sz, _vp, _vp, _vp, _u32p, _i, _i, _i, _vp, _vp, _vp]),
    "hmpc_bgv_encrypt_plain": (_i, [_vp, _vp, _i, _sz, _vp, _vp, _vp, _u32p, _i, _i, _i, _vp, _vp, _vp]),
    "hmpc_bgv_decrypt": (_i, [_vp, _i, _sz, _vp, _vp, _vp, _vp, _vp]),
    "hmpc_cast": (_i, [_vp, _vp, _sz, _vp, _vp, _vp]),
    "hmpc_sample_capability": (_i, [_vp, _vp, _i, _i, _i, _i, _sz, _vp, _vp]),
    "hmpc_drown_bits": (_i, [_u32p, _i, _i, _i, ctypes.POINTER(_i)]),
    "hmpc_bgv_encrypt_plaintext": (_i, [_vp, _vp, _vp, _i, _sz, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp]),
    "hmpc_bgv_decrypt_plaintext": (_i, [_vp, _vp, _i, _sz, _vp, _vp, _vp, _i, _vp, _vp]),
    "hmpc_batch_create": (_i, [_vp, ctypes.POINTER(_vp)]),
    "hmpc_batch_destroy": (_i, [_vp]),
    "hmpc_batch_add_polymul": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "hmpc_batch_add_bgv_encrypt": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _u32p, _i, _i, _i, _vp, _vp]),
    "hmpc_batch_run": (_i, [_vp, _i, _vp]),
    "hmpc_batch_info": (_i, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "hmpc_server_create": (_i, [_vp, _sz, _i, ctypes.POINTER(_vp)]),
    "hmpc_server_destroy": (_i, [_vp]),
    "hmpc_server_add_key": (_i, [_vp, _vp, _i, _vp, _vp, _u32p, _i, _i, _i, ctypes.POINTER(_i)]),
    "hmpc_server_submit_encrypt": (_i, [_vp, _i, _sz, _vp, _u32p, ctypes.POINTER(_u64)]),
    "hmpc_server_submit_polymul": (_i, [_vp, _vp, _i, _sz, _vp, _vp, ctypes.POINTER(_u64)]),
    "hmpc_server_poll": (_i, [_vp, _u64, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_double)]),
    "hmpc_server_wait": (_i, [_vp, _u64]),
    "hmpc_server_result": (_i, [_vp, _u64, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "hmpc_server_release": (_i, [_vp, _u64]),
    "hmpc_server_completed": (_i, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(ctypes.c_double), _sz, ctypes.POINTER(_sz)]),
    "hmpc_server_stats": (_i, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_u64),
                               ctypes.POINTER(_u64)]),
    "hmpc_context_set_timing": (_i, [_vp, _i]),
    "hmpc_context_timing": (_i, [_vp, _i, ctypes.c_char_p, _sz, ctypes.POINTER(_sz), ctypes.POINTER(ctypes.c_double)]),
    "hmpc_event_create": (_i, [_vp, ctypes.POINTER(_vp)]),
    "hmpc_event_destroy": (_i, [_vp, _vp]),
    "hmpc_event_record": (_i, [_vp, _vp, _vp]),
    "hmpc_event_elapsed_ms": (_i, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


class HmpcError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"hmpc error {code}: {message}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -m hmpc_amd.build`")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int):
    if rc != 0:
        raise HmpcError(rc, load().hmpc_last_error().decode())


def _arr(values, n=None) -> np.ndarray:
    a = np.zeros(n if n is not None else len(values), dtype=np.uint32)
    a[: len(values)] = values
    return a


def _p(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def int_to_words(value: int, words: int) -> np.ndarray:
    return np.array([(int(value) >> (32 * i)) & 0xFFFFFFFF for i in range(words)], dtype=np.uint32)


def words_to_int(words) -> int:
    return sum(int(w) << (32 * i) for i, w in enumerate(words))


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
        ("compute_units", ctypes.c_int),
        ("wavefront_size", ctypes.c_int),
        ("clock_khz", ctypes.c_int),
        ("memory_clock_khz", ctypes.c_int),
        ("memory_bus_width", ctypes.c_int),
        ("global_memory", ctypes.c_size_t),
        ("lds_per_workgroup", ctypes.c_size_t),
    ]


class Buffer:
    """Raw device allocation owned by a context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        check(load().hmpc_malloc(ctx.handle, max(self.nbytes, 4), ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, ctx: "Context", a: np.ndarray) -> "Buffer":
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.nbytes)
        if a.nbytes:
            check(load().hmpc_memcpy_h2d(ctx.handle, b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_numpy(self, shape, dtype=np.uint32) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if out.nbytes:
            check(load().hmpc_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr is not None and _lib is not None and self.ctx.handle:
            _lib.hmpc_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """hmpc::comp::queue analogue: one device, one RNG capability (key, nonce)."""

    def __init__(self, device: int = 0, key=None):
        load()
        h = _vp()
        k = _arr(key, 8) if key is not None else None
        check(_lib.hmpc_context_create(device, _p(k) if k is not None else None, ctypes.byref(h)))
        self.handle = h.value

    def close(self):
        if self.handle and _lib is not None:
            _lib.hmpc_context_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def wait(self):
        check(_lib.hmpc_context_wait(self.handle))

    def info(self) -> dict:
        d = DeviceInfo()
        check(_lib.hmpc_context_info(self.handle, ctypes.byref(d)))
        return {f: (getattr(d, f).decode() if isinstance(getattr(d, f), bytes) else getattr(d, f)) for f, _ in d._fields_}

    @property
    def rng(self):
        k = np.zeros(8, np.uint32)
        n = np.zeros(3, np.uint32)
        check(_lib.hmpc_context_get_rng(self.handle, _p(k), _p(n)))
        return k, n

    def set_rng(self, key=None, nonce=None):
        k = _arr(key, 8) if key is not None else None
        n = _arr(nonce, 3) if nonce is not None else None
        check(_lib.hmpc_context_set_rng(self.handle, _p(k) if k is not None else None, _p(n) if n is not None else None))

    def buffer(self, nbytes: int) -> Buffer:
        return Buffer(self, nbytes)

    def upload(self, a: np.ndarray) -> Buffer:
        return Buffer.from_numpy(self, a)

    def set_timing(self, enabled: bool):
        """Record HIP events around every fast-kernel launch (clears earlier records)."""
        check(_lib.hmpc_context_set_timing(self.handle, int(enabled)))

    def kernel_times(self) -> dict:
        """{kernel name: (launches, average ms)} of the launches recorded since set_timing."""
        out = {}
        name = ctypes.create_string_buffer(128)
        i = 0
        while True:
            n, ms = _sz(), ctypes.c_double()
            if _lib.hmpc_context_timing(self.handle, i, name, 128, ctypes.byref(n), ctypes.byref(ms)) != 0:
                break
            out[name.value.decode()] = (n.value, ms.value)
            i += 1
        return out

    def stream(self) -> "Stream":
        return Stream(self)

    def event(self) -> "Event":
        return Event(self)

    def sample_words(self, key, offset: int, count: int, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or Buffer(self, 4 * count)
        k = _arr(key, 8)
        check(_lib.hmpc_sample_words(self.handle, _p(k), offset, count, out.ptr, _sp(stream)))
        return out

    def cipher_encrypt(self, limbs: int, count: int, key, nonce, src: Buffer, out: Buffer | None = None, counter: int = 0,
                       stream=None) -> Buffer:
        """limb-planar u32[limbs][count] -> element-major u32[count][limbs] XOR ChaCha20(key, nonce)"""
        out = out or Buffer(self, 4 * limbs * count)
        k, n = _arr(key, 8), _arr(nonce, 3)
        check(_lib.hmpc_cipher_encrypt(self.handle, limbs, count, _p(k), _p(n), counter, src.ptr, out.ptr, _sp(stream)))
        return out

    def cipher_decrypt(self, limbs: int, count: int, key, nonce, src: Buffer, out: Buffer | None = None, counter: int = 0,
                       stream=None) -> Buffer:
        """element-major u32[count][limbs] -> limb-planar u32[limbs][count]"""
        out = out or Buffer(self, 4 * limbs * count)
        k, n = _arr(key, 8), _arr(nonce, 3)
        check(_lib.hmpc_cipher_decrypt(self.handle, limbs, count, _p(k), _p(n), counter, src.ptr, out.ptr, _sp(stream)))
        return out

    def host_buffer(self, nbytes: int) -> "HostBuffer":
        return HostBuffer(self, nbytes)


class HostBuffer:
    """Page-locked host memory (hmpc_host_alloc) with asynchronous copies to and from the device."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        check(load().hmpc_host_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def array(self, dtype=np.uint32) -> np.ndarray:
        """a numpy view of the pinned memory (valid while this object lives)"""
        n = self.nbytes // np.dtype(dtype).itemsize
        buf = (ctypes.c_char * self.nbytes).from_address(self.ptr)
        return np.frombuffer(buf, dtype=dtype, count=n)

    def to_device(self, dst: Buffer, nbytes: int | None = None, stream=None):
        n = self.nbytes if nbytes is None else int(nbytes)
        check(_lib.hmpc_memcpy_h2d_async(self.ctx.handle, dst.ptr, self.ptr, n, _sp(stream)))

    def from_device(self, src: Buffer, nbytes: int | None = None, stream=None):
        n = self.nbytes if nbytes is None else int(nbytes)
        check(_lib.hmpc_memcpy_d2h_async(self.ctx.handle, self.ptr, src.ptr, n, _sp(stream)))

    def free(self):
        if self.ptr is not None and _lib is not None and self.ctx.handle:
            _lib.hmpc_host_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _sp(stream):
    return None if stream is None else stream.handle


class Stream:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = _vp()
        check(_lib.hmpc_stream_create(ctx.handle, ctypes.byref(h)))
        self.handle = h.value

    def synchronize(self):
        check(_lib.hmpc_stream_synchronize(self.ctx.handle, self.handle))

    def close(self):
        if self.handle and _lib is not None and self.ctx.handle:
            _lib.hmpc_stream_destroy(self.ctx.handle, self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = _vp()
        check(_lib.hmpc_event_create(ctx.handle, ctypes.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        check(_lib.hmpc_event_record(self.ctx.handle, self.handle, _sp(stream)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(_lib.hmpc_event_elapsed_ms(self.ctx.handle, self.handle, end.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.handle and _lib is not None and self.ctx.handle:
                _lib.hmpc_event_destroy(self.ctx.handle, self.handle)
        except Exception:
            pass


class Modulus:
    """A registered modulus q (hmpc::ints::mod<q> at run time) with resident tables."""

    ADD, SUB, MUL = 0, 1, 2

    def __init__(self, ctx: Context, q: int):
        self.ctx = ctx
        self.q = int(q)
        words = max(1, (self.q.bit_length() + 31) // 32)
        qa = int_to_words(self.q, words)
        h = _vp()
        check(_lib.hmpc_modulus_create(ctx.handle, _p(qa), words, ctypes.byref(h)))
        self.handle = h.value
        L, b = _i(), _i()
        check(_lib.hmpc_modulus_info(self.handle, ctypes.byref(L), ctypes.byref(b), None, None, None))
        self.limbs, self.bit_size = L.value, b.value

    def close(self):
        if self.handle and _lib is not None:
            _lib.hmpc_modulus_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- host helpers
    def constants(self):
        npr = np.zeros(1, np.uint32)
        r1 = np.zeros(self.limbs, np.uint32)
        r2 = np.zeros(self.limbs, np.uint32)
        check(_lib.hmpc_modulus_info(self.handle, None, None, _p(npr), _p(r1), _p(r2)))
        return int(npr[0]), words_to_int(r1), words_to_int(r2)

    def set_kernels(self, variant: int):
        """0: fast register-radix kernels where supported, 1: portable LDS radix-2 kernels."""
        check(_lib.hmpc_modulus_set_kernels(self.handle, variant))

    def set_schedule(self, log_n: int, elems_per_thread: int, persistent: int = -1):
        """elements per thread (0 = default); persistent 1 = job kernel, 0 = one launch per pass,
        -1 = unchanged, -2 = default"""
        check(_lib.hmpc_modulus_set_schedule(self.handle, log_n, elems_per_thread, persistent))

    def set_bgv_schedule(self, log_n: int, elems_per_thread: int = 0, tiles: int = -1):
        """fast BGV: elements per thread (0 = default); tiles 1 = two launches over tiles,
        0 = persistent jobs, -1 = default"""
        check(_lib.hmpc_modulus_set_bgv_schedule(self.handle, log_n, elems_per_thread, tiles))

    def schedule(self, log_n: int):
        e, p, f = _i(), _i(), _i()
        check(_lib.hmpc_modulus_get_schedule(self.handle, log_n, ctypes.byref(e), ctypes.byref(p), ctypes.byref(f)))
        return {"elems_per_thread": e.value, "persistent": bool(p.value), "fast": bool(f.value)}

    def autotune_bgv(self, log_n: int, n: int, stream=None) -> dict:
        e, t, ms = _i(), _i(), ctypes.c_float()
        check(_lib.hmpc_autotune_bgv(self.ctx.handle, self.handle, log_n, n, _sp(stream), ctypes.byref(e), ctypes.byref(t),
                                     ctypes.byref(ms)))
        return {"elems_per_thread": e.value, "tiles": bool(t.value), "ms": ms.value}

    def autotune(self, log_n: int, pairs: int, stream=None) -> dict:
        e, p, ms = _i(), _i(), ctypes.c_float()
        check(_lib.hmpc_autotune(self.handle, log_n, pairs, _sp(stream), ctypes.byref(e), ctypes.byref(p), ctypes.byref(ms)))
        return {"elems_per_thread": e.value, "persistent": bool(p.value), "ms": ms.value}

    def prepare(self, log_n: int):
        check(_lib.hmpc_modulus_prepare(self.handle, log_n))

    def root(self, log_n: int, inverse: bool = False) -> int:
        out = np.zeros(self.limbs, np.uint32)
        check(_lib.hmpc_modulus_root(self.handle, log_n, int(inverse), _p(out)))
        return words_to_int(out)

    def twiddles(self, log_n: int, inverse: bool = False) -> np.ndarray:
        p = _vp()
        check(_lib.hmpc_modulus_twiddles(self.handle, log_n, int(inverse), ctypes.byref(p)))
        out = np.empty((self.limbs, 1 << log_n), np.uint32)
        check(_lib.hmpc_memcpy_d2h(self.ctx.handle, out.ctypes.data, p.value, out.nbytes))
        return out

    def mont(self, value: int) -> np.ndarray:
        """Montgomery form value*R mod q as L host words."""
        R = 1 << (32 * self.limbs)
        return int_to_words(int(value) % self.q * R % self.q, self.limbs)

    # ---------------------------------------------------------------- tensors ([L][count] u32)
    def upload(self, planes: np.ndarray) -> Buffer:
        assert planes.shape[0] == self.limbs
        return Buffer.from_numpy(self.ctx, np.ascontiguousarray(planes, dtype=np.uint32))

    def download(self, buf: Buffer, count: int) -> np.ndarray:
        return buf.to_numpy((self.limbs, count))

    def alloc(self, count: int) -> Buffer:
        return Buffer(self.ctx, 4 * self.limbs * count)

    def ntt(self, log_n, n, x: Buffer, out: Buffer | None = None, inverse=False, plain=False, stream=None) -> Buffer:
        out = out or self.alloc(n << log_n)
        if plain:
            check(_lib.hmpc_ntt_plain(self.handle, log_n, n, int(inverse), x.ptr, out.ptr, None, _sp(stream)))
        else:
            check(_lib.hmpc_ntt(self.handle, log_n, n, int(inverse), x.ptr, out.ptr, _sp(stream)))
        return out

    def pointwise(self, op, count, a: Buffer, b: Buffer, b_count=None, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(count)
        check(_lib.hmpc_pointwise(self.handle, op, count, a.ptr, b.ptr, b_count or count, out.ptr, _sp(stream)))
        return out

    def pointwise_scalar(self, op, count, a: Buffer, c_mont: np.ndarray, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(count)
        c = np.ascontiguousarray(c_mont, dtype=np.uint32)
        check(_lib.hmpc_pointwise_scalar(self.handle, op, count, a.ptr, _p(c), out.ptr, _sp(stream)))
        return out

    def ciphertext_op(self, op, count, a0: Buffer, a1: Buffer, b0: Buffer, b1: Buffer, b_count=None,
                      out0: Buffer | None = None, out1: Buffer | None = None, stream=None):
        """(a0 op b0, a1 op b1) for ciphertexts a, b (or a plaintext b0 = b1 broadcast with period b_count)"""
        out0 = out0 or self.alloc(count)
        out1 = out1 or self.alloc(count)
        check(_lib.hmpc_ciphertext_op(self.handle, op, count, a0.ptr, a1.ptr, b0.ptr, b1.ptr,
                                      count if b_count is None else b_count, out0.ptr, out1.ptr, _sp(stream)))
        return out0, out1

    def ciphertext_scalar(self, op, count, a0: Buffer, a1: Buffer, c_mont: np.ndarray, out0: Buffer | None = None,
                          out1: Buffer | None = None, stream=None):
        out0 = out0 or self.alloc(count)
        out1 = out1 or self.alloc(count)
        c = np.ascontiguousarray(c_mont, dtype=np.uint32)
        check(_lib.hmpc_ciphertext_scalar(self.handle, op, count, a0.ptr, a1.ptr, _p(c), out0.ptr, out1.ptr, _sp(stream)))
        return out0, out1

    def fma(self, count, a: Buffer, b: Buffer, c: Buffer, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(count)
        check(_lib.hmpc_fma(self.handle, count, a.ptr, b.ptr, c.ptr, out.ptr, _sp(stream)))
        return out

    def reduce(self, op, count, x: Buffer, out: Buffer | None = None, stream=None) -> Buffer:
        """sum (op ADD) or product (op MUL) of `count` elements, as one element"""
        out = out or self.alloc(1)
        check(_lib.hmpc_reduce(self.handle, op, count, x.ptr, out.ptr, _sp(stream)))
        return out

    def matvec(self, batch, rows, cols, M: Buffer, v: Buffer, broadcast_matrix=False, broadcast_vector=False,
               out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(batch * rows)
        check(_lib.hmpc_matvec(self.handle, batch, rows, cols, M.ptr, int(broadcast_matrix), v.ptr, int(broadcast_vector),
                               out.ptr, _sp(stream)))
        return out

    def from_u32(self, count, words: Buffer, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(count)
        check(_lib.hmpc_from_u32(self.handle, count, words.ptr, out.ptr, _sp(stream)))
        return out

    def to_unsigned(self, count, x: Buffer, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(count)
        check(_lib.hmpc_to_unsigned(self.handle, count, x.ptr, out.ptr, _sp(stream)))
        return out

    def sample_uniform(self, key, offset, count, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(count)
        k = _arr(key, 8)
        check(_lib.hmpc_sample_uniform(self.handle, _p(k), offset, count, out.ptr, _sp(stream)))
        return out

    def sample_cbd(self, key, offset, bits, count, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(count)
        k = _arr(key, 8)
        check(_lib.hmpc_sample_cbd(self.handle, _p(k), offset, bits, count, out.ptr, _sp(stream)))
        return out

    def sample_cbd_ntt(self, bits, log_n, n, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(n << log_n)
        check(_lib.hmpc_sample_cbd_ntt(self.ctx.handle, self.handle, bits, log_n, n, out.ptr, _sp(stream)))
        return out

    def polymul(self, log_n, n, a: Buffer, b: Buffer, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or self.alloc(n << log_n)
        check(_lib.hmpc_polymul(self.handle, log_n, n, a.ptr, b.ptr, out.ptr, _sp(stream)))
        return out

    def polymul_noalloc(self, log_n, n, a: Buffer, b: Buffer, out: Buffer, stream=None) -> Buffer:
        """hmpc_polymul_noalloc: the caller passes every buffer; the library allocates nothing."""
        check(_lib.hmpc_polymul_noalloc(self.handle, log_n, n, a.ptr, b.ptr, out.ptr, _sp(stream)))
        return out

    def bgv_encrypt(self, log_n, n, pk_a: Buffer, pk_b: Buffer, messages: Buffer, p_mont: np.ndarray,
                    bits=(2, 40, 40), c0: Buffer | None = None, c1: Buffer | None = None, plain=False, stream=None):
        c0 = c0 or self.alloc(n << log_n)
        c1 = c1 or self.alloc(n << log_n)
        pm = np.ascontiguousarray(p_mont, dtype=np.uint32)
        fn = _lib.hmpc_bgv_encrypt_plain if plain else _lib.hmpc_bgv_encrypt
        check(fn(self.ctx.handle, self.handle, log_n, n, pk_a.ptr, pk_b.ptr, messages.ptr, _p(pm),
                 bits[0], bits[1], bits[2], c0.ptr, c1.ptr, _sp(stream)))
        return c0, c1

    def bgv_decrypt(self, log_n, n, s_ntt: Buffer, c0: Buffer, c1: Buffer, out: Buffer | None = None, stream=None) -> Buffer:
        out = out or Buffer(self.ctx, 4 * (n << log_n))
        check(_lib.hmpc_bgv_decrypt(self.handle, log_n, n, s_ntt.ptr, c0.ptr, c1.ptr, out.ptr, _sp(stream)))
        return out

    # -- hmpc's general plaintext path
    DISTRIBUTIONS = {"uniform": 0, "centered_binomial": 1, "binomial": 2, "drown_signed": 3, "drown_unsigned": 4}
    LAYOUTS = {"element": 0, "ntt": 1, "intt": 2}

    def cast(self, to: "Modulus", count, x: Buffer, out: Buffer | None = None, stream=None) -> Buffer:
        """mod<to>(mod<self>) element-wise with the centred lift (hmpc_cast)."""
        out = out or to.alloc(count)
        check(_lib.hmpc_cast(self.handle, to.handle, count, x.ptr, out.ptr, _sp(stream)))
        return out

    def sample_capability(self, dist: str, bits: int, layout: str, log_n: int, n: int, out: Buffer | None = None,
                          stream=None) -> Buffer:
        out = out or self.alloc(n << log_n)
        check(_lib.hmpc_sample_capability(self.ctx.handle, self.handle, self.DISTRIBUTIONS[dist], bits, self.LAYOUTS[layout],
                                          log_n, n, out.ptr, _sp(stream)))
        return out

    def bgv_encrypt_plaintext(self, pmod: "Modulus", log_n, n, pk_a: Buffer, pk_b: Buffer, plaintext: Buffer,
                              plaintext_ntt: bool, bits_u=2, v_dist="centered_binomial", bits_v=40, bits_w=40,
                              c0: Buffer | None = None, c1: Buffer | None = None, stream=None):
        c0 = c0 or self.alloc(n << log_n)
        c1 = c1 or self.alloc(n << log_n)
        check(_lib.hmpc_bgv_encrypt_plaintext(self.ctx.handle, self.handle, pmod.handle, log_n, n, pk_a.ptr, pk_b.ptr,
                                              plaintext.ptr, int(plaintext_ntt), bits_u, self.DISTRIBUTIONS[v_dist], bits_v,
                                              bits_w, c0.ptr, c1.ptr, _sp(stream)))
        return c0, c1

    def bgv_decrypt_plaintext(self, pmod: "Modulus", log_n, n, s_ntt: Buffer, c0: Buffer, c1: Buffer, plaintext_ntt: bool,
                              out: Buffer | None = None, stream=None) -> Buffer:
        out = out or pmod.alloc(n << log_n)
        check(_lib.hmpc_bgv_decrypt_plaintext(self.handle, pmod.handle, log_n, n, s_ntt.ptr, c0.ptr, c1.ptr,
                                              int(plaintext_ntt), out.ptr, _sp(stream)))
        return out


def drown_bits(bound: int, signed: bool, security: int = 80) -> int:
    """hmpc_drown_bits: width of drown_uniform's draw for a bound."""
    words = max(1, (int(bound).bit_length() + 31) // 32)
    b = _arr([(int(bound) >> (32 * i)) & 0xFFFFFFFF for i in range(words)], words)
    out = ctypes.c_int(0)
    check(load().hmpc_drown_bits(_p(b), words, int(bool(signed)), security, ctypes.byref(out)))
    return out.value


class Batch:
    """A list of independent jobs (products, encryptions) of uneven degree and modulus run by
    persistent kernels, optionally replayed from a captured hipGraph."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = _vp()
        check(_lib.hmpc_batch_create(ctx.handle, ctypes.byref(h)))
        self.handle = h.value
        self._keep = []

    def add_polymul(self, mod: Modulus, log_n: int, a: Buffer, b: Buffer, out: Buffer):
        check(_lib.hmpc_batch_add_polymul(self.handle, mod.handle, log_n, a.ptr, b.ptr, out.ptr))
        self._keep += [a, b, out]

    def add_bgv_encrypt(self, mod: Modulus, log_n: int, pk_a: Buffer, pk_b: Buffer, messages: Buffer, p_mont,
                        c0: Buffer, c1: Buffer, bits=(2, 40, 40)):
        pm = np.ascontiguousarray(p_mont, dtype=np.uint32)
        check(_lib.hmpc_batch_add_bgv_encrypt(self.handle, mod.handle, log_n, pk_a.ptr, pk_b.ptr, messages.ptr, _p(pm),
                                              bits[0], bits[1], bits[2], c0.ptr, c1.ptr))
        self._keep += [pk_a, pk_b, messages, c0, c1]

    def run(self, graph: bool = True, stream=None):
        check(_lib.hmpc_batch_run(self.handle, int(graph), _sp(stream)))

    def info(self) -> dict:
        j, c, b = _sz(), _sz(), _sz()
        check(_lib.hmpc_batch_info(self.handle, ctypes.byref(j), ctypes.byref(c), ctypes.byref(b)))
        return {"jobs": j.value, "classes": c.value, "device_bytes": b.value}

    def close(self):
        if self.handle and _lib is not None:
            _lib.hmpc_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class View:
    """Non-owning device pointer (offset into a buffer, or memory owned by the server)."""

    def __init__(self, ctx: Context, ptr: int):
        self.ctx = ctx
        self.ptr = ptr

    def to_numpy(self, shape, dtype=np.uint32) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            check(load().hmpc_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out


class Server:
    """Continuous-batching request loop (hmpc_server_*): encryption and product requests admitted
    within an HBM budget and packed into overlapping persistent launches."""

    QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, ctx: Context, budget_bytes: int, slots: int = 3):
        self.ctx = ctx
        h = _vp()
        check(_lib.hmpc_server_create(ctx.handle, budget_bytes, slots, ctypes.byref(h)))
        self.handle = h.value
        self._keep = []
        self._ids = (_u64 * 4096)()
        self._lat = (ctypes.c_double * 4096)()

    def add_key(self, mod: Modulus, log_n: int, pk_a: Buffer, pk_b: Buffer, p_mont, bits=(2, 40, 40)) -> int:
        k = _i()
        pm = np.ascontiguousarray(p_mont, dtype=np.uint32)
        check(_lib.hmpc_server_add_key(self.handle, mod.handle, log_n, pk_a.ptr, pk_b.ptr, _p(pm), bits[0], bits[1], bits[2],
                                       ctypes.byref(k)))
        self._keep += [pk_a, pk_b]
        return k.value

    def submit_encrypt(self, key: int, k: int, messages_ptr: int, nonce) -> int:
        r = _u64()
        nz = _arr(nonce, 3)
        check(_lib.hmpc_server_submit_encrypt(self.handle, key, k, messages_ptr, _p(nz), ctypes.byref(r)))
        return r.value

    def submit_polymul(self, mod: Modulus, log_n: int, n: int, a_ptr: int, b_ptr: int) -> int:
        r = _u64()
        check(_lib.hmpc_server_submit_polymul(self.handle, mod.handle, log_n, n, a_ptr, b_ptr, ctypes.byref(r)))
        return r.value

    def poll(self, request: int):
        st, lat = _i(), ctypes.c_double()
        check(_lib.hmpc_server_poll(self.handle, request, ctypes.byref(st), ctypes.byref(lat)))
        return st.value, lat.value

    def wait(self, request: int):
        check(_lib.hmpc_server_wait(self.handle, request))

    def result(self, request: int):
        c0, c1 = _vp(), _vp()
        check(_lib.hmpc_server_result(self.handle, request, ctypes.byref(c0), ctypes.byref(c1)))
        return View(self.ctx, c0.value), (View(self.ctx, c1.value) if c1.value else None)

    def release(self, request: int):
        check(_lib.hmpc_server_release(self.handle, request))

    def completed(self):
        """[(request, latency_us)] finished since the last call"""
        out = []
        while True:
            n = _sz()
            check(_lib.hmpc_server_completed(self.handle, self._ids, self._lat, 4096, ctypes.byref(n)))
            out += [(self._ids[i], self._lat[i]) for i in range(n.value)]
            if n.value < 4096:
                return out

    def stats(self) -> dict:
        q, r, b = _sz(), _sz(), _sz()
        la, jo = _u64(), _u64()
        check(_lib.hmpc_server_stats(self.handle, ctypes.byref(q), ctypes.byref(r), ctypes.byref(b), ctypes.byref(la),
                                     ctypes.byref(jo)))
        return {"queued": q.value, "running": r.value, "bytes_in_use": b.value, "launches": la.value, "jobs": jo.value}

    def close(self):
        if self.handle and _lib is not None:
            _lib.hmpc_server_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
