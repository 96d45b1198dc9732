"""CPU-side checks of the C ABI: the library builds, loads and exports every declared symbol."""

import ctypes
import os
import re

import pytest

import hmpc_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hmpc_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hmpc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "hmpc_ntt" in names and "hmpc_bgv_encrypt" in names and len(names) >= 30


def test_library_exports_every_declared_symbol():
    if not os.path.exists(hmpc_amd.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(hmpc_amd.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_symbols()) == set(hmpc_amd.SIGNATURES)


def test_load_and_error_path_without_gpu():
    if not os.path.exists(hmpc_amd.LIB_PATH):
        pytest.skip("library not built")
    lib = hmpc_amd.load()
    # argument validation happens before any device call
    rc = lib.hmpc_context_create(0, None, None)
    assert rc == 1
    assert b"null" in lib.hmpc_last_error()


def test_cpp_interface_compiles(tmp_path):
    """include/hmpc_amd/comp.hpp (the hmpc::comp mirror) compiles as C++17 with warnings as errors."""
    import subprocess

    src = os.path.join(ROOT, "tests", "cpp", "facade_test.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        f"-I{os.path.join(ROOT, 'include')}", src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_argument_validation_without_gpu():
    """null handles and out-of-range arguments are rejected by the host side of every newer entry
    point before any device work (status codes, messages through hmpc_last_error)"""
    if not os.path.exists(hmpc_amd.LIB_PATH):
        pytest.skip("library not built")
    lib = hmpc_amd.load()
    key = (ctypes.c_uint32 * 8)()
    nonce = (ctypes.c_uint32 * 3)()
    one = (ctypes.c_uint32 * 4)(1)
    calls = {
        "hmpc_ciphertext_op": lambda: lib.hmpc_ciphertext_op(None, 0, 16, None, None, None, None, 16, None, None, None),
        "hmpc_ciphertext_scalar": lambda: lib.hmpc_ciphertext_scalar(None, 2, 16, None, None, one, None, None, None),
        "hmpc_cipher_encrypt": lambda: lib.hmpc_cipher_encrypt(None, 4, 16, key, nonce, 0, None, None, None),
        "hmpc_cipher_decrypt": lambda: lib.hmpc_cipher_decrypt(None, 4, 16, key, nonce, 0, None, None, None),
        "hmpc_fma": lambda: lib.hmpc_fma(None, 16, None, None, None, None, None),
        "hmpc_reduce": lambda: lib.hmpc_reduce(None, 0, 16, None, None, None),
        "hmpc_matvec": lambda: lib.hmpc_matvec(None, 1, 4, 4, None, 0, None, 0, None, None),
        "hmpc_host_alloc": lambda: lib.hmpc_host_alloc(None, 64, None),
        "hmpc_memcpy_h2d_async": lambda: lib.hmpc_memcpy_h2d_async(None, None, None, 64, None),
        "hmpc_memcpy_d2h_async": lambda: lib.hmpc_memcpy_d2h_async(None, None, None, 64, None),
    }
    for name, call in calls.items():
        rc = call()
        assert rc != 0, name
        assert lib.hmpc_last_error(), name
