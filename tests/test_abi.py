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
