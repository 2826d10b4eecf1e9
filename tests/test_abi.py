"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pkgpu_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "pkgpu_ntt_nr" in names and "pkgpu_commit" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    import provekit_amd as pk
    if not os.path.exists(pk.LIB_PATH):
        pytest.fail(f"{pk.LIB_PATH} not built (run python -m provekit_amd.build)")
    lib = ctypes.CDLL(pk.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import provekit_amd as pk
    bound = set(pk.SIGNATURES) | {"pkgpu_last_error", "pkgpu_abi_version"}
    missing = [n for n in declared_symbols() if n not in bound]
    assert not missing, missing


def test_abi_version_and_error_string():
    import provekit_amd as pk
    L = pk.lib()
    assert L.pkgpu_abi_version() == 1
    assert isinstance(L.pkgpu_last_error(), bytes)


def test_argument_errors_without_device():
    # Host-side validation runs before any device work, so these fail the same
    # way with or without a GPU.
    import provekit_amd as pk
    L = pk.lib()
    assert L.pkgpu_rs_next_order(5, None) == pk.ERR_ARG
    assert b"null" in L.pkgpu_last_error()
    out = ctypes.c_size_t()
    assert L.pkgpu_rs_next_order((1 << 28) + 1, ctypes.byref(out)) == pk.ERR_SIZE
    assert L.pkgpu_rs_next_order(1000, ctypes.byref(out)) == pk.OK and out.value == 1024
    assert L.pkgpu_skyscraper_hash_many(31, None, 0, None, 0) == pk.ERR_SIZE
    assert b"unsupported message size 31" in L.pkgpu_last_error()
