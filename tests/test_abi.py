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


def test_integration_rust_block_matches_header():
    """INTEGRATION.md's pkgpu-sys extern block declares every entry point of
    the header with the same number of parameters."""
    hdr = open(os.path.join(ROOT, "include", "provekit_gpu.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    c_params = {}
    for m in re.finditer(r"\b(pkgpu_\w+)\s*\(([^;{]*?)\)\s*;", hdr):
        args = " ".join(m.group(2).split())
        c_params[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    a = doc.index('extern "C" {')
    block = doc[a:doc.index("\n}", a)]
    rust_params = {}
    for m in re.finditer(r"pub fn (pkgpu_\w+)\(([^)]*)\)", block):
        args = " ".join(m.group(2).split())
        rust_params[m.group(1)] = 0 if not args else args.count(",") + 1
    assert set(rust_params) == set(c_params), sorted(set(c_params) ^ set(rust_params))
    bad = {n: (c_params[n], rust_params[n]) for n in c_params if c_params[n] != rust_params[n]}
    assert not bad, bad


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
