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


def test_transpose_and_eq_argument_errors_without_device():
    # The interned transpose and the interned At * eq(alpha) product check
    # their arguments on the host before any device work.
    import numpy as np

    import provekit_amd as pk
    L = pk.lib()
    u32p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rs, cols, ids = (np.zeros(1, np.uint32), np.array([2], np.uint32), np.zeros(1, np.uint32))
    trs, tcs, tis = np.zeros(2, np.uint32), np.zeros(1, np.uint32), np.zeros(1, np.uint32)
    # column index beyond num_cols (host form validates like SparseMatrix)
    assert L.pkgpu_sparse_transpose_interned(1, 2, u32p(rs), 1, u32p(cols), u32p(ids),
                                             u32p(trs), u32p(tcs), u32p(tis)) == pk.ERR_ARG
    assert b"column index out of range" in L.pkgpu_last_error()
    # decreasing row starts
    rs2, cols2 = np.array([1, 0], np.uint32), np.array([0, 1], np.uint32)
    ids2 = np.zeros(2, np.uint32)
    assert L.pkgpu_sparse_transpose_interned(2, 2, u32p(rs2), 2, u32p(cols2), u32p(ids2),
                                             u32p(trs), u32p(np.zeros(2, np.uint32)),
                                             u32p(np.zeros(2, np.uint32))) == pk.ERR_ARG
    assert b"bad row starts" in L.pkgpu_last_error()
    # entries without ids, outputs missing
    assert L.pkgpu_sparse_transpose_interned(1, 3, u32p(rs), 1, u32p(cols), None,
                                             u32p(trs), u32p(tcs), u32p(tis)) == pk.ERR_ARG
    assert L.pkgpu_sparse_transpose_interned_device(1, 3, u32p(rs), 1, u32p(cols), u32p(ids),
                                                    u32p(trs), u32p(tcs), None, None) == pk.ERR_ARG
    assert b"null output" in L.pkgpu_last_error()
    # the eq table cannot cover more constraints than 2^m
    assert L.pkgpu_multiply_transposed_by_eq_alpha_interned_device(
        3, 5, u32p(np.zeros(3, np.uint32)), 0, None, None, None, 0, None, 2, None, None) == pk.ERR_SIZE
    assert b"num_entries <= full_size" in L.pkgpu_last_error()
    # ids without an interner table
    assert L.pkgpu_multiply_transposed_by_eq_alpha_interned_device(
        1, 1, u32p(rs), 1, u32p(np.zeros(1, np.uint32)), u32p(ids), None, 4, None, 0,
        u32p(np.zeros(8, np.uint32)), None) == pk.ERR_ARG
    assert b"null interner" in L.pkgpu_last_error()
