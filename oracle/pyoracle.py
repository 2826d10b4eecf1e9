"""ORACLE — test infrastructure only.  ctypes loader for liboracle.so, the C
restatement of the reference path (oracle/*.c).  Imported only by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, always as the
checker or the timed CPU baseline, never as the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None
_sz, _vp, _i, _u64 = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64


class OFr(ctypes.Structure):
    _fields_ = [("l", ctypes.c_uint64 * 4)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        sig = {
            "or_set_threads": ([_i], None),
            "or_get_threads": ([], _i),
            "or_root_of_unity": ([_sz], OFr),
            "or_roots_table": ([_sz, _vp], None),
            "or_ntt_nr": ([_vp, _sz, _sz, _sz], None),
            "or_intt_rn": ([_vp, _sz], None),
            "or_intt_nr": ([_vp, _sz], None),
            "or_reverse_order": ([_vp, _sz], None),
            "or_rs_next_order": ([_sz, ctypes.POINTER(_sz)], _i),
            "or_rs_evaluation_points": ([_sz, _vp, _sz, _vp], None),
            "or_rs_interleaved_encode": ([_vp, _sz, _sz, _vp, _sz, _sz, _vp], _i),
            "or_sky_sbox": ([ctypes.c_uint8], ctypes.c_uint8),
            "or_sky_ss": ([_i, _vp, _vp, _vp, _vp], None),
            "or_sky_bb": ([_i, _vp, _vp, _vp, _vp], None),
            "or_sky_permute": ([_vp, _vp, _vp, _vp], None),
            "or_sky_compress_spec": ([_vp, _vp, _vp], None),
            "or_sky_compress": ([_vp, _vp, _vp], None),
            "or_scalar_sqr": ([_vp, _vp], None),
            "or_scalar_mul": ([_vp, _vp, _vp], None),
            "or_whir_r1cs_params": ([_sz, _sz, _sz, ctypes.POINTER(_sz), ctypes.POINTER(_sz)], None),
            "or_whir_r1cs_shape": ([_vp, _sz, _sz, _vp, _sz, _vp], _i),
            "or_sky_compress_many": ([_vp, _vp, _sz], None),
            "or_sky_hash_many": ([_sz, _vp, _sz, _vp], _i),
            "or_f64_to_u256": ([ctypes.c_double, _vp], None),
            "or_pow_threshold": ([ctypes.c_double, _vp], _i),
            "or_pow_verify": ([_vp, ctypes.c_double, _u64], _i),
            "or_pow_solve": ([_vp, ctypes.c_double, _u64, ctypes.POINTER(_u64)], _i),
            "or_merkle_tree": ([_vp, _sz, _vp], None),
            "or_commit": ([_vp, _sz, _sz, _i, _sz, _vp, _vp], _i),
            "or_p2_permutation": ([_vp, _vp], None),
            "or_p2_hash": ([_vp, _sz], OFr),
            "or_p2_merkle_levels": ([_vp, _sz, _vp], None),
            "or_eq_evals": ([_vp, _sz, _sz, _vp], None),
            "or_calculate_eq": ([_vp, _vp, _sz], OFr),
            "or_sparse_mul": ([_sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp], None),
            "or_sparse_transpose": ([_sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp], _i),
            "or_varint_decode": ([_vp, _sz, _i, _i, ctypes.POINTER(_u64), ctypes.POINTER(_sz)], _i),
            "or_interner_decode": ([_vp, _sz, _vp, _sz, ctypes.POINTER(_sz)], _i),
            "or_sparse_matrix_decode": ([_vp, _sz, _vp, _sz, _sz, _sz, _vp, _vp, _vp,
                                         ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_sz)], _i),
            "or_r1cs_decode": ([_vp, _sz, _vp, _vp, _vp, _vp, _vp], _i),
            "or_sumcheck_r1cs_round": ([_vp, _vp, _vp, _vp, _sz, _vp, _vp], None),
            "or_eval_cubic": ([_vp, OFr], OFr),
            "or_sumcheck_prove": ([_vp, _vp, _vp, _sz, _vp, _u64, _vp, _vp, _vp, _vp, _vp], _i),
            "or_sumcheck_plant_fault": ([_i], None),
            "or_transcript_run": ([_u64, _vp, _sz, _vp, _sz], None),
            "or_fr_mul": ([OFr, OFr], OFr),
            "or_fr_add": ([OFr, OFr], OFr),
            "or_fr_sub": ([OFr, OFr], OFr),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _a(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.uint64)


def _p(a: np.ndarray):
    return a.ctypes.data if a is not None and a.size else None


def _fr(o: OFr) -> np.ndarray:
    return np.array(list(o.l), dtype=np.uint64)


def _ofr(x) -> OFr:
    o = OFr()
    a = np.asarray(x, dtype=np.uint64).reshape(4)
    for i in range(4):
        o.l[i] = int(a[i])
    return o


def fr_mul(a, b) -> np.ndarray:
    return _fr(lib().or_fr_mul(_ofr(a), _ofr(b)))


def fr_add(a, b) -> np.ndarray:
    return _fr(lib().or_fr_add(_ofr(a), _ofr(b)))


def fr_sub(a, b) -> np.ndarray:
    return _fr(lib().or_fr_sub(_ofr(a), _ofr(b)))


def set_threads(n: int) -> None:
    lib().or_set_threads(n)


def threads() -> int:
    return lib().or_get_threads()


# ---- ntt
def root_of_unity(n: int) -> np.ndarray:
    return _fr(lib().or_root_of_unity(n))


def roots_table(order: int) -> np.ndarray:
    out = np.zeros((max(order // 2, 0), 4), dtype=np.uint64)
    lib().or_roots_table(order, _p(out))
    return out


def ntt_nr(values: np.ndarray, codeword_size: int, num_groups: int) -> None:
    lib().or_ntt_nr(_p(values), values.shape[0], codeword_size, num_groups)


def intt_rn(values: np.ndarray) -> None:
    lib().or_intt_rn(_p(values), values.shape[0])


def intt_nr(values: np.ndarray) -> None:
    lib().or_intt_nr(_p(values), values.shape[0])


def reverse_order(values: np.ndarray, codeword_size: Optional[int] = None) -> None:
    n = values.shape[0] if codeword_size is None else codeword_size
    assert n <= values.shape[0]
    lib().or_reverse_order(_p(values), n)


def rs_next_order(size: int) -> Optional[int]:
    o = _sz()
    return o.value if lib().or_rs_next_order(size, ctypes.byref(o)) else None


def rs_evaluation_points(codeword_length: int, indices) -> np.ndarray:
    idx = np.ascontiguousarray(indices, dtype=np.uint64)
    out = np.zeros((len(idx), 4), dtype=np.uint64)
    lib().or_rs_evaluation_points(codeword_length, _p(idx), len(idx), _p(out))
    return out


def rs_interleaved_encode(messages, masks, codeword_length: int) -> np.ndarray:
    M = len(messages)
    if M == 0:
        return np.zeros((0, 4), np.uint64)
    msgs = _a(np.concatenate([np.asarray(m, dtype=np.uint64).reshape(-1, 4) for m in messages]))
    mlen = msgs.shape[0] // M
    mk = _a(masks).reshape(-1, 4) if masks is not None and len(masks) else np.zeros((0, 4), np.uint64)
    out = np.zeros((M * codeword_length, 4), dtype=np.uint64)
    rc = lib().or_rs_interleaved_encode(_p(msgs), M, mlen, _p(mk), mk.shape[0], codeword_length, _p(out))
    if rc:
        raise ValueError("interleaved_encode failed")
    return out


# ---- skyscraper (raw 256-bit integers as 4 LE u64 limbs)
def sbox(v: int) -> int:
    return lib().or_sky_sbox(v)


def _pair(fn, *args):
    ol = np.zeros(4, np.uint64)
    orr = np.zeros(4, np.uint64)
    fn(*args, _p(ol), _p(orr))
    return ol, orr


# (argument arrays are bound to locals so they outlive the foreign call)
def sky_ss(rnd: int, l, r):
    la, ra = _a(l), _a(r)
    return _pair(lib().or_sky_ss, rnd, _p(la), _p(ra))


def sky_bb(rnd: int, l, r):
    la, ra = _a(l), _a(r)
    return _pair(lib().or_sky_bb, rnd, _p(la), _p(ra))


def sky_permute(l, r):
    la, ra = _a(l), _a(r)
    return _pair(lib().or_sky_permute, _p(la), _p(ra))


def sky_compress_spec(l, r) -> np.ndarray:
    la, ra = _a(l), _a(r)
    o = np.zeros(4, np.uint64)
    lib().or_sky_compress_spec(_p(la), _p(ra), _p(o))
    return o


def sky_compress(l, r) -> np.ndarray:
    la, ra = _a(l), _a(r)
    o = np.zeros(4, np.uint64)
    lib().or_sky_compress(_p(la), _p(ra), _p(o))
    return o


def scalar_sqr(a) -> np.ndarray:
    aa = _a(a)
    o = np.zeros(4, np.uint64)
    lib().or_scalar_sqr(_p(aa), _p(o))
    return o


def scalar_mul(a, b) -> np.ndarray:
    aa, bb = _a(a), _a(b)
    o = np.zeros(4, np.uint64)
    lib().or_scalar_mul(_p(aa), _p(bb), _p(o))
    return o


def whir_r1cs_params(num_witnesses: int, num_constraints: int, w1_size: int):
    m, m0 = ctypes.c_size_t(), ctypes.c_size_t()
    lib().or_whir_r1cs_params(num_witnesses, num_constraints, w1_size, ctypes.byref(m),
                              ctypes.byref(m0))
    return m.value, m0.value


def whir_r1cs_shape(witness, num_vars: int, blinding=None) -> np.ndarray:
    """The witness shaping of WhirR1CSScheme::commit -> (2^num_vars, 4)."""
    w = _a(witness).reshape(-1, 4)
    g = _a(blinding).reshape(-1, 4) if blinding is not None else None
    out = np.zeros((1 << num_vars, 4), np.uint64)
    rc = lib().or_whir_r1cs_shape(_p(w), w.shape[0], num_vars, _p(g),
                                  g.shape[0] // 4 if g is not None else 0, _p(out))
    if rc == -1:
        raise ValueError("R1CS witness length exceeds scheme capacity")
    if rc == -2:
        raise IndexError("blinding coefficients run past the padded witness")
    return out


def sky_compress_many(messages: bytes) -> bytes:
    m = np.frombuffer(messages, np.uint8)
    out = np.zeros(len(m) // 2, np.uint8)
    lib().or_sky_compress_many(_p(m), _p(out), len(m) // 64)
    return out.tobytes()


def sky_hash_many(size: int, data: bytes, count: int) -> bytes:
    m = np.frombuffer(data, np.uint8)
    out = np.zeros(32 * count, np.uint8)
    if lib().or_sky_hash_many(size, _p(m), count, _p(out)):
        raise ValueError("unsupported size")
    return out.tobytes()


def f64_to_u256(f: float) -> np.ndarray:
    o = np.zeros(4, np.uint64)
    lib().or_f64_to_u256(float(f), _p(o))
    return o


def pow_threshold(difficulty: float) -> np.ndarray:
    o = np.zeros(4, np.uint64)
    if lib().or_pow_threshold(float(difficulty), _p(o)):
        raise ValueError("Difficulty must be in the range [0, 80)")
    return o


def pow_verify(challenge, difficulty: float, nonce: int) -> bool:
    ch = _a(challenge)
    r = lib().or_pow_verify(_p(ch), float(difficulty), int(nonce))
    if r < 0:
        raise ValueError("Difficulty must be in the range [0, 80)")
    return bool(r)


def pow_solve(challenge, difficulty: float, limit: int = 1 << 40) -> Optional[int]:
    n = _u64(0)
    ch = _a(challenge)
    r = lib().or_pow_solve(_p(ch), float(difficulty), int(limit), ctypes.byref(n))
    if r < 0:
        raise ValueError("Difficulty must be in the range [0, 80)")
    return int(n.value) if r else None


def merkle_tree(leaf_digests: np.ndarray) -> np.ndarray:
    ld = np.ascontiguousarray(leaf_digests, dtype=np.uint8).reshape(-1, 32)
    n = ld.shape[0]
    nodes = np.zeros((2 * n - 1, 32), np.uint8)
    lib().or_merkle_tree(_p(ld), n, _p(nodes))
    return nodes


def commit(coeffs: np.ndarray, blowup: int, coset: bool, leaf_elems: int):
    c = _a(coeffs).reshape(-1, 4)
    n = c.shape[0]
    N = n * blowup
    cw = np.zeros((N, 4), np.uint64)
    leaves = N // leaf_elems
    nodes = np.zeros((2 * leaves - 1, 32), np.uint8)
    if lib().or_commit(_p(c), n, blowup, int(coset), leaf_elems, _p(cw), _p(nodes)):
        raise ValueError("commit failed")
    return nodes[-1].tobytes(), cw, nodes


# ---- poseidon2 (Montgomery field elements)
def p2_permutation(state) -> np.ndarray:
    s = _a(state).reshape(4, 4)
    o = np.zeros((4, 4), np.uint64)
    lib().or_p2_permutation(_p(s), _p(o))
    return o


def p2_hash(inputs) -> np.ndarray:
    x = _a(inputs).reshape(-1, 4)
    return _fr(lib().or_p2_hash(_p(x), x.shape[0]))


def p2_merkle_levels(leaves) -> np.ndarray:
    x = _a(leaves).reshape(-1, 4)
    n = x.shape[0]
    out = np.zeros((2 * n - 1, 4), np.uint64)
    lib().or_p2_merkle_levels(_p(x), n, _p(out))
    return out


# ---- sumcheck / sparse
def eq_evals(r, num_entries: int) -> np.ndarray:
    rr = _a(r).reshape(-1, 4)
    out = np.zeros((num_entries, 4), np.uint64)
    lib().or_eq_evals(_p(rr), rr.shape[0], num_entries, _p(out))
    return out


def calculate_eq(r, alpha) -> np.ndarray:
    a = _a(r).reshape(-1, 4)
    b = _a(alpha).reshape(-1, 4)
    return _fr(lib().or_calculate_eq(_p(a), _p(b), a.shape[0]))


def sparse_mul(num_rows: int, num_cols: int, row_starts, cols, vals, z) -> np.ndarray:
    rs = np.ascontiguousarray(row_starts, dtype=np.uint32)
    cs = np.ascontiguousarray(cols, dtype=np.uint32)
    vs = _a(vals).reshape(-1, 4)
    zz = _a(z).reshape(-1, 4)
    out = np.zeros((num_rows, 4), np.uint64)
    lib().or_sparse_mul(num_rows, num_cols, _p(rs), cs.shape[0], _p(cs), _p(vs), _p(zz), _p(out))
    return out


def sparse_transpose(num_rows: int, num_cols: int, row_starts, cols, vals):
    """(t_row_starts, t_cols, t_vals) of the transposed matrix (num_cols rows)."""
    rs = np.ascontiguousarray(row_starts, dtype=np.uint32)
    cs = np.ascontiguousarray(cols, dtype=np.uint32)
    vs = np.ascontiguousarray(vals, dtype=np.uint64).reshape(-1, 4)
    nnz = cs.shape[0]
    trs = np.zeros(num_cols, np.uint32)
    tcs = np.zeros(nnz, np.uint32)
    tvs = np.zeros((nnz, 4), np.uint64)
    if lib().or_sparse_transpose(num_rows, num_cols, _p(rs), nnz, _p(cs), _p(vs), _p(trs), _p(tcs), _p(tvs)):
        raise ValueError("column index out of range")
    return trs, tcs, tvs


class WireError(ValueError):
    pass


def varint_decode(data: bytes, max_bytes: int, bits: int):
    """(value, bytes used) of the LEB128 varint at the start of data."""
    b = np.frombuffer(data, np.uint8)
    v, used = _u64(0), _sz(0)
    if lib().or_varint_decode(_p(b), b.size, max_bytes, bits, ctypes.byref(v), ctypes.byref(used)):
        raise WireError("bad varint")
    return int(v.value), int(used.value)


def interner_decode(blob: bytes) -> np.ndarray:
    b = np.frombuffer(blob, np.uint8)
    n = _sz(0)
    if lib().or_interner_decode(_p(b), b.size, None, 0, ctypes.byref(n)):
        raise WireError("malformed interner")
    out = np.zeros((n.value, 4), np.uint64)
    if lib().or_interner_decode(_p(b), b.size, _p(out), n.value, ctypes.byref(n)):
        raise WireError("malformed interner")
    return out


def sparse_matrix_decode(blob: bytes, interner):
    """(num_rows, num_cols, row_starts, cols, vals) of a postcard SparseMatrix."""
    b = np.frombuffer(blob, np.uint8)
    it = _a(interner).reshape(-1, 4)
    nr, nc, nz = _sz(0), _sz(0), _sz(0)
    sz = (ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nz))
    if lib().or_sparse_matrix_decode(_p(b), b.size, _p(it), it.shape[0], 0, 0, None, None, None, *sz):
        raise WireError("malformed sparse matrix")
    rs = np.zeros(max(nr.value, 1), np.uint32)
    cs = np.zeros(max(nz.value, 1), np.uint32)
    vs = np.zeros((max(nz.value, 1), 4), np.uint64)
    if lib().or_sparse_matrix_decode(_p(b), b.size, _p(it), it.shape[0], nr.value, nz.value,
                                     _p(rs), _p(cs), _p(vs), *sz):
        raise WireError("malformed sparse matrix")
    return nr.value, nc.value, rs[: nr.value], cs[: nz.value], vs[: nz.value]


def r1cs_decode(blob: bytes) -> dict:
    b = np.frombuffer(blob, np.uint8)
    sizes = np.zeros(11, np.uint64)
    nulls = (ctypes.c_void_p * 3)()
    if lib().or_r1cs_decode(_p(b), b.size, _p(sizes), nulls, nulls, nulls, None):
        raise WireError("malformed r1cs")
    mats = []
    rs_p, cs_p, vs_p = (ctypes.c_void_p * 3)(), (ctypes.c_void_p * 3)(), (ctypes.c_void_p * 3)()
    caps = np.zeros(6, np.uint64)
    keep = []
    for m in range(3):
        nr, nz = int(sizes[2 + 3 * m]), int(sizes[4 + 3 * m])
        arrs = (np.zeros(max(nr, 1), np.uint32), np.zeros(max(nz, 1), np.uint32),
                np.zeros((max(nz, 1), 4), np.uint64))
        keep.append(arrs)
        rs_p[m], cs_p[m], vs_p[m] = _p(arrs[0]), _p(arrs[1]), _p(arrs[2])
        caps[2 * m], caps[2 * m + 1] = nr, nz
    t0 = time.perf_counter()
    if lib().or_r1cs_decode(_p(b), b.size, _p(sizes), rs_p, cs_p, vs_p, _p(caps)):
        raise WireError("malformed r1cs")
    fill_s = time.perf_counter() - t0
    for m in range(3):
        nr, nc, nz = (int(x) for x in sizes[2 + 3 * m: 5 + 3 * m])
        mats.append((nr, nc, keep[m][0][:nr], keep[m][1][:nz], keep[m][2][:nz]))
    return {"num_public_inputs": int(sizes[0]), "interner_len": int(sizes[1]),
            "a": mats[0], "b": mats[1], "c": mats[2], "fill_seconds": fill_s}


def sumcheck_round(a, b, c, eq, fold=None) -> np.ndarray:
    """One sumcheck_fold_map_reduce with the R1CS map; a, b, c, eq are folded
    in place (C-contiguous uint64 (n, 4) arrays)."""
    out = np.zeros((3, 4), np.uint64)
    f = _a(fold).reshape(4) if fold is not None else None
    lib().or_sumcheck_r1cs_round(_p(a), _p(b), _p(c), _p(eq), a.shape[0], _p(f), _p(out))
    return out


def eval_cubic(poly, x) -> np.ndarray:
    p = _a(poly).reshape(4, 4)
    o = OFr()
    for i in range(4):
        o.l[i] = int(np.asarray(x, dtype=np.uint64).reshape(4)[i])
    return _fr(lib().or_eval_cubic(_p(p), o))


def sumcheck_prove(a, b, c, m: int, blinding=None, seed: int = 0):
    """Full sumcheck (oracle/transcript.c).  Returns dict with coeffs (m,4,4),
    alphas (m,4), r (m,4), final (4,), rho, sum_g, blinding_eval, state (4,)
    each, ok (bool).  m = 0 pads a, b, c with a zero to 2 elements
    (pad_to_pow2_len_min2)."""
    A, B, C = (_a(x).reshape(-1, 4).copy() for x in (a, b, c))
    if m == 0:
        A, B, C = (np.concatenate([x[:1], np.zeros((1, 4), np.uint64)]) for x in (A, B, C))
    coeffs = np.zeros((max(m, 1), 4, 4), np.uint64)
    alphas = np.zeros((max(m, 1), 4), np.uint64)
    rr = np.zeros((max(m, 1), 4), np.uint64)
    fin = np.zeros(4, np.uint64)
    extra = np.zeros((4, 4), np.uint64)
    g = _a(blinding).reshape(-1, 4) if blinding is not None else None
    rc = lib().or_sumcheck_prove(_p(A), _p(B), _p(C), m, _p(g), seed, _p(coeffs), _p(alphas),
                                 _p(rr), _p(fin), _p(extra))
    if rc == -2:
        raise ValueError("blinding polynomial needs m >= 1")
    return {"coeffs": coeffs[:m], "alphas": alphas[:m], "r": rr[:m], "final": fin,
            "rho": extra[0], "sum_g": extra[1], "blinding_eval": extra[2], "state": extra[3],
            "ok": rc == 0}


def plant_fault(kind: int) -> None:
    """Plant a mistake in the oracle's ZK sumcheck (0 removes it); for the
    tests that show the verifier's checks catch it."""
    lib().or_sumcheck_plant_fault(kind)


def transcript_run(seed: int, absorbs, n_squeeze: int) -> np.ndarray:
    ab = _a(absorbs).reshape(-1, 4) if absorbs is not None and len(absorbs) else np.zeros((0, 4), np.uint64)
    out = np.zeros((n_squeeze, 4), np.uint64)
    lib().or_transcript_run(seed, _p(ab), ab.shape[0], _p(out), n_squeeze)
    return out
