"""provekit_amd — MI355X (gfx950) implementation of ProveKit's WHIR commitment
hot path, exposed through the C ABI in include/provekit_gpu.h.

This Python module is a thin ctypes binding with the reference's names
(``ntt_nr``, ``intt_rn``, ``RSFr``, ``skyscraper.compress_many``,
``SkyscraperHashEngine``, ...), used by the tests and the benchmark.  Every
call runs the HIP kernels in ``libprovekit_gpu.so``; there is no CPU path.  If
the library is missing, calls raise instead of falling back.

Field elements are numpy ``uint64`` arrays of shape (n, 4): the Montgomery
limbs of ark-ff's ``ark_bn254::Fr``.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libprovekit_gpu.so")

# status codes (include/provekit_gpu.h)
OK, ERR_ARG, ERR_NOT_POW2, ERR_SIZE, ERR_OOM, ERR_DEVICE, ERR_BUDGET, ERR_STATE, ERR_INTERNAL = range(9)


class ProvekitGpuError(RuntimeError):
    """A pkgpu_* call failed; ``code`` is the pkgpu_status, the message is the
    library's (the reference's panic text where the reference panics)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[pkgpu status {code}] {msg}")
        self.code = code


_lib: Optional[ctypes.CDLL] = None

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_i = ctypes.c_int
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)

# name -> argtypes; every function returns int (pkgpu_status) except the two below
SIGNATURES = {
    "pkgpu_init": [_i],
    "pkgpu_device_synchronize": [],
    "pkgpu_malloc": [ctypes.POINTER(_vp), _sz],
    "pkgpu_free": [_vp],
    "pkgpu_memcpy_h2d": [_vp, _vp, _sz, _vp],
    "pkgpu_memcpy_d2h": [_vp, _vp, _sz, _vp],
    "pkgpu_stream_synchronize": [_vp],
    "pkgpu_ntt_nr": [_vp, _sz, _sz, _sz],
    "pkgpu_ntt_nr_device": [_vp, _sz, _sz, _sz, _vp],
    "pkgpu_intt_rn": [_vp, _sz],
    "pkgpu_intt_rn_device": [_vp, _sz, _vp],
    "pkgpu_intt_nr": [_vp, _sz],
    "pkgpu_intt_nr_device": [_vp, _sz, _vp],
    "pkgpu_reverse_order": [_vp, _sz, _sz],
    "pkgpu_reverse_order_device": [_vp, _sz, _sz, _vp],
    "pkgpu_rs_next_order": [_sz, ctypes.POINTER(_sz)],
    "pkgpu_rs_generator": [_sz, _vp],
    "pkgpu_rs_evaluation_points": [_sz, _sz, _vp, _sz, _vp],
    "pkgpu_rs_interleaved_encode": [_vp, _sz, _sz, _vp, _sz, _sz, _vp],
    "pkgpu_rs_interleaved_encode_device": [_vp, _sz, _sz, _vp, _sz, _sz, _i, _vp, _vp],
    "pkgpu_skyscraper_compress": [_vp, _vp, _vp],
    "pkgpu_skyscraper_compress_many": [_vp, _sz, _vp, _sz],
    "pkgpu_skyscraper_hash_many": [_sz, _vp, _sz, _vp, _sz],
    "pkgpu_skyscraper_hash_many_device": [_sz, _vp, _sz, _vp, _vp],
    "pkgpu_skyscraper_merkle_device": [_vp, _sz, _vp, _vp],
    "pkgpu_merkle_open_device": [_vp, _vp, _sz, _sz, _vp, _sz, _vp, _vp, _vp],
    "pkgpu_skyscraper_pow_threshold": [ctypes.c_double, _vp],
    "pkgpu_skyscraper_pow_verify": [_vp, ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(_i)],
    "pkgpu_skyscraper_pow_solve": [_vp, ctypes.c_double, ctypes.c_uint64,
                                   ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_i), _vp],
    "pkgpu_poseidon2_permutation_many": [_vp, _sz, _vp],
    "pkgpu_poseidon2_hash": [_vp, _sz, _vp],
    "pkgpu_poseidon2_hash_many": [_vp, _sz, _sz, _vp],
    "pkgpu_poseidon2_hash_many_device": [_vp, _sz, _sz, _vp, _vp],
    "pkgpu_poseidon2_merkle": [_vp, _sz, _vp, _vp],
    "pkgpu_poseidon2_merkle_device": [_vp, _sz, _vp, _vp],
    "pkgpu_poseidon2_merkle_graph": [_vp, _sz, _vp, ctypes.POINTER(_vp)],
    "pkgpu_graph_launch": [_vp, _vp],
    "pkgpu_graph_destroy": [_vp],
    "pkgpu_sparse_mul": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _sz, _vp],
    "pkgpu_sparse_mul_device": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp],
    "pkgpu_sparse_mul_device_noalloc": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp],
    "pkgpu_sparse_mul_interned": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _sz, _vp, _sz, _vp],
    "pkgpu_sparse_mul_interned_device": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _sz, _vp, _vp, _vp],
    "pkgpu_interner_decode_device": [_vp, _sz, _vp, ctypes.POINTER(_sz), _vp],
    "pkgpu_sparse_matrix_decode_device": [_vp, _sz, _vp, _sz, _vp, _vp],
    "pkgpu_r1cs_decode_device": [_vp, _sz, ctypes.POINTER(_sz), _vp, ctypes.POINTER(_sz), _vp, _vp],
    "pkgpu_sparse_transpose": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp],
    "pkgpu_sparse_transpose_device": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp],
    "pkgpu_multiply_transposed_by_eq_alpha_device": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _sz, _vp,
                                                     _vp],
    "pkgpu_sparse_transpose_interned": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp],
    "pkgpu_sparse_transpose_interned_device": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp],
    "pkgpu_multiply_transposed_by_eq_alpha_interned_device": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _sz,
                                                              _vp, _sz, _vp, _vp],
    "pkgpu_calculate_witness_bounds": [_sz, _sz, _vp, _sz, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _sz,
                                       _vp, _vp, _vp, ctypes.POINTER(_sz)],
    "pkgpu_eq_evaluations": [_vp, _sz, _sz, _vp],
    "pkgpu_eq_evaluations_device": [_vp, _sz, _sz, _vp, _vp],
    "pkgpu_sumcheck_r1cs_round_device": [_vp, _vp, _vp, _vp, _sz, _vp, _vp, _vp],
    "pkgpu_sumcheck_prove_device": [_vp, _vp, _vp, _sz, _vp, ctypes.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pkgpu_ntt_autotune": [_i, _i, _i, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)],
    "pkgpu_ntt_tuning_load": [ctypes.c_char_p],
    "pkgpu_ntt_tuning_get": [_i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "pkgpu_batch_plan_create": [_sz, _vp, _vp, _vp, _vp, _sz, _vp, _vp, _i, ctypes.POINTER(_vp)],
    "pkgpu_batch_run": [_vp, _vp],
    "pkgpu_batch_graph_launch": [_vp, _vp],
    "pkgpu_batch_roots": [_vp, _vp, _vp],
    "pkgpu_batch_plan_destroy": [_vp],
    "pkgpu_server_create": [_i, _sz, _i, ctypes.POINTER(_vp)],
    "pkgpu_server_submit": [_vp, _vp, ctypes.POINTER(ctypes.c_uint64)],
    "pkgpu_server_wait": [_vp, ctypes.c_uint64, ctypes.c_double, _vp, _vp, _vp, _vp],
    "pkgpu_server_poll": [_vp, ctypes.c_uint64, ctypes.POINTER(_i)],
    "pkgpu_server_open": [_vp, ctypes.c_uint64, _vp, _sz, _vp, _vp],
    "pkgpu_server_release": [_vp, ctypes.c_uint64],
    "pkgpu_server_stats": [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)],
    "pkgpu_server_destroy": [_vp],
    "pkgpu_commit": [_vp, _sz, _sz, _i, _sz, _vp, _vp, _vp],
    "pkgpu_commit_device": [_vp, _sz, _sz, _i, _sz, _vp, _vp, _vp],
    "pkgpu_whir_r1cs_scheme_size": [_sz, _sz, _sz, ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "pkgpu_whir_r1cs_commit": [_vp, _sz, _sz, _vp, _sz, _sz, _i, _sz, _vp, _vp, _vp, _vp],
    "pkgpu_whir_r1cs_commit_device": [_vp, _sz, _sz, _vp, _sz, _sz, _i, _sz, _vp, _vp, _vp, _vp],
    "pkgpu_random_fr_device": [_vp, ctypes.c_uint64, _sz, _vp, _vp],
    "pkgpu_scalar_mul_many": [_vp, _vp, _vp, _sz],
    "pkgpu_scalar_sqr_many": [_vp, _vp, _sz],
    "pkgpu_scalar_mul_many_device": [_vp, _vp, _vp, _sz, _vp],
    "pkgpu_scalar_sqr_many_device": [_vp, _vp, _sz, _vp],
}


def lib() -> ctypes.CDLL:
    """Load libprovekit_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ProvekitGpuError(
                ERR_INTERNAL,
                f"{LIB_PATH} is missing: build it with `python -m provekit_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.pkgpu_last_error.argtypes = []
        L.pkgpu_last_error.restype = ctypes.c_char_p
        L.pkgpu_abi_version.argtypes = []
        L.pkgpu_abi_version.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != OK:
        msg = lib().pkgpu_last_error().decode(errors="replace")
        raise ProvekitGpuError(rc, msg)


def _call(name: str, *args) -> None:
    _check(getattr(lib(), name)(*args))


_initialised = False


def init(device: int = 0) -> None:
    """pkgpu_init: select and set up the device for this thread."""
    global _initialised
    _call("pkgpu_init", device)
    _initialised = True


def _ensure_init() -> None:
    if not _initialised:
        init(int(os.environ.get("PKGPU_DEVICE", "0")))


def _fr_array(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    if a.ndim == 1:
        a = a.reshape(-1, 4)
    if a.ndim != 2 or a.shape[1] != 4:
        raise ValueError("field arrays have shape (n, 4)")
    return a


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ----------------------------------------------------------------- ntt crate
def ntt_nr(values: np.ndarray, codeword_size: int, num_groups: int) -> None:
    """ntt::ntt_nr (ntt/src/ntt.rs:105): in place, natural -> bit-reversed."""
    _ensure_init()
    if values.dtype != np.uint64 or not values.flags.c_contiguous:
        raise ValueError("values must be a C-contiguous uint64 (n, 4) array")
    _call("pkgpu_ntt_nr", _ptr(values), values.size // 4, codeword_size, num_groups)


def ntt_autotune(log_min: int, log_max: int, max_skip: int = 2, save_path: Optional[str] = None) -> float:
    """Time every NTT pass shape for the given lengths on the current device
    and keep the fastest per (length, stages); returns the seconds spent."""
    _ensure_init()
    t = ctypes.c_double()
    _call("pkgpu_ntt_autotune", log_min, log_max, max_skip,
          save_path.encode() if save_path else None, ctypes.byref(t))
    return t.value


def ntt_tuning_load(path: str) -> None:
    _call("pkgpu_ntt_tuning_load", path.encode())


def ntt_tuning_get(log_codeword: int, stages: int) -> dict:
    a, b, c = _i(), _i(), _i()
    _call("pkgpu_ntt_tuning_get", log_codeword, stages, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return {"tile_log": a.value, "bpt": b.value, "min_run": c.value}


def intt_rn(values: np.ndarray) -> None:
    """ntt::intt_rn (ntt/src/ntt.rs:326): in place, bit-reversed -> natural."""
    _ensure_init()
    if values.dtype != np.uint64 or not values.flags.c_contiguous:
        raise ValueError("values must be a C-contiguous uint64 (n, 4) array")
    _call("pkgpu_intt_rn", _ptr(values), values.size // 4)


def intt_nr(values: np.ndarray) -> None:
    """ntt's intt_nr (ntt/src/ntt.rs:333-348): in place, natural-order
    evaluations -> coefficients in bit-reversed order."""
    _ensure_init()
    if values.dtype != np.uint64 or not values.flags.c_contiguous:
        raise ValueError("values must be a C-contiguous uint64 (n, 4) array")
    _call("pkgpu_intt_nr", _ptr(values), values.size // 4)


def reverse_order(values: np.ndarray, codeword_size: Optional[int] = None) -> None:
    """ntt's reverse_order (ntt/src/ntt.rs:310-323): the first codeword_size
    elements (all by default) in bit-reversed order, in place."""
    _ensure_init()
    if values.dtype != np.uint64 or not values.flags.c_contiguous:
        raise ValueError("values must be a C-contiguous uint64 (n, 4) array")
    n = values.size // 4
    _call("pkgpu_reverse_order", _ptr(values), n, n if codeword_size is None else codeword_size)


# ----------------------------------------------------------------- RSFr
class RSFr:
    """ReedSolomon<Fr> plugin RSFr (provekit/common/src/ntt.rs:12-105)."""

    def next_order(self, size: int) -> Optional[int]:
        out = _sz()
        rc = lib().pkgpu_rs_next_order(size, ctypes.byref(out))
        if rc == ERR_SIZE:
            return None
        _check(rc)
        return out.value

    def generator(self, codeword_length: int) -> np.ndarray:
        out = np.zeros(4, dtype=np.uint64)
        _call("pkgpu_rs_generator", codeword_length, _ptr(out))
        return out

    def evaluation_points(self, masked_message_length: int, codeword_length: int,
                          indices: Sequence[int]) -> np.ndarray:
        _ensure_init()
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        out = np.zeros((len(idx), 4), dtype=np.uint64)
        _call("pkgpu_rs_evaluation_points", masked_message_length, codeword_length,
              _ptr(idx), len(idx), _ptr(out))
        return out

    def interleaved_encode(self, messages: Sequence[np.ndarray], masks: np.ndarray,
                           codeword_length: int) -> np.ndarray:
        _ensure_init()
        if len(messages) == 0:
            return np.zeros((0, 4), dtype=np.uint64)
        msgs = [_fr_array(m) for m in messages]
        mlen = msgs[0].shape[0]
        for m in msgs:
            if m.shape[0] != mlen:
                raise ProvekitGpuError(ERR_ARG, "assertion `left == right` failed: message lengths differ")
        masks = _fr_array(masks) if masks is not None and len(masks) else np.zeros((0, 4), np.uint64)
        ptrs = (ctypes.c_void_p * len(msgs))(*[_ptr(m) for m in msgs])
        out = np.zeros((len(msgs) * codeword_length, 4), dtype=np.uint64)
        _call("pkgpu_rs_interleaved_encode", ptrs, len(msgs), mlen, _ptr(masks),
              masks.shape[0], codeword_length, _ptr(out))
        return out


# ----------------------------------------------------------------- Skyscraper
class skyscraper:  # noqa: N801 - namespace mirroring the `skyscraper` crate
    """skyscraper::simple (skyscraper/core/src/simple.rs)."""

    @staticmethod
    def compress(l, r) -> np.ndarray:
        _ensure_init()
        a = np.ascontiguousarray(l, dtype=np.uint64).reshape(4)
        b = np.ascontiguousarray(r, dtype=np.uint64).reshape(4)
        out = np.zeros(4, dtype=np.uint64)
        _call("pkgpu_skyscraper_compress", _ptr(a), _ptr(b), _ptr(out))
        return out

    @staticmethod
    def compress_many(messages: bytes) -> bytes:
        _ensure_init()
        m = np.frombuffer(messages, dtype=np.uint8)
        out = np.zeros(len(messages) // 2, dtype=np.uint8)
        _call("pkgpu_skyscraper_compress_many", _ptr(m) if len(m) else None, len(m),
              _ptr(out) if len(out) else None, len(out))
        return out.tobytes()


class pow:  # noqa: N801 - namespace mirroring `skyscraper::pow`
    """skyscraper::pow (skyscraper/core/src/pow.rs:18-45).  A challenge is any
    256-bit integer as 4 LE u64."""

    @staticmethod
    def threshold(difficulty: float) -> np.ndarray:
        out = np.zeros(4, dtype=np.uint64)
        _call("pkgpu_skyscraper_pow_threshold", float(difficulty), _ptr(out))
        return out

    @staticmethod
    def verify(challenge, difficulty: float, nonce: int) -> bool:
        _ensure_init()
        c = np.ascontiguousarray(challenge, dtype=np.uint64).reshape(4)
        ok = _i(0)
        _call("pkgpu_skyscraper_pow_verify", _ptr(c), float(difficulty), int(nonce),
              ctypes.byref(ok))
        return bool(ok.value)

    @staticmethod
    def solve(challenge, difficulty: float, max_nonces: int = 0, stream=None) -> Optional[int]:
        """The smallest nonce passing threshold(difficulty + 0.01); None if
        none below max_nonces (0: unbounded, as the reference)."""
        _ensure_init()
        c = np.ascontiguousarray(challenge, dtype=np.uint64).reshape(4)
        n, f = ctypes.c_uint64(0), _i(0)
        _call("pkgpu_skyscraper_pow_solve", _ptr(c), float(difficulty), int(max_nonces),
              ctypes.byref(n), ctypes.byref(f), stream)
        return int(n.value) if f.value else None


class SkyscraperPoW:
    """spongefish PowStrategy (provekit/common/src/skyscraper/pow.rs:8-36)."""

    def __init__(self, challenge: bytes, bits: float):
        if not (0.0 <= bits < 60.0):
            raise ValueError("bits must be smaller than 60")
        if len(challenge) != 32:
            raise ValueError("challenge must be 32 bytes")
        self.challenge = bytes(challenge)
        self.bits = float(bits)

    def _c(self) -> np.ndarray:
        return np.frombuffer(self.challenge, dtype=np.uint64)

    def check(self, nonce: int) -> bool:
        return pow.verify(self._c(), self.bits, nonce)

    def solution(self, nonce: int) -> dict:
        return {"challenge": self.challenge, "nonce": int(nonce)}

    def solve(self) -> dict:
        return self.solution(pow.solve(self._c(), self.bits))


class SkyscraperHashEngine:
    """whir HashEngine plugin (provekit/common/src/skyscraper/whir.rs:29-101)."""

    def name(self) -> str:
        return "skyscraper"

    def supports_size(self, size: int) -> bool:
        return size > 0 and size % 32 == 0

    def preferred_batch_size(self) -> int:
        return 1 << 16

    def hash_many(self, size: int, input: bytes, count: int) -> bytes:
        _ensure_init()
        m = np.frombuffer(input, dtype=np.uint8)
        out = np.zeros(32 * count, dtype=np.uint8)
        _call("pkgpu_skyscraper_hash_many", size, _ptr(m) if len(m) else None, len(m),
              _ptr(out) if len(out) else None, count)
        return out.tobytes()


# ----------------------------------------------------------------- Poseidon2
def poseidon2_permutation(state) -> np.ndarray:
    """poseidon2_permutation (poseidon2/src/permutation.rs:174): (4, 4) -> (4, 4);
    also accepts (n, 4, 4) for n states."""
    _ensure_init()
    s = np.ascontiguousarray(state, dtype=np.uint64)
    single = s.ndim == 2
    s = s.reshape(-1, 4, 4)
    out = np.zeros_like(s)
    _call("pkgpu_poseidon2_permutation_many", _ptr(s), s.shape[0], _ptr(out))
    return out[0] if single else out


def poseidon2_hash(inputs) -> np.ndarray:
    """poseidon2_hash (poseidon2/src/hash.rs:25)."""
    _ensure_init()
    x = _fr_array(inputs) if len(inputs) else np.zeros((0, 4), np.uint64)
    out = np.zeros(4, dtype=np.uint64)
    _call("pkgpu_poseidon2_hash", _ptr(x) if x.size else None, x.shape[0], _ptr(out))
    return out


def poseidon2_hash_many(messages: np.ndarray) -> np.ndarray:
    """(count, len, 4) messages -> (count, 4) digests."""
    _ensure_init()
    m = np.ascontiguousarray(messages, dtype=np.uint64)
    count, ln = m.shape[0], m.shape[1]
    out = np.zeros((count, 4), dtype=np.uint64)
    _call("pkgpu_poseidon2_hash_many", _ptr(m) if m.size else None, ln, count, _ptr(out))
    return out


def poseidon2_merkle(leaves: np.ndarray, want_nodes: bool = True):
    """Poseidon2 2-to-1 Merkle tree; returns (root (4,), nodes (2n-1, 4) or None)."""
    _ensure_init()
    x = _fr_array(leaves)
    n = x.shape[0]
    nodes = np.zeros((2 * n - 1, 4), dtype=np.uint64) if want_nodes else None
    root = np.zeros(4, dtype=np.uint64)
    _call("pkgpu_poseidon2_merkle", _ptr(x), n, _ptr(nodes) if nodes is not None else None, _ptr(root))
    return root, nodes


class Graph:
    """An executable hipGraph made by one of the *_graph entry points."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream=None) -> None:
        _call("pkgpu_graph_launch", self.handle, stream)

    def destroy(self) -> None:
        if self.handle:
            lib().pkgpu_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def poseidon2_merkle_graph(leaves: "DeviceBuffer", n: int, nodes: "DeviceBuffer") -> Graph:
    h = ctypes.c_void_p()
    _call("pkgpu_poseidon2_merkle_graph", leaves.ptr, n, nodes.ptr, ctypes.byref(h))
    return Graph(h.value)


# ----------------------------------------------------------------- R1CS / sumcheck
class SparseMatrix:
    """CSR as provekit's SparseMatrix (provekit/common/src/sparse_matrix.rs:56-71)
    with interned values expanded: row_starts (u32, one per row), cols (u32),
    vals ((nnz, 4) Montgomery limbs)."""

    def __init__(self, num_rows, num_cols, row_starts, cols, vals):
        self.num_rows, self.num_cols = int(num_rows), int(num_cols)
        self.row_starts = np.ascontiguousarray(row_starts, dtype=np.uint32)
        self.cols = np.ascontiguousarray(cols, dtype=np.uint32)
        self.vals = _fr_array(vals) if len(vals) else np.zeros((0, 4), np.uint64)

    @property
    def nnz(self):
        return self.cols.shape[0]

    def _args(self):
        return (_ptr(self.row_starts) if self.num_rows else None, self.nnz,
                _ptr(self.cols) if self.nnz else None, _ptr(self.vals) if self.nnz else None)

    def __mul__(self, z):
        """HydratedSparseMatrix * &[Fr] (sparse_matrix.rs:606-624)."""
        _ensure_init()
        zz = _fr_array(z) if len(z) else np.zeros((0, 4), np.uint64)
        out = np.zeros((self.num_rows, 4), dtype=np.uint64)
        rs, nnz, cols, vals = self._args()
        _call("pkgpu_sparse_mul", self.num_rows, self.num_cols, rs, nnz, cols, vals,
              _ptr(zz) if zz.size else None, zz.shape[0], _ptr(out) if out.size else None)
        return out

    def transpose(self) -> "SparseMatrix":
        """SparseMatrix::transpose (sparse_matrix.rs:409-449): entries in
        (column, row) order."""
        _ensure_init()
        trs = np.zeros(self.num_cols, np.uint32)
        tcs = np.zeros(self.nnz, np.uint32)
        tvs = np.zeros((self.nnz, 4), np.uint64)
        rs, nnz, cols, vals = self._args()
        _call("pkgpu_sparse_transpose", self.num_rows, self.num_cols, rs, nnz, cols, vals,
              _ptr(trs) if trs.size else None, _ptr(tcs) if tcs.size else None,
              _ptr(tvs) if tvs.size else None)
        return SparseMatrix(self.num_cols, self.num_rows, trs, tcs, tvs)


def transpose_r1cs_matrices(A: SparseMatrix, B: SparseMatrix, C: SparseMatrix):
    """transpose_r1cs_matrices (provekit/common/src/utils/sumcheck.rs:194-201)."""
    return A.transpose(), B.transpose(), C.transpose()


def multiply_transposed_by_eq_alpha(at: SparseMatrix, bt: SparseMatrix, ct: SparseMatrix, alpha,
                                    num_constraints: int) -> list:
    """multiply_transposed_by_eq_alpha (sumcheck.rs:203-227): [At, Bt, Ct] times
    eq(alpha, .) truncated to num_constraints entries (the transposed
    matrices' column count)."""
    _ensure_init()
    al = _fr_array(alpha) if len(alpha) else np.zeros((0, 4), np.uint64)
    dal = DeviceBuffer.from_array(al) if al.size else None
    outs = []
    for t in (at, bt, ct):
        if t.num_cols != num_constraints:
            raise ValueError("transposed matrix column count must equal the constraint count")
        bufs = [DeviceBuffer.from_array(x) for x in (t.row_starts, t.cols, t.vals) if x.size]
        it = iter(bufs)
        rs = next(it).ptr if t.row_starts.size else None
        cs = next(it).ptr if t.cols.size else None
        vs = next(it).ptr if t.vals.size else None
        out = DeviceBuffer(32 * max(t.num_rows, 1))
        _call("pkgpu_multiply_transposed_by_eq_alpha_device", t.num_rows, t.num_cols, rs, t.nnz, cs, vs,
              dal.ptr if dal is not None else None, al.shape[0], out.ptr, None)
        outs.append(out.download(shape=(t.num_rows, 4)) if t.num_rows else np.zeros((0, 4), np.uint64))
    return outs


class PkCsr(ctypes.Structure):
    """pkgpu_csr (include/provekit_gpu.h)."""
    _fields_ = [("num_rows", _sz), ("num_cols", _sz), ("nnz", _sz), ("row_starts", _vp),
                ("cols", _vp), ("vals", _vp)]


def _csr_download(c: "PkCsr", bufs) -> SparseMatrix:
    rs = bufs[0].download(np.uint32)[: c.num_rows] if c.num_rows else np.zeros(0, np.uint32)
    cs = bufs[1].download(np.uint32)[: c.nnz] if c.nnz else np.zeros(0, np.uint32)
    vs = bufs[2].download(shape=(-1, 4))[: c.nnz] if c.nnz else np.zeros((0, 4), np.uint64)
    return SparseMatrix(c.num_rows, c.num_cols, rs, cs, vs)


def _csr_alloc(c: "PkCsr"):
    bufs = [DeviceBuffer(4 * max(c.num_rows, 1)), DeviceBuffer(4 * max(c.nnz, 1)),
            DeviceBuffer(32 * max(c.nnz, 1))]
    c.row_starts, c.cols, c.vals = bufs[0].ptr, bufs[1].ptr, bufs[2].ptr
    return bufs


def decode_interner(blob: bytes) -> np.ndarray:
    """Interner (interner.rs:6-10, postcard + serde_ark) -> (n, 4) Montgomery Fr."""
    _ensure_init()
    b = np.frombuffer(blob, np.uint8)
    n = _sz()
    _call("pkgpu_interner_decode_device", _ptr(b) if b.size else None, b.size, None, ctypes.byref(n), None)
    if not n.value:
        return np.zeros((0, 4), np.uint64)
    out = DeviceBuffer(32 * n.value)
    _call("pkgpu_interner_decode_device", _ptr(b), b.size, out.ptr, ctypes.byref(n), None)
    return out.download(shape=(n.value, 4))


def decode_sparse_matrix(blob: bytes, interner) -> SparseMatrix:
    """SparseMatrix (sparse_matrix.rs:149-290, postcard) decoded on the device,
    values hydrated through `interner` ((n, 4) Montgomery Fr)."""
    _ensure_init()
    b = np.frombuffer(blob, np.uint8)
    it = np.ascontiguousarray(interner, dtype=np.uint64).reshape(-1, 4)
    dit = DeviceBuffer.from_array(it) if it.size else None
    c = PkCsr()
    _call("pkgpu_sparse_matrix_decode_device", _ptr(b) if b.size else None, b.size,
          dit.ptr if dit else None, it.shape[0], ctypes.byref(c), None)
    bufs = _csr_alloc(c)
    _call("pkgpu_sparse_matrix_decode_device", _ptr(b) if b.size else None, b.size,
          dit.ptr if dit else None, it.shape[0], ctypes.byref(c), None)
    return _csr_download(c, bufs)


def decode_r1cs(blob: bytes) -> dict:
    """R1CS (r1cs.rs:59-66, postcard) decoded on the device: num_public_inputs,
    the interner ((n, 4) Montgomery Fr) and the hydrated a, b, c."""
    _ensure_init()
    b = np.frombuffer(blob, np.uint8)
    npub, ni = _sz(), _sz()
    mats = (PkCsr * 3)()
    args = (_ptr(b) if b.size else None, b.size, ctypes.byref(npub))
    _call("pkgpu_r1cs_decode_device", *args, None, ctypes.byref(ni), mats, None)
    inter = DeviceBuffer(32 * max(ni.value, 1))
    bufs = [_csr_alloc(mats[i]) for i in range(3)]
    _call("pkgpu_r1cs_decode_device", *args, inter.ptr, ctypes.byref(ni), mats, None)
    return {"num_public_inputs": npub.value,
            "interner": inter.download(shape=(-1, 4))[: ni.value],
            "a": _csr_download(mats[0], bufs[0]), "b": _csr_download(mats[1], bufs[1]),
            "c": _csr_download(mats[2], bufs[2])}


def calculate_witness_bounds(A: SparseMatrix, B: SparseMatrix, witness):
    """calculate_witness_bounds (provekit/common/src/utils/sumcheck.rs:163-179)."""
    _ensure_init()
    w = _fr_array(witness)
    n = 1
    while n < A.num_rows:
        n <<= 1
    outs = [np.zeros((n, 4), np.uint64) for _ in range(3)]
    plen = _sz()
    ra, na, ca, va = A._args()
    rb, nb, cb, vb = B._args()
    _call("pkgpu_calculate_witness_bounds", A.num_rows, A.num_cols, ra, na, ca, va, rb, nb, cb, vb,
          _ptr(w), w.shape[0], _ptr(outs[0]), _ptr(outs[1]), _ptr(outs[2]), ctypes.byref(plen))
    return tuple(o[: plen.value] for o in outs)


def sparse_mul_interned(num_rows: int, num_cols: int, row_starts, cols, ids, interner, z) -> np.ndarray:
    """HydratedSparseMatrix * &[Fr] on interned storage (pkgpu_sparse_mul_interned):
    entry i's value is interner[ids[i]]."""
    _ensure_init()
    rs = np.ascontiguousarray(row_starts, dtype=np.uint32)
    cs = np.ascontiguousarray(cols, dtype=np.uint32)
    ii = np.ascontiguousarray(ids, dtype=np.uint32)
    tab = _fr_array(interner) if len(interner) else np.zeros((0, 4), np.uint64)
    zz = _fr_array(z) if len(z) else np.zeros((0, 4), np.uint64)
    out = np.zeros((num_rows, 4), dtype=np.uint64)
    _call("pkgpu_sparse_mul_interned", num_rows, num_cols, _ptr(rs) if rs.size else None, cs.shape[0],
          _ptr(cs) if cs.size else None, _ptr(ii) if ii.size else None,
          _ptr(tab) if tab.size else None, tab.shape[0], _ptr(zz) if zz.size else None, zz.shape[0],
          _ptr(out) if out.size else None)
    return out


def sparse_transpose_interned(num_rows: int, num_cols: int, row_starts, cols, ids):
    """SparseMatrix::transpose on interned storage (sparse_matrix.rs:409-449):
    returns (t_row_starts, t_cols, t_ids), the ids moved with their entries."""
    _ensure_init()
    rs = np.ascontiguousarray(row_starts, dtype=np.uint32)
    cs = np.ascontiguousarray(cols, dtype=np.uint32)
    ii = np.ascontiguousarray(ids, dtype=np.uint32)
    trs = np.zeros(num_cols, np.uint32)
    tcs = np.zeros(cs.shape[0], np.uint32)
    tis = np.zeros(cs.shape[0], np.uint32)
    _call("pkgpu_sparse_transpose_interned", num_rows, num_cols, _ptr(rs) if rs.size else None,
          cs.shape[0], _ptr(cs) if cs.size else None, _ptr(ii) if ii.size else None,
          _ptr(trs) if trs.size else None, _ptr(tcs) if tcs.size else None,
          _ptr(tis) if tis.size else None)
    return trs, tcs, tis


def eq_evaluations(r, num_entries: int) -> np.ndarray:
    """calculate_evaluations_over_boolean_hypercube_for_eq (sumcheck.rs:108-120)."""
    _ensure_init()
    rr = _fr_array(r) if len(r) else np.zeros((0, 4), np.uint64)
    out = np.zeros((num_entries, 4), dtype=np.uint64)
    _call("pkgpu_eq_evaluations", _ptr(rr) if rr.size else None, rr.shape[0], num_entries,
          _ptr(out) if out.size else None)
    return out


def sumcheck_r1cs_round(a: "DeviceBuffer", b: "DeviceBuffer", c: "DeviceBuffer",
                        eq: "DeviceBuffer", size: int, fold=None, stream=None) -> np.ndarray:
    """One sumcheck_fold_map_reduce round on device-resident MLEs -> (3, 4)."""
    out = np.zeros((3, 4), dtype=np.uint64)
    f = np.ascontiguousarray(fold, dtype=np.uint64).reshape(4) if fold is not None else None
    _call("pkgpu_sumcheck_r1cs_round_device", a.ptr, b.ptr, c.ptr, eq.ptr, size,
          _ptr(f) if f is not None else None, _ptr(out), stream)
    return out


def sumcheck_prove(a, b, c, m: int, blinding=None, seed: int = 0, stream=None) -> dict:
    """Whole sumcheck on the device (pkgpu_sumcheck_prove_device).  a, b, c are
    host (2^m, 4) arrays (copied to the device) or DeviceBuffers (consumed).
    Returns coeffs, alphas, r, final, rho and blinding_eval."""
    _ensure_init()
    bufs = []
    for x in (a, b, c):
        bufs.append(x if isinstance(x, DeviceBuffer) else DeviceBuffer.from_array(_fr_array(x)))
    coeffs = np.zeros((max(m, 1), 4, 4), np.uint64)
    alphas = np.zeros((max(m, 1), 4), np.uint64)
    r = np.zeros((max(m, 1), 4), np.uint64)
    fin = np.zeros(4, np.uint64)
    rho = np.zeros(4, np.uint64)
    beval = np.zeros(4, np.uint64)
    g = np.ascontiguousarray(blinding, dtype=np.uint64).reshape(-1, 4) if blinding is not None else None
    _call("pkgpu_sumcheck_prove_device", bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, m,
          _ptr(g) if g is not None else None, seed, _ptr(coeffs), _ptr(alphas), _ptr(r), _ptr(fin),
          _ptr(rho), _ptr(beval), stream)
    return {"coeffs": coeffs[:m], "alphas": alphas[:m], "r": r[:m], "final": fin, "rho": rho,
            "blinding_eval": beval}


# ----------------------------------------------------------------- commit
def commit(coeffs: np.ndarray, blowup: int, coset: bool = True, leaf_elems: int = 16,
           want_codeword: bool = True, want_nodes: bool = True):
    """Commit one polynomial (pkgpu_commit).  Returns (root bytes, codeword or
    None, nodes (2*leaves-1, 32) uint8 or None)."""
    _ensure_init()
    c = _fr_array(coeffs)
    n = c.shape[0]
    N = n * blowup
    leaves = N // leaf_elems
    cw = np.zeros((N, 4), dtype=np.uint64) if want_codeword else None
    nodes = np.zeros((2 * leaves - 1, 32), dtype=np.uint8) if want_nodes else None
    root = np.zeros(32, dtype=np.uint8)
    _call("pkgpu_commit", _ptr(c), n, blowup, int(bool(coset)), leaf_elems,
          _ptr(cw) if cw is not None else None, _ptr(nodes) if nodes is not None else None,
          _ptr(root))
    return root.tobytes(), cw, nodes


def whir_r1cs_scheme_size(num_witnesses: int, num_constraints: int, w1_size: int):
    """WhirR1CSScheme::new_from_dimensions sizes -> (m, m0)."""
    m, m0 = ctypes.c_size_t(), ctypes.c_size_t()
    _call("pkgpu_whir_r1cs_scheme_size", num_witnesses, num_constraints, w1_size,
          ctypes.byref(m), ctypes.byref(m0))
    return m.value, m0.value


def whir_r1cs_commit(witness, num_vars: int, blinding=None, blowup: int = 4, coset: bool = True,
                     leaf_elems: int = 16):
    """WhirR1CSScheme::commit (pkgpu_whir_r1cs_commit).  Returns (root bytes,
    shaped polynomial (2^num_vars, 4), codeword, nodes (2*leaves-1, 32))."""
    _ensure_init()
    w = _fr_array(witness) if len(witness) else np.zeros((0, 4), np.uint64)
    g = np.ascontiguousarray(blinding, dtype=np.uint64).reshape(-1, 4) if blinding is not None else None
    m0 = g.shape[0] // 4 if g is not None else 0
    n = 1 << num_vars
    N = n * blowup
    poly = np.zeros((n, 4), np.uint64)
    cw = np.zeros((N, 4), np.uint64)
    nodes = np.zeros((2 * (N // leaf_elems) - 1, 32), np.uint8)
    root = np.zeros(32, np.uint8)
    _call("pkgpu_whir_r1cs_commit", _ptr(w) if w.size else None, w.shape[0], num_vars,
          _ptr(g) if g is not None else None, m0, blowup, int(bool(coset)), leaf_elems,
          _ptr(poly), _ptr(cw), _ptr(nodes), _ptr(root))
    return root.tobytes(), poly, cw, nodes


def random_fr_device(n: int, seed: int, stream_name: str = "", first: int = 0,
                     out: "DeviceBuffer" = None, stream=None) -> "DeviceBuffer":
    """n seeded uniform Fr elements generated on the device
    (pkgpu_random_fr_device; key and definition in inputs.wide_key /
    inputs.random_fr_wide)."""
    from .inputs import wide_key
    _ensure_init()
    key = np.frombuffer(wide_key(seed, stream_name), dtype=np.uint8).copy()
    buf = out if out is not None else DeviceBuffer(32 * max(n, 1))
    _call("pkgpu_random_fr_device", _ptr(key), first, n, buf.ptr, stream)
    return buf


def scalar_mul_many(a, b) -> np.ndarray:
    """bn254_multiplier::scalar_mul element-wise (pkgpu_scalar_mul_many)."""
    _ensure_init()
    A = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
    B = np.ascontiguousarray(b, dtype=np.uint64).reshape(-1, 4)
    if A.shape != B.shape:
        raise ValueError("operands differ in length")
    out = np.zeros_like(A)
    _call("pkgpu_scalar_mul_many", _ptr(A), _ptr(B), _ptr(out), A.shape[0])
    return out


def scalar_sqr_many(a) -> np.ndarray:
    """bn254_multiplier::scalar_sqr element-wise (pkgpu_scalar_sqr_many)."""
    _ensure_init()
    A = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros_like(A)
    _call("pkgpu_scalar_sqr_many", _ptr(A), _ptr(out), A.shape[0])
    return out


# ----------------------------------------------------------------- device memory
class DeviceBuffer:
    """A device allocation (pkgpu_malloc) for the *_device entry points."""

    def __init__(self, nbytes: int):
        _ensure_init()
        p = ctypes.c_void_p()
        _call("pkgpu_malloc", ctypes.byref(p), nbytes)
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than the buffer")
        _call("pkgpu_memcpy_h2d", self.ptr, _ptr(a), a.nbytes, stream)

    def download(self, dtype=np.uint64, shape=None, nbytes=None, stream=None) -> np.ndarray:
        nb = self.nbytes if nbytes is None else nbytes
        out = np.empty(nb // np.dtype(dtype).itemsize, dtype=dtype)
        _call("pkgpu_memcpy_d2h", _ptr(out), self.ptr, nb, stream)
        return out.reshape(shape) if shape is not None else out

    def free(self) -> None:
        if self.ptr:
            lib().pkgpu_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    _call("pkgpu_device_synchronize")


def commit_device(coeffs: DeviceBuffer, n: int, blowup: int, coset: bool, leaf_elems: int,
                  codeword: DeviceBuffer, nodes: DeviceBuffer, stream=None) -> None:
    _call("pkgpu_commit_device", coeffs.ptr, n, blowup, int(bool(coset)), leaf_elems,
          codeword.ptr, nodes.ptr, stream)


def ntt_nr_device(values: DeviceBuffer, length: int, codeword_size: int, num_groups: int,
                  stream=None) -> None:
    _call("pkgpu_ntt_nr_device", values.ptr, length, codeword_size, num_groups, stream)


def intt_rn_device(values: DeviceBuffer, length: int, stream=None) -> None:
    _call("pkgpu_intt_rn_device", values.ptr, length, stream)


def intt_nr_device(values: DeviceBuffer, length: int, stream=None) -> None:
    _call("pkgpu_intt_nr_device", values.ptr, length, stream)


def reverse_order_device(values: DeviceBuffer, length: int, codeword_size: int, stream=None) -> None:
    _call("pkgpu_reverse_order_device", values.ptr, length, codeword_size, stream)


def hash_many_device(size: int, inp: DeviceBuffer, count: int, out: DeviceBuffer,
                     stream=None) -> None:
    _call("pkgpu_skyscraper_hash_many_device", size, inp.ptr, count, out.ptr, stream)


def merkle_device(leaves: DeviceBuffer, n: int, nodes: DeviceBuffer, stream=None) -> None:
    _call("pkgpu_skyscraper_merkle_device", leaves.ptr, n, nodes.ptr, stream)


def merkle_open_device(codeword: DeviceBuffer, nodes: DeviceBuffer, leaves: int, leaf_elems: int,
                       leaf_indices, stream=None):
    """Openings from a resident tree: (values (n, leaf_elems, 4), paths (n, depth, 32 bytes))."""
    idx = np.ascontiguousarray(leaf_indices, dtype=np.uint64)
    n = idx.shape[0]
    depth = max(leaves.bit_length() - 1, 0)
    values = np.zeros((n, leaf_elems, 4), np.uint64)
    paths = np.zeros((n, depth, 32), np.uint8)
    _call("pkgpu_merkle_open_device", codeword.ptr, nodes.ptr, leaves, leaf_elems,
          _ptr(idx) if n else None, n, _ptr(values) if n else None,
          _ptr(paths) if paths.size else None, stream)
    return values, paths


class BatchCommit:
    """A batched commitment plan (pkgpu_batch_plan_create) over device buffers:
    polys is a list of (coeffs DeviceBuffer, n, blowup, coset)."""

    def __init__(self, polys, leaf_elems: int = 16, tile_log: int = 0):
        _ensure_init()
        self.count = len(polys)
        self.leaf_elems = leaf_elems
        self.polys = polys
        self.codewords, self.nodes = [], []
        for (c, n, blowup, coset) in polys:
            N = n * blowup
            self.codewords.append(DeviceBuffer(32 * N))
            self.nodes.append(DeviceBuffer(32 * (2 * (N // leaf_elems) - 1)))
        arr = lambda vals, t: (t * len(vals))(*vals)  # noqa: E731
        h = ctypes.c_void_p()
        _call("pkgpu_batch_plan_create", self.count, arr([p[0].ptr for p in polys], _vp),
              arr([p[1] for p in polys], _sz), arr([p[2] for p in polys], _sz),
              arr([int(bool(p[3])) for p in polys], _i), leaf_elems,
              arr([b.ptr for b in self.codewords], _vp), arr([b.ptr for b in self.nodes], _vp),
              tile_log, ctypes.byref(h))
        self.handle = h.value

    def run(self, stream=None) -> None:
        _call("pkgpu_batch_run", self.handle, stream)

    def graph_launch(self, stream=None) -> None:
        _call("pkgpu_batch_graph_launch", self.handle, stream)

    def roots(self, stream=None) -> list:
        out = np.zeros(32 * self.count, dtype=np.uint8)
        _call("pkgpu_batch_roots", self.handle, _ptr(out), stream)
        return [out[32 * i:32 * i + 32].tobytes() for i in range(self.count)]

    def destroy(self) -> None:
        if self.handle:
            lib().pkgpu_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PkJob(ctypes.Structure):
    """pkgpu_job (include/provekit_gpu.h)."""
    _fields_ = [("m", _sz), ("row_starts", _vp * 3), ("cols", _vp * 3), ("vals", _vp * 3),
                ("nnz", _sz * 3), ("z", _vp), ("seed", ctypes.c_uint64), ("blowup", _sz),
                ("coset", _i), ("leaf_elems", _sz)]


class PkJobResult(ctypes.Structure):
    _fields_ = [("root", ctypes.c_uint64 * 4), ("final_claim", ctypes.c_uint64 * 4),
                ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double),
                ("run_ms", ctypes.c_double), ("status", _i)]


class Server:
    """The serving loop (pkgpu_server_*).  submit() takes (A, B, C) as
    SparseMatrix-like triples of arrays (or device buffers) and z."""

    def __init__(self, hbm_budget_bytes: int, slots: int = 8, device: int = 0):
        _ensure_init()
        h = ctypes.c_void_p()
        _call("pkgpu_server_create", device, hbm_budget_bytes, slots, ctypes.byref(h))
        self.handle = h.value
        self._keep = {}

    def submit(self, m: int, mats, z, seed: int, blowup: int = 4, coset: bool = True,
               leaf_elems: int = 16) -> int:
        """mats: three (row_starts, cols, vals) triples, each array either a
        numpy array (host) or a DeviceBuffer; z likewise."""
        def ptr(x):
            return x.ptr if isinstance(x, DeviceBuffer) else _ptr(x)
        j = PkJob()
        j.m = m
        keep = []
        for k, (rs, cs, vs) in enumerate(mats):
            if not isinstance(rs, DeviceBuffer):
                rs = np.ascontiguousarray(rs, dtype=np.uint32)
                cs = np.ascontiguousarray(cs, dtype=np.uint32)
                vs = _fr_array(vs)
            keep += [rs, cs, vs]
            j.row_starts[k] = ptr(rs)
            j.cols[k] = ptr(cs)
            j.vals[k] = ptr(vs)
            j.nnz[k] = (cs.nbytes // 4) if isinstance(cs, DeviceBuffer) else cs.shape[0]
        if not isinstance(z, DeviceBuffer):
            z = _fr_array(z)
        keep.append(z)
        j.z = ptr(z)
        j.seed, j.blowup, j.coset, j.leaf_elems = seed, blowup, int(bool(coset)), leaf_elems
        jid = ctypes.c_uint64()
        _call("pkgpu_server_submit", self.handle, ctypes.byref(j), ctypes.byref(jid))
        self._keep[jid.value] = (keep, m)
        return jid.value

    def wait(self, jid: int, timeout_ms: float = 0.0) -> dict:
        m = self._keep[jid][1]
        res = PkJobResult()
        coeffs = np.zeros((max(m, 1), 4, 4), np.uint64)
        alphas = np.zeros((max(m, 1), 4), np.uint64)
        r = np.zeros((max(m, 1), 4), np.uint64)
        _call("pkgpu_server_wait", self.handle, jid, timeout_ms, ctypes.byref(res), _ptr(coeffs),
              _ptr(alphas), _ptr(r))
        self._keep[jid] = (None, m)
        return {"root": bytes(np.array(res.root, dtype=np.uint64).tobytes()),
                "final": np.array(res.final_claim, dtype=np.uint64),
                "coeffs": coeffs[:m], "alphas": alphas[:m], "r": r[:m],
                "latency_ms": res.latency_ms, "queue_ms": res.queue_ms, "run_ms": res.run_ms}

    def poll(self, jid: int) -> bool:
        d = ctypes.c_int()
        _call("pkgpu_server_poll", self.handle, jid, ctypes.byref(d))
        return bool(d.value)

    def open(self, jid: int, leaf_indices, leaf_elems: int, depth: int):
        idx = np.ascontiguousarray(leaf_indices, dtype=np.uint64)
        vals = np.zeros((len(idx), leaf_elems, 4), np.uint64)
        paths = np.zeros((len(idx), max(depth, 1), 32), np.uint8)
        _call("pkgpu_server_open", self.handle, jid, _ptr(idx), len(idx), _ptr(vals), _ptr(paths))
        return vals, paths[:, :depth]

    def release(self, jid: int) -> None:
        _call("pkgpu_server_release", self.handle, jid)
        self._keep.pop(jid, None)

    def stats(self) -> dict:
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        _call("pkgpu_server_stats", self.handle, ctypes.byref(a), ctypes.byref(b))
        return {"bytes_in_use": a.value, "peak_in_flight": b.value}

    def destroy(self) -> None:
        if self.handle:
            lib().pkgpu_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
