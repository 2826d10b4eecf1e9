"""Skyscraper proof of work (skyscraper/core/src/pow.rs:18-45, generic.rs:42-71,
provekit/common/src/skyscraper/pow.rs:8-36).

CPU: the oracle's f64_to_u256 against the reference's own vectors
(pow.rs:94-109), its threshold against a Python restatement over IEEE
doubles (libm's exp2, which the reference's f64::exp2 calls on Linux), and
its solve against a plain scan from nonce 0 (the reference returns the
smallest passing nonce: every chunk below the smallest hit is scanned in
full).  GPU: threshold, verify and solve through the C ABI against the
oracle, including answers past the first few search windows and bounded
searches that stop short of the answer."""
import ctypes
import math

import numpy as np
import pytest

from tests.util import P

M64 = (1 << 64) - 1
_libm = ctypes.CDLL("libm.so.6")
_libm.exp2.restype = ctypes.c_double
_libm.exp2.argtypes = [ctypes.c_double]


def f64_to_u256_py(f: float) -> int:
    if math.isnan(f) or f == math.inf:
        return (1 << 256) - 1
    if f <= 0:
        return 0
    if f >= 2.0**52:
        return min(int(f), (1 << 256) - 1)  # exact: the float is an integer
    return int(math.floor(f + 0.5))


def threshold_py(d: float) -> int:
    modulus = float(0x30644E72E131A029) * 2.0**192
    return f64_to_u256_py(_libm.exp2(-d) * modulus)


def as_int(a) -> int:
    return sum(int(x) << (64 * i) for i, x in enumerate(np.asarray(a, dtype=np.uint64)))


# pow.rs:94-109 (test_f64_to_u256)
F64_VECTORS = [
    (0.0, [0, 0, 0, 0]),
    (-1.7976931348623157e308, [0, 0, 0, 0]),
    (0.49, [0, 0, 0, 0]),
    (0.50, [1, 0, 0, 0]),
    (1.0, [1, 0, 0, 0]),
    (2.0**128, [0, 0, 1, 0]),
    (math.inf, [M64] * 4),
    (-42.0, [0, 0, 0, 0]),
    (float("nan"), [M64] * 4),
]

CHALLENGES = {
    "max": np.array([M64] * 4, dtype=np.uint64),  # pow.rs:114
    "42s": np.frombuffer(bytes([42] * 32), dtype=np.uint64).copy(),  # skyscraper/pow.rs:40
    "zero": np.zeros(4, dtype=np.uint64),
    "rand": np.array([0x0123456789ABCDEF, 0xFEDCBA9876543210, 0x0F1E2D3C4B5A6978, 0x1122334455667788],
                     dtype=np.uint64),
}


def test_oracle_f64_to_u256_reference_vectors(oracle):
    for f, want in F64_VECTORS:
        assert list(map(int, oracle.f64_to_u256(f))) == want, f
        assert f64_to_u256_py(f) == sum(w << (64 * i) for i, w in enumerate(want))


def test_oracle_threshold_closed_form(oracle):
    for d in [0.0, 0.01, 1.0, 3.0, math.pi, 10.0, 10.01, 20.5, 52.0, 60.0, 79.99]:
        t = as_int(oracle.pow_threshold(d))
        assert t == threshold_py(d), d
        # |{x < t}| / p is 2^-d, up to p's top limb standing for p and two
        # double roundings
        assert abs(((t << 200) // P) / 2**200 * 2.0**d - 1.0) < 2.0**-50, d
    for bad in [-0.5, 80.0, 100.0, float("nan")]:
        with pytest.raises(ValueError):
            oracle.pow_threshold(bad)


def test_oracle_threshold_ratio(oracle):
    # threshold(d) / threshold(d + 1) = 2 for integer d (exact powers of two)
    for d in range(0, 70):
        a, b = as_int(oracle.pow_threshold(d)), as_int(oracle.pow_threshold(d + 1))
        assert a == 2 * b or a == 2 * b + 1, d


@pytest.mark.parametrize("name", list(CHALLENGES))
@pytest.mark.parametrize("d", [1.0, 4.0, 6.5])
def test_oracle_solve_is_smallest_passing_nonce(oracle, name, d):
    ch = CHALLENGES[name]
    thr = as_int(oracle.pow_threshold(d + 0.01))
    n = 0
    while as_int(oracle.sky_compress(ch, [n, 0, 0, 0])) >= thr:
        n += 1
    assert oracle.pow_solve(ch, d) == n
    assert oracle.pow_verify(ch, d + 0.01, n)


def test_oracle_reference_solve_verify(oracle):
    # pow.rs:111-118 and provekit/common/src/skyscraper/pow.rs:38-46
    ch = CHALLENGES["max"]
    assert oracle.pow_solve(ch, 0.0) == 0
    assert oracle.pow_verify(ch, 0.0, 12345)
    n = oracle.pow_solve(ch, math.pi)
    assert oracle.pow_verify(ch, math.pi, n)
    n = oracle.pow_solve(CHALLENGES["42s"], 10.0)
    assert oracle.pow_verify(CHALLENGES["42s"], 10.0, n)
    assert oracle.pow_solve(ch, 16.0, limit=1) is None


def test_skyscraper_pow_bits_bound():
    import provekit_amd as pk

    with pytest.raises(ValueError):
        pk.SkyscraperPoW(bytes(32), 60.0)
    with pytest.raises(ValueError):
        pk.SkyscraperPoW(bytes(32), -1.0)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_gpu_threshold_matches_oracle(gpu, oracle):
    for d in [0.0, 0.01, 1.0, math.pi, 10.0, 10.01, 33.3, 79.99]:
        assert np.array_equal(gpu.pow.threshold(d), oracle.pow_threshold(d)), d
    with pytest.raises(gpu.ProvekitGpuError):
        gpu.pow.threshold(80.0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CHALLENGES))
@pytest.mark.parametrize("d", [0.0, 1.0, math.pi, 10.0, 14.0])
def test_gpu_solve_matches_oracle(gpu, oracle, name, d):
    ch = CHALLENGES[name]
    want = oracle.pow_solve(ch, d)
    got = gpu.pow.solve(ch, d)
    print(f"[pow] {name} d={d:.3f} nonce={got}", flush=True)
    assert got == want
    assert gpu.pow.verify(ch, d, got) == oracle.pow_verify(ch, d, got)


@pytest.mark.gpu
def test_gpu_solve_past_first_windows(gpu, oracle):
    # ~2^19 expected hashes: the answer lies beyond the first 2^16-nonce window
    for seed in range(3):
        ch = np.random.default_rng(seed).integers(0, 2**63, 4, dtype=np.uint64)
        want = oracle.pow_solve(ch, 19.0)
        got = gpu.pow.solve(ch, 19.0)
        print(f"[pow] seed={seed} d=19 nonce={got}", flush=True)
        assert got == want
        assert gpu.pow.verify(ch, 19.01, got)
        if got > 0:
            # the nonce just below fails at the prover's threshold
            assert not oracle.pow_verify(ch, 19.01, got - 1)


@pytest.mark.gpu
def test_gpu_solve_bounded(gpu, oracle):
    ch = CHALLENGES["rand"]
    want = oracle.pow_solve(ch, 12.0)
    assert want is not None and want > 0
    assert gpu.pow.solve(ch, 12.0, max_nonces=want) is None
    assert gpu.pow.solve(ch, 12.0, max_nonces=want + 1) == want


@pytest.mark.gpu
def test_gpu_verify_matches_oracle(gpu, oracle):
    rng = np.random.default_rng(7)
    for _ in range(16):
        ch = rng.integers(0, 2**63, 4, dtype=np.uint64)
        d = float(rng.uniform(0.0, 6.0))
        nonce = int(rng.integers(0, 2**62))
        assert gpu.pow.verify(ch, d, nonce) == oracle.pow_verify(ch, d, nonce)


@pytest.mark.gpu
def test_gpu_skyscraper_pow_strategy(gpu):
    # provekit/common/src/skyscraper/pow.rs:38-46
    pw = gpu.SkyscraperPoW(bytes([42] * 32), 10.0)
    sol = pw.solve()
    assert sol["challenge"] == bytes([42] * 32)
    assert pw.check(sol["nonce"])
