"""WhirR1CSScheme sizes and commit (provekit/r1cs-compiler/src/whir_r1cs.rs:55-84,
provekit/prover/src/whir_r1cs.rs:82-140), and the bn254_multiplier entry
points: the oracle against the reference's rules, the GPU bit-exact against
the oracle."""
import random

import numpy as np
import pytest

from provekit_amd.inputs import random_fr
from tests.util import P, to_int, unmont


def ceil_log2(n):
    return max(n - 1, 0).bit_length()


# ---------------------------------------------------------------- CPU: oracle
def test_scheme_sizes(oracle):
    """new_from_dimensions: at least 13 witness variables, at least one
    sumcheck variable, and w1's zero padding always holds the 4 m0 blinding
    coefficients (the reason for the extra variable, whir_r1cs.rs:75-78)."""
    # hand-worked cases: w1 fills 2^13 exactly, so one more variable
    assert oracle.whir_r1cs_params(8192 + 10, 100, 8192) == (14, 7)
    assert oracle.whir_r1cs_params(100, 1, 50) == (13, 1)
    assert oracle.whir_r1cs_params(1 << 20, 1 << 20, 1 << 19) == (20, 20)  # w1 fills 2^19
    assert oracle.whir_r1cs_params(1 << 19, 1 << 20, (1 << 19) - 100) == (19, 20)  # 100 >= 80 free
    assert oracle.whir_r1cs_params((1 << 19) + 3, 5, (1 << 19) - 3) == (20, 3)
    rng = random.Random(5)
    for _ in range(2000):
        nw = rng.randrange(1, 1 << rng.randrange(1, 24))
        w1 = rng.randrange(0, nw + 1)
        nc = rng.randrange(0, 1 << rng.randrange(1, 24))
        m, m0 = oracle.whir_r1cs_params(nw, nc, w1)
        assert m >= 13 and m0 >= 1
        assert m0 == max(ceil_log2(nc), 1)
        assert (1 << m) - w1 >= 4 * m0
        assert m in (max(ceil_log2(w1), ceil_log2(nw - w1), 13),
                     max(ceil_log2(w1), ceil_log2(nw - w1), 13) + 1)


def test_witness_shape(oracle):
    """The commit's polynomial: the witness, its blinding coefficients at
    offset = witness size (row-major, i * 4 + j), zeros up to 2^num_vars."""
    w = random_fr(37, seed=1, stream="shape-w")
    g = random_fr(4 * 3, seed=1, stream="shape-g")
    out = oracle.whir_r1cs_shape(w, 6, g)
    assert out.shape == (64, 4)
    assert np.array_equal(out[:37], w)
    assert np.array_equal(out[37:49], g)
    assert not out[49:].any()
    out2 = oracle.whir_r1cs_shape(w, 6)  # the w2 commitment: no blinding
    assert np.array_equal(out2[:37], w) and not out2[37:].any()
    with pytest.raises(ValueError):
        oracle.whir_r1cs_shape(random_fr(65, seed=2, stream="x"), 6)
    with pytest.raises(IndexError):
        oracle.whir_r1cs_shape(random_fr(60, seed=2, stream="x"), 6, g)


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_scheme_sizes(gpu, oracle):
    rng = random.Random(6)
    for _ in range(200):
        nw = rng.randrange(1, 1 << 22)
        w1 = rng.randrange(0, nw + 1)
        nc = rng.randrange(0, 1 << 22)
        assert gpu.whir_r1cs_scheme_size(nw, nc, w1) == oracle.whir_r1cs_params(nw, nc, w1)


@pytest.mark.gpu
@pytest.mark.parametrize("witness_size,num_vars,m0", [(1000, 13, 10), (8192 - 40, 13, 10),
                                                      (5000, 14, 0), ((1 << 18) + 7, 19, 18)])
def test_gpu_whir_r1cs_commit(gpu, oracle, witness_size, num_vars, m0):
    """Shaped polynomial, codeword, every tree node and the root against the
    oracle's shaping and commit."""
    w = random_fr(witness_size, seed=witness_size, stream="wc-w")
    g = random_fr(4 * m0, seed=m0, stream="wc-g") if m0 else None
    print(f"[case] whir_r1cs_commit ws={witness_size} num_vars={num_vars} m0={m0}", flush=True)
    root, poly, cw, nodes = gpu.whir_r1cs_commit(w, num_vars, g, 4, True, 16)
    epoly = oracle.whir_r1cs_shape(w, num_vars, g)
    eroot, ecw, enodes = oracle.commit(epoly, 4, True, 16)
    assert np.array_equal(poly, epoly)
    assert np.array_equal(cw, ecw)
    assert np.array_equal(nodes, enodes)
    assert root == eroot


@pytest.mark.gpu
def test_gpu_whir_r1cs_commit_errors(gpu):
    g = random_fr(4 * 8, seed=1, stream="wc-g")
    with pytest.raises(gpu.ProvekitGpuError) as e:
        gpu.whir_r1cs_commit(random_fr((1 << 13) + 1, seed=1, stream="x"), 13)
    assert "exceeds scheme capacity" in str(e.value)
    with pytest.raises(gpu.ProvekitGpuError) as e:
        gpu.whir_r1cs_commit(random_fr((1 << 13) - 31, seed=1, stream="x"), 13, g)
    assert "index out of bounds" in str(e.value)


@pytest.mark.gpu
def test_gpu_commit_then_zk_sumcheck(gpu, oracle):
    """The w1 path of prove_noir: commit the witness with its blinding
    univariates, run the ZK sumcheck with the same univariates; the
    blinding_eval it sends is the dot product of expand_powers::<4>(alpha)
    with the committed polynomial at offset = witness size
    (whir_r1cs.rs:613-618)."""
    m = 6
    n = 1 << m
    rng = random.Random(8)
    a = [rng.randrange(P) for _ in range(n)]
    b = [rng.randrange(P) for _ in range(n)]
    c = [x * y % P for x, y in zip(a, b)]
    from tests.util import mont
    A, B, C = (np.array([mont(v) for v in x], dtype=np.uint64) for x in (a, b, c))
    g = random_fr(4 * m, seed=3, stream="zk-g")
    witness = random_fr(3000, seed=3, stream="zk-w")
    nv, m0 = gpu.whir_r1cs_scheme_size(3000 + 500, n, 3000)
    assert m0 == m
    _, poly, _, _ = gpu.whir_r1cs_commit(witness, nv, g)
    out = gpu.sumcheck_prove(A, B, C, m, blinding=g, seed=4)
    alphas = [unmont(x) for x in out["alphas"]]
    coeffs = [unmont(x) for x in poly[3000: 3000 + 4 * m]]
    weights = [pow(al, j, P) for al in alphas for j in range(4)]
    assert unmont(out["blinding_eval"]) == sum(w * v for w, v in zip(weights, coeffs)) % P
    exp = oracle.sumcheck_prove(A, B, C, m, g, seed=4)
    assert np.array_equal(out["blinding_eval"], exp["blinding_eval"])


@pytest.mark.gpu
def test_gpu_scalar_mul_sqr(gpu, oracle):
    """bn254_multiplier::scalar_mul / scalar_sqr on the device: the same bytes
    as the oracle's restatement (not canonical), on random inputs up to
    2^256 - 2p and on edge integers."""
    from tests.util import limbs
    out_max = (1 << 256) - 2 * P
    rng = random.Random(9)
    edge = [0, 1, P - 1, P, 2 * P - 1, out_max - 1, (1 << 255) - 1, 1 << 192]
    xs = edge + [rng.randrange(out_max) for _ in range(4000)]
    ys = edge[::-1] + [rng.randrange(out_max) for _ in range(4000)]
    A = np.array([limbs(x) for x in xs], dtype=np.uint64)
    B = np.array([limbs(y) for y in ys], dtype=np.uint64)
    print("[case] scalar_mul/sqr", flush=True)
    got = gpu.scalar_mul_many(A, B)
    for i in range(len(xs)):
        assert np.array_equal(got[i], oracle.scalar_mul(A[i], B[i])), (xs[i], ys[i])
    sq_in = A.copy()
    sq_in[:, :] = np.array([limbs(x % (2 * P)) for x in xs], dtype=np.uint64)
    got = gpu.scalar_sqr_many(sq_in)
    for i in range(len(xs)):
        assert np.array_equal(got[i], oracle.scalar_sqr(sq_in[i]))
        assert to_int(got[i]) < out_max
