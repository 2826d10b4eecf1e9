"""R1CS products, eq tables and the sumcheck: the oracle pinned by the
reference's tests (provekit/common/src/utils/sumcheck.rs:260-435) and by
big-integer closed forms / the verifier's final equation
(provekit/verifier/src/whir_r1cs.rs:249-257); the GPU bit-exact against it."""
import random

import numpy as np
import pytest

from provekit_amd.inputs import random_fr
from tests.util import P, mont, unmont


def fe(v):
    return mont(v % P)


def ints(arr):
    return [unmont(x) for x in np.asarray(arr).reshape(-1, 4)]


def frs(vals):
    return np.array([fe(v) for v in vals], dtype=np.uint64).reshape(-1, 4)


def eq_py(r, x):
    acc = 1
    for ri, xi in zip(r, x):
        acc = acc * (ri * xi + (1 - ri) * (1 - xi)) % P
    return acc


# ---------------------------------------------------------------- CPU: oracle
def test_calculate_eq_kats(oracle, kats):
    k = kats["calculate_eq"]
    assert unmont(oracle.calculate_eq(frs(k["r"]), frs(k["alpha"]))) == k["out"]
    r = frs([0, 1, 1, 0])
    assert unmont(oracle.calculate_eq(r, frs([0, 1, 1, 0]))) == 1
    assert unmont(oracle.calculate_eq(r, frs([1, 0, 0, 1]))) == 0


def test_eq_table_kats(oracle, kats):
    r = [2, 3, 4, 5]
    table = ints(oracle.eq_evals(frs(r), 16))
    for i, v in enumerate(table):
        x = [(i >> (3 - j)) & 1 for j in range(4)]  # r[0] on the top bit
        assert v == eq_py(r, x)
    assert ints(oracle.eq_evals(frs(r), 10)) == table[:10]
    assert ints(oracle.eq_evals(frs(kats["eval_eq_left_only"]["r"]), 1)) == [kats["eval_eq_left_only"]["out"] % P]
    assert ints(oracle.eq_evals(frs(kats["eq_single_entry"]["r"]), 1)) == [kats["eq_single_entry"]["out"]]
    assert ints(oracle.eq_evals(np.zeros((0, 4), np.uint64), 1)) == [1]


def csr_from_rows(rows, ncols):
    starts, cols, vals = [], [], []
    for row in rows:
        starts.append(len(cols))
        for v, c in sorted(row, key=lambda t: t[1]):
            cols.append(c)
            vals.append(v)
    return np.array(starts, np.uint32), np.array(cols, np.uint32), frs(vals) if vals else np.zeros((0, 4), np.uint64)


def transpose_rows(rows, ncols):
    out = [[] for _ in range(ncols)]
    for r, row in enumerate(rows):
        for v, c in row:
            out[c].append((v, r))
    return out


def test_r1cs_3x4_transposed_by_eq_alpha(oracle, kats):
    # multiply_transposed_by_eq_alpha (sumcheck.rs:417-435) on the 3x4 R1CS
    k = kats["r1cs_3x4"]
    eqa = oracle.eq_evals(frs(k["alpha"]), 3)
    for name, key in (("A", "expected_a"), ("B", "expected_b"), ("C", "expected_c")):
        t = transpose_rows(k[name], 4)
        rs, cs, vs = csr_from_rows(t, 3)
        got = ints(oracle.sparse_mul(4, 3, rs, cs, vs, eqa))
        assert got == [v % P for v in k[key]], name


def map_py(a, b, c, e):
    (a0, a1), (b0, b1), (c0, c1), (e0, e1) = a, b, c, e
    f0 = e0 * (a0 * b0 - c0)
    fm1 = (2 * e0 - e1) * ((2 * a0 - a1) * (2 * b0 - b1) - (2 * c0 - c1))
    finf = (e1 - e0) * (a1 - a0) * (b1 - b0)
    return [f0 % P, fm1 % P, finf % P]


def round_py(mles, fold):
    size = len(mles[0])
    if fold is not None:
        q = size // 4
        for m in mles:
            for i in range(q):
                m[i] = (m[i] + fold * (m[i + 2 * q] - m[i])) % P
                m[i + q] = (m[i + q] + fold * (m[i + 3 * q] - m[i + q])) % P
        h = q
    else:
        h = size // 2
    tot = [0, 0, 0]
    for i in range(h):
        o = map_py(*[(m[i], m[i + h]) for m in mles])
        tot = [(x + y) % P for x, y in zip(tot, o)]
    return tot


def test_round_brute_force(oracle):
    rng = random.Random(7)
    size = 16
    mles = [[rng.randrange(P) for _ in range(size)] for _ in range(4)]
    arrs = [frs(m).copy() for m in mles]
    assert ints(oracle.sumcheck_round(*arrs)) == round_py([list(m) for m in mles], None)
    f = rng.randrange(P)
    py = [list(m) for m in mles]
    got = ints(oracle.sumcheck_round(*arrs, fold=fe(f)))
    assert got == round_py(py, f)
    for arr, m in zip(arrs, py):  # folded halves written back
        assert ints(arr[: size // 2]) == m[: size // 2]


def mle_eval(vals, point):
    v = list(vals)
    for x in point:
        h = len(v) // 2
        v = [(v[i] + x * (v[i + h] - v[i])) % P for i in range(h)]
    return v[0]


@pytest.mark.parametrize("m", [1, 2, 3, 6])
def test_plain_sumcheck_verifier_equation(oracle, m):
    rng = random.Random(m)
    n = 1 << m
    a, b, c = ([rng.randrange(P) for _ in range(n)] for _ in range(3))
    out = oracle.sumcheck_prove(frs(a), frs(b), frs(c), m, None, seed=5)
    assert out["ok"]
    r, alpha = ints(out["r"]), ints(out["alphas"])
    # the claimed sum is sum eq(r,x)(a b - c); each round's cubic hits it
    claim = sum(eq_py(r, [(i >> (m - 1 - j)) & 1 for j in range(m)]) * (a[i] * b[i] - c[i]) for i in range(n)) % P
    co = [ints(x) for x in out["coeffs"]]
    S = claim
    for k in range(m):
        assert (2 * co[k][0] + co[k][1] + co[k][2] + co[k][3]) % P == S
        x = alpha[k]
        S = (co[k][0] + x * (co[k][1] + x * (co[k][2] + x * co[k][3]))) % P
    assert unmont(out["final"]) == S
    # verifier's final check: f(alpha) = eq(r, alpha) (a(alpha) b(alpha) - c(alpha))
    fa = eq_py(r, alpha) * (mle_eval(a, alpha) * mle_eval(b, alpha) - mle_eval(c, alpha)) % P
    assert unmont(out["final"]) == fa


def zk_verify(out, m, a, b, c, g):
    """The reference verifier on a ZK-mode transcript (run_sumcheck_verifier,
    provekit/verifier/src/whir_r1cs.rs:266-312, and the final check at
    :249-257), with Python integers: the round sums start at rho * sum_g and
    hit every cubic's h(0) + h(1); f(alpha) = S_final - rho * blinding_eval
    must equal eq(r, alpha) (a(alpha) b(alpha) - c(alpha)).  Also the values
    the prover sends: sum_g = 2^(m-1) sum_i (g_i(0) + g_i(1)) (the sum of
    sum_i g_i(x_i) over the cube) and blinding_eval = sum_i g_i(alpha_i) (the
    dot product with the committed univariates, prefix_covector.rs:156-166).
    Returns the list of the checks that fail."""
    bad = []
    r, alpha = ints(out["r"]), ints(out["alphas"])
    rho, sum_g, beval = unmont(out["rho"]), unmont(out["sum_g"]), unmont(out["blinding_eval"])
    gi = [g[4 * i: 4 * i + 4] for i in range(m)]

    def cub(p, x):
        return (p[0] + x * (p[1] + x * (p[2] + x * p[3]))) % P

    if sum_g != pow(2, m - 1, P) * sum(cub(q, 0) + cub(q, 1) for q in gi) % P:
        bad.append("sum_g")
    if beval != sum(cub(q, x) for q, x in zip(gi, alpha)) % P:
        bad.append("blinding_eval")
    S = rho * sum_g % P
    co = [ints(x) for x in out["coeffs"]]
    for k in range(m):
        if (cub(co[k], 0) + cub(co[k], 1)) % P != S:
            bad.append(f"round {k}")
        S = cub(co[k], alpha[k])
    if unmont(out["final"]) != S:
        bad.append("final claim")
    f_alpha = (S - rho * beval) % P
    if f_alpha != eq_py(r, alpha) * (mle_eval(a, alpha) * mle_eval(b, alpha) - mle_eval(c, alpha)) % P:
        bad.append("final equation")
    return bad


def zk_case(m, seed, satisfied=True):
    rng = random.Random(seed)
    n = 1 << m
    a = [rng.randrange(P) for _ in range(n)]
    b = [rng.randrange(P) for _ in range(n)]
    c = [x * y % P for x, y in zip(a, b)]  # c = a o b (calculate_witness_bounds)
    if not satisfied:
        c[rng.randrange(n)] ^= 1
    g = [rng.randrange(P) for _ in range(4 * m)]
    return a, b, c, g


@pytest.mark.parametrize("m", [1, 2, 3, 5, 7])
def test_zk_mode_passes_reference_verifier(oracle, m):
    a, b, c, g = zk_case(m, 11 + m)
    out = oracle.sumcheck_prove(frs(a), frs(b), frs(c), m, frs(g), seed=3)
    assert out["ok"]
    assert zk_verify(out, m, a, b, c, g) == []
    # rho is squeezed after sum_g: init, squeeze(m) -> r, absorb(sum_g), squeeze(1)
    assert np.array_equal(oracle.transcript_run(3, None, m), out["r"])


def test_zk_mode_unsatisfied_witness_fails_final_equation(oracle):
    """With c != a o b the claim rho * sum_g is false and no honest round
    polynomials reach the verifier's final equation."""
    m = 5
    a, b, c, g = zk_case(m, 4, satisfied=False)
    out = oracle.sumcheck_prove(frs(a), frs(b), frs(c), m, frs(g), seed=9)
    assert zk_verify(out, m, a, b, c, g) == ["final equation"]


@pytest.mark.parametrize("fault,caught", [(1, {"sum_g", "final equation"}),
                                          (2, {"sum_g", "final equation"}),
                                          (3, {"final equation"}),
                                          (4, {"blinding_eval", "final equation"})])
def test_zk_mode_planted_mistakes_are_caught(oracle, fault, caught):
    """A wrong prefix multiplier, suffix multiplier, sign of the rho * g mix
    or a dropped term of blinding_eval, planted in the oracle, fails the
    verifier's checks (the round identity alone holds by construction)."""
    m = 5
    a, b, c, g = zk_case(m, 21)
    oracle.plant_fault(fault)
    try:
        out = oracle.sumcheck_prove(frs(a), frs(b), frs(c), m, frs(g), seed=3)
    finally:
        oracle.plant_fault(0)
    assert out["ok"]  # the identity c0 + c0 + c1 + c2 + c3 = S cannot see these
    assert set(zk_verify(out, m, a, b, c, g)) == caught


def test_zk_mode_needs_a_variable(oracle):
    z = frs([1])
    with pytest.raises(ValueError):
        oracle.sumcheck_prove(z, z, z, 0, frs([1, 2, 3, 4]), seed=1)


def test_plain_mode_zero_variables(oracle):
    """m = 0: the one-element MLEs are padded with a zero (pad_to_pow2_len_min2),
    so the claim is eq(x=0)(a0 b0 - c0) with the empty eq point's table [1, 0]."""
    a, b, c = 5, 7, 11
    out = oracle.sumcheck_prove(frs([a]), frs([b]), frs([c]), 0, None, seed=2)
    assert out["ok"]
    assert unmont(out["final"]) == (a * b - c) % P


# ---------------------------------------------------------------- GPU
def random_csr(rng, rows, cols, seed):
    starts, cidx = [], []
    for _ in range(rows):
        starts.append(len(cidx))
        k = rng.integers(1, 5)
        cidx.extend(sorted(rng.choice(cols, size=k, replace=False).tolist()))
    vals = random_fr(len(cidx), seed=seed, stream="csr")
    return np.array(starts, np.uint32), np.array(cidx, np.uint32), vals


@pytest.mark.gpu
@pytest.mark.parametrize("log_rows", [0, 4, 12, 18])
def test_gpu_sparse_mul(gpu, oracle, log_rows):
    rng = np.random.default_rng(log_rows)
    rows = cols = 1 << log_rows
    cols = max(cols, 4)
    rs, cs, vs = random_csr(rng, rows, cols, log_rows)
    z = random_fr(cols, seed=1, stream="z")
    A = gpu.SparseMatrix(rows, cols, rs, cs, vs)
    assert np.array_equal(A * z, oracle.sparse_mul(rows, cols, rs, cs, vs, z))


def long_row_cases():
    """Row-length shapes around the device's per-lane limit (32 entries) and
    its 2048-entry chunks: (name, rows, cols, row lengths)."""
    return [
        ("just-long", 40, 100, [33, 1, 32, 34] + [2] * 36),
        ("one-huge", 10, 30000, [1, 2, 25000, 3, 0, 0, 4, 1, 1, 1]),
        ("chunk-aligned", 6, 9000, [2048, 2048, 4096, 1, 33, 2047]),
        ("adjacent-long", 8, 5000, [1500, 1500, 1500, 40, 1, 3000, 0, 60]),
        ("many-long-in-chunk", 70, 200, [33] * 70),
        ("long-after-empty", 7, 4000, [0, 0, 0, 3500, 0, 2, 0]),
    ]


@pytest.mark.gpu
@pytest.mark.parametrize("case", long_row_cases(), ids=lambda c: c[0])
def test_gpu_sparse_mul_long_rows(gpu, oracle, case):
    name, rows, cols, lens = case
    rng = np.random.default_rng(len(name))
    rs, cs = [], []
    for k in lens:
        rs.append(len(cs))
        cs.extend(sorted(rng.choice(cols, size=k, replace=False).tolist()))
    rs, cs = np.array(rs, np.uint32), np.array(cs, np.uint32)
    vs = random_fr(len(cs), seed=3, stream="long-rows") if len(cs) else np.zeros((0, 4), np.uint64)
    z = random_fr(cols, seed=6, stream="long-rows-z")
    print(f"[spmv] {name} rows={rows} nnz={len(cs)}", flush=True)
    A = gpu.SparseMatrix(rows, cols, rs, cs, vs)
    assert np.array_equal(A * z, oracle.sparse_mul(rows, cols, rs, cs, vs, z))


@pytest.mark.gpu
def test_gpu_sparse_mul_constant_column_transposed(gpu, oracle):
    # a transposed R1CS matrix whose first row (the constant-one column)
    # holds an entry for most constraints: 2^18 rows, one row of ~2^17 entries
    n = 1 << 18
    rng = np.random.default_rng(21)
    rs, cs, vs = random_csr(rng, n, n, 21)
    rows_with_one = np.sort(rng.choice(n, size=n // 2, replace=False))
    # add column 0 to half the rows (keeping columns sorted and unique)
    new_rs, new_cs, new_vs = [], [], []
    ends = np.append(rs[1:], len(cs))
    mark = np.zeros(n, bool)
    mark[rows_with_one] = True
    extra = random_fr(n // 2, seed=8, stream="const-col")
    ei = 0
    for r in range(n):
        new_rs.append(len(new_cs))
        seg_c, seg_v = cs[rs[r]:ends[r]], vs[rs[r]:ends[r]]
        if mark[r] and (len(seg_c) == 0 or seg_c[0] != 0):
            new_cs.append(0)
            new_vs.append(extra[ei])
            ei += 1
        new_cs.extend(seg_c.tolist())
        new_vs.extend(seg_v)
    rs2, cs2 = np.array(new_rs, np.uint32), np.array(new_cs, np.uint32)
    vs2 = np.array(new_vs, np.uint64).reshape(-1, 4)
    t = oracle.sparse_transpose(n, n, rs2, cs2, vs2)
    z = random_fr(n, seed=9, stream="const-col-z")
    print(f"[spmv] transposed constant column: longest row {int(np.diff(np.append(t[0], len(t[1]))).max())}",
          flush=True)
    At = gpu.SparseMatrix(n, n, *t)
    assert np.array_equal(At * z, oracle.sparse_mul(n, n, *t, z))


@pytest.mark.gpu
def test_gpu_sparse_mul_device_malformed_stays_in_bounds(gpu, oracle):
    # device-resident matrices are not validated: row ranges are clipped to
    # [0, nnz] and entries with a column >= ncols are skipped
    rows, ncols = 64, 32
    rng = np.random.default_rng(9)
    cs = rng.integers(0, 48, size=200).astype(np.uint32)       # some columns >= ncols
    rs = np.sort(rng.integers(0, 230, size=rows)).astype(np.uint32)
    rs[5] = 250                                               # past nnz, and not monotone
    vs = random_fr(200, seed=4, stream="bad-csr")
    z = random_fr(ncols, seed=5, stream="bad-csr-z")
    nnz = 200
    # expected: the same clipping on the host
    keep_rs, keep_cs, keep_vs = [], [], []
    for r in range(rows):
        e = min(int(rs[r + 1]), nnz) if r + 1 < rows else nnz
        b = min(int(rs[r]), e)
        keep_rs.append(len(keep_cs))
        for i in range(b, e):
            if cs[i] < ncols:
                keep_cs.append(cs[i])
                keep_vs.append(vs[i])
    exp = oracle.sparse_mul(rows, ncols, np.array(keep_rs, np.uint32), np.array(keep_cs, np.uint32),
                            np.array(keep_vs, np.uint64).reshape(-1, 4), z)
    bufs = [gpu.DeviceBuffer.from_array(x) for x in (rs, cs, vs, z)]
    out = gpu.DeviceBuffer(32 * rows)
    gpu._call("pkgpu_sparse_mul_device", rows, ncols, bufs[0].ptr, nnz, bufs[1].ptr, bufs[2].ptr,
              bufs[3].ptr, out.ptr, None)
    gpu.synchronize()
    assert np.array_equal(out.download(shape=(rows, 4)), exp)


@pytest.mark.gpu
@pytest.mark.parametrize("rows,n_values", [(1, 1), (1000, 16), (1 << 16, 3), (1 << 18, 5000)])
def test_gpu_sparse_mul_interned(gpu, oracle, rows, n_values):
    """The interned form (ids into an interner, the reference's storage)
    against the oracle on the hydrated matrix, with long rows (chunked path)."""
    rng = np.random.default_rng(rows)
    cols = max(rows, 4096)
    rs, cs, _ = random_csr(rng, rows, cols, 5)
    if rows >= 1000:  # a few long rows across chunks
        extra = rng.choice(cols, size=3000, replace=False)
        cs = np.concatenate([cs[: rs[1]], np.sort(extra).astype(np.uint32), cs[rs[1]:]])
        rs = rs.copy()
        rs[1:] += 3000
    interner = random_fr(n_values, seed=n_values, stream="interner")
    ids = rng.integers(0, n_values, size=cs.shape[0]).astype(np.uint32)
    z = random_fr(cols, seed=rows, stream="interned-z")
    print(f"[case] interned spmv rows={rows} values={n_values} nnz={cs.shape[0]}", flush=True)
    got = gpu.sparse_mul_interned(rows, cols, rs, cs, ids, interner, z)
    exp = oracle.sparse_mul(rows, cols, rs, cs, np.ascontiguousarray(interner[ids]), z)
    assert np.array_equal(got, exp)
    # device form, ids past the table read as zero (in bounds)
    bad = ids.copy()
    bad[::7] = n_values + 5
    bufs = [gpu.DeviceBuffer.from_array(x) for x in (rs, cs, bad, interner, z)]
    out = gpu.DeviceBuffer(32 * rows)
    gpu._call("pkgpu_sparse_mul_interned_device", rows, cols, bufs[0].ptr, cs.shape[0], bufs[1].ptr,
              bufs[2].ptr, bufs[3].ptr, n_values, bufs[4].ptr, out.ptr, None)
    gpu.synchronize()
    vals0 = np.ascontiguousarray(interner[np.minimum(bad, n_values - 1)])
    vals0[bad >= n_values] = 0
    assert np.array_equal(out.download(shape=(rows, 4)), oracle.sparse_mul(rows, cols, rs, cs, vals0, z))
    with pytest.raises(gpu.ProvekitGpuError) as e:
        gpu.sparse_mul_interned(rows, cols, rs, cs, bad, interner, z)
    assert "Value not in interner." in str(e.value)


@pytest.mark.gpu
def test_gpu_sparse_mul_errors_and_kats(gpu, kats):
    import provekit_amd as pk
    A = gpu.SparseMatrix(2, 3, [0, 1], [0, 2], frs([5, 6]))
    with pytest.raises(pk.ProvekitGpuError) as e:
        A * frs([1, 2])
    assert "Vector length does not match number of columns." in str(e.value)
    k = kats["r1cs_3x4"]
    eqa = gpu.eq_evaluations(frs(k["alpha"]), 3)
    for name, key in (("A", "expected_a"), ("B", "expected_b"), ("C", "expected_c")):
        rs, cs, vs = csr_from_rows(transpose_rows(k[name], 4), 3)
        assert ints(gpu.SparseMatrix(4, 3, rs, cs, vs) * eqa) == [v % P for v in k[key]]


@pytest.mark.gpu
def test_gpu_witness_bounds(gpu, oracle):
    rng = np.random.default_rng(3)
    rows, cols = 1000, 1024
    A = gpu.SparseMatrix(rows, cols, *random_csr(rng, rows, cols, 10))
    B = gpu.SparseMatrix(rows, cols, *random_csr(rng, rows, cols, 11))
    w = random_fr(cols, seed=2, stream="w")
    a, b, c = gpu.calculate_witness_bounds(A, B, w)
    assert a.shape == (1024, 4)
    ea = oracle.sparse_mul(rows, cols, A.row_starts, A.cols, A.vals, w)
    eb = oracle.sparse_mul(rows, cols, B.row_starts, B.cols, B.vals, w)
    assert np.array_equal(a[:rows], ea) and not a[rows:].any()
    assert np.array_equal(b[:rows], eb) and not b[rows:].any()
    assert ints(c[:rows]) == [x * y % P for x, y in zip(ints(ea), ints(eb))] and not c[rows:].any()


@pytest.mark.gpu
@pytest.mark.parametrize("m", [0, 1, 2, 5, 12, 13, 20])
def test_gpu_eq(gpu, oracle, m):
    r = random_fr(m, seed=m, stream="eq") if m else np.zeros((0, 4), np.uint64)
    n = 1 << m
    assert np.array_equal(gpu.eq_evaluations(r, n), oracle.eq_evals(r, n))
    if n > 3:
        assert np.array_equal(gpu.eq_evaluations(r, n - 3), oracle.eq_evals(r, n - 3))


@pytest.mark.gpu
@pytest.mark.parametrize("m,entries", [(13, 1), (13, 5), (14, 129), (20, 1000), (40, 3), (62, 70)])
def test_gpu_eq_truncated_large_m(gpu, oracle, m, entries):
    """A few entries of an eq table over many variables: the half tables hold
    only the entries the truncated table reaches (at m = 62 the full upper
    half alone would be 64 GiB)."""
    r = random_fr(m, seed=m, stream="eqt")
    print(f"[case] eq m={m} entries={entries}", flush=True)
    assert np.array_equal(gpu.eq_evaluations(r, entries), oracle.eq_evals(r, entries))


@pytest.mark.gpu
@pytest.mark.parametrize("m", [1, 2, 6, 16])
def test_gpu_round(gpu, oracle, m):
    n = 1 << m
    arrs = [random_fr(n, seed=10 * m + j, stream="round") for j in range(4)]
    bufs = [gpu.DeviceBuffer.from_array(x) for x in arrs]
    ref = [x.copy() for x in arrs]
    got = gpu.sumcheck_r1cs_round(*bufs, n)
    exp = oracle.sumcheck_round(*ref)
    assert np.array_equal(got, exp)
    if n >= 4:
        f = random_fr(1, seed=m, stream="fold")[0]
        got = gpu.sumcheck_r1cs_round(*bufs, n, fold=f)
        exp = oracle.sumcheck_round(*ref, fold=f)
        assert np.array_equal(got, exp)
        for buf, x in zip(bufs, ref):
            assert np.array_equal(buf.download(shape=(n, 4))[: n // 2], x[: n // 2])


@pytest.mark.gpu
@pytest.mark.parametrize("m,mode", [(1, "plain"), (2, "plain"), (3, "plain"), (5, "plain"),
                                    (10, "plain"), (11, "plain"), (12, "plain"), (16, "plain"),
                                    (1, "ref"), (3, "ref"), (5, "ref"), (11, "ref"), (12, "ref")])
def test_gpu_sumcheck_prove(gpu, oracle, m, mode):
    n = 1 << m
    a = random_fr(n, seed=m, stream="sa")
    b = random_fr(n, seed=m, stream="sb")
    if mode == "plain":
        c = random_fr(n, seed=m, stream="sc")
        g = None
    else:
        c = np.array([fe(unmont(x) * unmont(y)) for x, y in zip(a, b)], dtype=np.uint64)
        g = random_fr(4 * m, seed=m, stream="g")
    print(f"[case] sumcheck m={m} mode={mode}", flush=True)
    got = gpu.sumcheck_prove(a, b, c, m, blinding=g, seed=77)
    exp = oracle.sumcheck_prove(a, b, c, m, g, seed=77)
    assert exp["ok"]
    assert np.array_equal(got["r"], exp["r"])
    assert np.array_equal(got["coeffs"], exp["coeffs"])
    assert np.array_equal(got["alphas"], exp["alphas"])
    assert np.array_equal(got["final"], exp["final"])
    assert np.array_equal(got["rho"], exp["rho"])
    assert np.array_equal(got["blinding_eval"], exp["blinding_eval"])
    if mode == "ref" and m <= 5:
        # the device's own output passes the reference verifier
        got["sum_g"] = exp["sum_g"]
        assert zk_verify(got, m, ints(a), ints(b), ints(c), ints(g)) == []


@pytest.mark.gpu
def test_gpu_sumcheck_zero_variables(gpu, oracle):
    """m = 0 in plain mode: one-element MLEs run zero-padded to two (the
    reference's pad_to_pow2_len_min2), nothing read past the caller's
    element; a blinding polynomial is rejected (the reference underflows)."""
    a, b, c = (random_fr(1, seed=k, stream="m0") for k in range(3))
    print("[case] sumcheck m=0", flush=True)
    got = gpu.sumcheck_prove(a, b, c, 0, seed=5)
    exp = oracle.sumcheck_prove(a, b, c, 0, None, seed=5)
    assert exp["ok"]
    assert np.array_equal(got["final"], exp["final"])
    with pytest.raises(gpu.ProvekitGpuError):
        gpu.sumcheck_prove(a, b, c, 0, blinding=random_fr(4, seed=1, stream="g"), seed=5)


@pytest.mark.gpu
@pytest.mark.parametrize("m", [20, 22])
def test_gpu_sumcheck_prove_large(gpu, oracle, m):
    """The whole plain sumcheck at 2^20 and 2^22 constraints (the grid rounds,
    the claim of round 0 and the one-workgroup tail at full size), bit-exact
    against the oracle: eq point, every round polynomial, every challenge and
    the final folded values."""
    n = 1 << m
    rng = np.random.default_rng(m)

    def canon(k):
        x = rng.integers(0, 1 << 63, size=(n, 4), dtype=np.uint64)
        x[:, 3] &= np.uint64((1 << 61) - 1)  # < 2^253 < p
        return x
    a, b, c = canon(0), canon(1), canon(2)
    print(f"[case] sumcheck m={m} plain (large)", flush=True)
    got = gpu.sumcheck_prove(a, b, c, m, blinding=None, seed=91)
    exp = oracle.sumcheck_prove(a, b, c, m, None, seed=91)
    assert exp["ok"]
    for key in ("r", "coeffs", "alphas", "final"):
        assert np.array_equal(got[key], exp[key]), key


@pytest.mark.gpu
def test_gpu_sparse_mul_2p22_rows(gpu, oracle):
    """HydratedSparseMatrix * z at 2^22 rows (1-4 entries per row, the
    benchmark's constraint count), bit-exact against the oracle."""
    from tests.test_transpose import random_csr as random_csr_np
    rows = cols = 1 << 22
    rng = np.random.default_rng(22)
    rs, cs, vs = random_csr_np(rng, rows, cols, seed=22)
    z = rng.integers(0, 1 << 63, size=(cols, 4), dtype=np.uint64)
    z[:, 3] &= np.uint64((1 << 61) - 1)  # < 2^253 < p
    print(f"[case] spmv 2^22 rows nnz={cs.shape[0]}", flush=True)
    A = gpu.SparseMatrix(rows, cols, rs, cs, vs)
    assert np.array_equal(A * z, oracle.sparse_mul(rows, cols, rs, cs, vs, z))
