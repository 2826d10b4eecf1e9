"""Randomized parity sweep on the GPU (not collected by pytest): fresh seeds
and shapes for every entry-point family, each case compared with the oracle
(bit-exact, or the same rejection where the reference panics).

    python tests/sweep_parity.py [seconds=180] [seed=12345] [family,family,...]
    PK_SWEEP_LARGE=1 python tests/sweep_parity.py ...   (larger shapes)
    PK_SWEEP_THREADS=8 python tests/sweep_parity.py ... (host threads at once)

Families: batch, commit, ntt, inverse, sumcheck, hash_many, poseidon2, spmv,
wire (valid, truncated, trailing bytes, a corrupted interned value),
whir_commit, rs_encode, scalar, transpose, eq, pow, serve (several jobs per
server, openings), open, interned, transposed_eq, p2hash, bounds, device (chains of _device calls on
a caller's stream).  Time-bounded;
prints each case before it runs and exits non-zero on any mismatch."""
import os
import random
import sys
import threading
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import provekit_amd as gpu  # noqa: E402
from provekit_amd.inputs import random_fr  # noqa: E402
from oracle import pyoracle as oracle  # noqa: E402

BUDGET = float(sys.argv[1]) if len(sys.argv) > 1 else 180.0
SEED = int(sys.argv[2]) if len(sys.argv) > 2 else 12345
THREADS = int(os.environ.get("PK_SWEEP_THREADS", "1"))
_local = threading.local()


class _ThreadRng:
    """One random.Random per host thread (PK_SWEEP_THREADS workers)."""

    def __getattr__(self, name):
        return getattr(_local.r, name)


rng = _ThreadRng()
_local.r = random.Random(SEED)
_lock = threading.Lock()
# PK_SWEEP_LARGE=1: the same families at sizes where the multi-pass NTTs,
# the sumcheck's grid rounds and the chunked SpMV rows take over
LARGE = os.environ.get("PK_SWEEP_LARGE") == "1"


def big(small, large):
    return large if LARGE else small
t_end = time.time() + BUDGET
fails = []
counts = {}


def case(name, **kw):
    with _lock:
        print(f"[case] {name} {kw}", flush=True)
        counts[name] = counts.get(name, 0) + 1


def fail(name, kw, msg):
    with _lock:
        print(f"[FAIL] {name} {kw}: {msg}", flush=True)
        fails.append((name, kw, msg))


def c_batch():
    count = rng.randint(1, big(14, 6))
    leaf = rng.choice([2, 4, 8, 16, 32])
    ks, bs, cs = [], [], []
    for _ in range(count):
        while True:
            k = rng.randint(0, big(14, 18))
            b = rng.choice([1, 2, 4, 8])
            if (1 << k) * b >= leaf:
                break
        ks.append(k), bs.append(b), cs.append(rng.random() < 0.5)
    tile = rng.choice([0, 0, 6, 8, 9, 10, 11])
    seed = rng.randrange(1 << 30)
    kw = dict(ks=ks, blowups=bs, cosets=cs, leaf=leaf, tile=tile, seed=seed)
    case("batch", **kw)
    polys, host = [], []
    for i, (k, b, c) in enumerate(zip(ks, bs, cs)):
        co = random_fr(1 << k, seed=seed + i, stream="fuzz-batch")
        host.append((co, b, c))
        polys.append((gpu.DeviceBuffer.from_array(co), 1 << k, b, c))
    batch = gpu.BatchCommit(polys, leaf_elems=leaf, tile_log=tile)
    if rng.random() < 0.5:
        batch.run()
    else:
        batch.graph_launch()
    roots = batch.roots()
    for i, (co, b, c) in enumerate(host):
        er, ecw, en = oracle.commit(co, b, c, leaf)
        N = co.shape[0] * b
        cw = batch.codewords[i].download(shape=(N, 4))
        nodes = batch.nodes[i].download(np.uint8).reshape(-1, 32)
        if not (np.array_equal(cw, ecw) and np.array_equal(nodes, en) and roots[i] == er):
            fail("batch", kw, f"poly {i}")
            break
    batch.destroy()


def c_commit():
    leaf = rng.choice([2, 4, 8, 16, 32, 64])
    while True:
        k = rng.randint(big(0, 14), big(17, 20))
        b = rng.choice([1, 2, 4, 8])
        if (1 << k) * b >= leaf:
            break
    c = rng.random() < 0.5
    seed = rng.randrange(1 << 30)
    kw = dict(k=k, blowup=b, coset=c, leaf=leaf, seed=seed)
    case("commit", **kw)
    co = random_fr(1 << k, seed=seed, stream="fuzz-commit")
    r, cw, nodes = gpu.commit(co, b, c, leaf)
    er, ecw, en = oracle.commit(co, b, c, leaf)
    if not (r == er and np.array_equal(cw, ecw) and np.array_equal(nodes, en)):
        fail("commit", kw, "root/codeword/nodes")


def c_ntt():
    log = rng.randint(big(0, 16), big(17, 21))
    m_log = rng.randint(0, min(3, big(20, 22) - log))
    groups = 1 << rng.randint(0, log) if log else 1
    seed = rng.randrange(1 << 30)
    n = 1 << log
    total = n << m_log
    kw = dict(codeword=n, interleave=1 << m_log, num_groups=groups, seed=seed)
    case("ntt", **kw)
    v = random_fr(total, seed=seed, stream="fuzz-ntt")
    a, b = v.copy(), v.copy()
    ea = eb = None
    try:
        gpu.ntt_nr(a, n, groups)
    except Exception as e:
        ea = type(e).__name__
    try:
        oracle.ntt_nr(b, n, groups)
    except Exception as e:
        eb = type(e).__name__
    if (ea is None) != (eb is None):
        fail("ntt", kw, f"error mismatch gpu={ea} oracle={eb}")
    elif ea is None and not np.array_equal(a, b):
        fail("ntt", kw, "values")


def c_inverse():
    log = rng.randint(big(0, 16), big(18, 21))
    seed = rng.randrange(1 << 30)
    kind = rng.choice(["intt_rn", "intt_nr", "reverse"])
    kw = dict(n=1 << log, kind=kind, seed=seed)
    case("inverse", **kw)
    v = random_fr(1 << log, seed=seed, stream="fuzz-inv")
    a, b = v.copy(), v.copy()
    if kind == "intt_rn":
        gpu.intt_rn(a), oracle.intt_rn(b)
    elif kind == "intt_nr":
        gpu.intt_nr(a), oracle.intt_nr(b)
    else:
        cs = 1 << rng.randint(0, log)
        kw["codeword_size"] = cs
        gpu.reverse_order(a, cs), oracle.reverse_order(b, cs)
    if not np.array_equal(a, b):
        fail("inverse", kw, "values")


def c_sumcheck():
    m = rng.randint(big(0, 11), big(13, 18))
    zk = m >= 1 and rng.random() < 0.5
    seed = rng.randrange(1 << 30)
    tseed = rng.randrange(1 << 20)
    n = 1 << m
    kw = dict(m=m, zk=zk, seed=seed, transcript_seed=tseed)
    case("sumcheck", **kw)
    a = random_fr(n, seed=seed, stream="fuzz-a")
    b = random_fr(n, seed=seed, stream="fuzz-b")
    if rng.random() < 0.5:  # satisfied: c = a * b
        c = np.stack([oracle.fr_mul(a[i], b[i]) for i in range(n)]) if n <= 4096 else None
        if c is None:
            c = random_fr(n, seed=seed, stream="fuzz-c")
    else:
        c = random_fr(n, seed=seed, stream="fuzz-c")
    g = random_fr(4 * m, seed=seed, stream="fuzz-g") if zk else None
    got = gpu.sumcheck_prove(a, b, c, m, blinding=g, seed=tseed)
    exp = oracle.sumcheck_prove(a, b, c, m, blinding=g, seed=tseed)
    ok = all(np.array_equal(got[k], exp[k][: len(got[k])] if k in ("coeffs", "alphas", "r") else exp[k])
             for k in ("coeffs", "alphas", "r", "final", "rho", "blinding_eval"))
    if not ok:
        fail("sumcheck", kw, "transcript")


def c_hash_many():
    size = 32 * rng.choice([1, 2, 3, 4, 5, 8, 16, 31, 32])
    count = rng.randint(0, 3000)
    seed = rng.randrange(1 << 30)
    kw = dict(size=size, count=count, seed=seed)
    case("hash_many", **kw)
    data = random_fr(size // 32 * count, seed=seed, stream="fuzz-hm").tobytes() if count else b""
    eng = gpu.SkyscraperHashEngine()
    got = eng.hash_many(size, data, count)
    exp = oracle.sky_hash_many(size, data, count)
    if got != exp:
        fail("hash_many", kw, "digests")


def c_poseidon2():
    log = rng.randint(1, 14)
    seed = rng.randrange(1 << 30)
    kw = dict(leaves=1 << log, seed=seed)
    case("poseidon2_merkle", **kw)
    leaves = random_fr(1 << log, seed=seed, stream="fuzz-p2")
    got = gpu.poseidon2_merkle(leaves, want_nodes=True)
    exp = oracle.p2_merkle_levels(leaves)
    gn = got[1] if isinstance(got, tuple) else got
    if not np.array_equal(np.asarray(gn).reshape(-1), np.asarray(exp).reshape(-1)):
        fail("poseidon2_merkle", kw, "nodes")


def c_spmv():
    per_max = rng.choice(big([2, 5, 40, 300], [2, 5, 8, 3000]))
    rows = rng.randint(1, min(big(5000, 1 << 18), 4_000_000 // per_max))
    cols = rng.randint(1, big(5000, 1 << 18))
    seed = rng.randrange(1 << 30)
    nrng = np.random.default_rng(seed)
    per = nrng.integers(0, per_max, rows)
    rs = np.zeros(rows + 1, np.uint64)
    rs[1:] = np.cumsum(per)
    nnz = int(rs[-1])
    cs = nrng.integers(0, cols, nnz).astype(np.uint32)
    vs = random_fr(nnz, seed=seed, stream="fuzz-vals") if nnz else np.zeros((0, 4), np.uint64)
    z = random_fr(cols, seed=seed, stream="fuzz-z")
    kw = dict(rows=rows, cols=cols, nnz=nnz, seed=seed)
    case("spmv", **kw)
    rs = rs[:rows]
    A = gpu.SparseMatrix(rows, cols, rs, cs, vs)
    got = A * z
    exp = oracle.sparse_mul(rows, cols, rs, cs, vs, z)
    if not np.array_equal(np.asarray(got).reshape(-1, 4), np.asarray(exp).reshape(-1, 4)):
        fail("spmv", kw, "product")


from tests.test_wire import encode_r1cs, random_matrix, varint  # noqa: E402
from tests.util import P  # noqa: E402


def outcome(fn):
    try:
        return "ok", fn()
    except Exception as e:  # both sides' rejections
        return "err", type(e).__name__


def c_wire():
    seed = rng.randrange(1 << 30)
    nrng = np.random.default_rng(seed)
    ni = int(nrng.integers(1, 40))
    values = [int(x) for x in nrng.integers(0, 2**62, size=ni - 1)] + [P - 1]
    mats = [random_matrix(nrng, int(nrng.integers(0, 400)), int(nrng.integers(1, 600)), ni) for _ in range(3)]
    npub = int(nrng.integers(0, 5))
    blob = encode_r1cs(npub, values, mats)
    mode = rng.choice(["valid", "truncate", "trail", "value"])
    if mode == "truncate":
        blob = blob[: rng.randrange(len(blob))]
    elif mode == "trail":
        blob = blob + bytes(rng.randrange(256) for _ in range(rng.randint(1, 8)))
    elif mode == "value":  # one byte of one interned value (sizes untouched)
        start = len(varint(npub)) + len(varint(8 + 32 * ni)) + 8
        pos = start + rng.randrange(32 * ni)
        b = bytearray(blob)
        b[pos] ^= rng.randint(1, 255)
        blob = bytes(b)
    kw = dict(seed=seed, ni=ni, bytes=len(blob), mode=mode)
    case("wire", **kw)
    g = outcome(lambda: gpu.decode_r1cs(blob))
    o = outcome(lambda: oracle.r1cs_decode(blob))
    if g[0] != o[0]:
        fail("wire", kw, f"gpu {g} oracle {o[0]}")
        return
    if g[0] == "ok":
        got, want = g[1], o[1]
        if got["num_public_inputs"] != want["num_public_inputs"]:
            fail("wire", kw, "npub")
        for key in "abc":
            gm, (nr, nc, rs, cs, vs) = got[key], want[key]
            if not ((gm.num_rows, gm.num_cols) == (nr, nc) and np.array_equal(gm.row_starts, rs)
                    and np.array_equal(gm.cols, cs) and np.array_equal(gm.vals, vs)):
                fail("wire", kw, f"matrix {key}")


def c_whir_commit():
    nv = rng.randint(13, 15)
    m0 = rng.choice([0, 1, 2, 5, 8])
    n = 1 << nv
    nw = rng.randint(0, n - 4 * m0)
    blow = rng.choice([2, 4])
    leaf = rng.choice([8, 16])
    coset = rng.random() < 0.5
    seed = rng.randrange(1 << 30)
    kw = dict(num_vars=nv, m0=m0, witness=nw, blowup=blow, leaf=leaf, coset=coset, seed=seed)
    case("whir_commit", **kw)
    w = random_fr(nw, seed=seed, stream="fuzz-w") if nw else np.zeros((0, 4), np.uint64)
    g = random_fr(4 * m0, seed=seed, stream="fuzz-wg") if m0 else None
    root, poly, cw, nodes = gpu.whir_r1cs_commit(w, nv, g, blow, coset, leaf)
    epoly = oracle.whir_r1cs_shape(w, nv, g)
    eroot, ecw, enodes = oracle.commit(epoly, blow, coset, leaf)
    if not (np.array_equal(poly, epoly) and root == eroot and np.array_equal(cw, ecw)
            and np.array_equal(nodes, enodes)):
        fail("whir_commit", kw, "shape/commit")


def c_rs_encode():
    M = rng.randint(1, 4)
    L = rng.randint(1, 3000)
    r = rng.randint(0, 3)
    cw_len = 1 << rng.randint(0, 14)
    seed = rng.randrange(1 << 30)
    kw = dict(messages=M, length=L, masks=M * r, codeword=cw_len, seed=seed)
    case("rs_encode", **kw)
    msgs = [random_fr(L, seed=seed + i, stream="fuzz-rs") for i in range(M)]
    masks = random_fr(M * r, seed=seed, stream="fuzz-rsm") if r else np.zeros((0, 4), np.uint64)
    g = outcome(lambda: gpu.RSFr().interleaved_encode(msgs, masks, cw_len))
    o = outcome(lambda: oracle.rs_interleaved_encode(msgs, masks, cw_len))
    if g[0] != o[0]:
        fail("rs_encode", kw, f"gpu {g[0]} ({g[1] if g[0] == 'err' else ''}) oracle {o[0]}")
    elif g[0] == "ok" and not np.array_equal(g[1], o[1]):
        fail("rs_encode", kw, "codeword")


def c_scalar():
    n = rng.randint(1, 4000)
    seed = rng.randrange(1 << 30)
    kw = dict(n=n, seed=seed)
    case("scalar", **kw)
    a = random_fr(n, seed=seed, stream="fuzz-sa")
    b = random_fr(n, seed=seed, stream="fuzz-sb")
    gm, gs = gpu.scalar_mul_many(a, b), gpu.scalar_sqr_many(a)
    for i in range(0, n, max(1, n // 64)):
        if not (np.array_equal(gm[i], oracle.scalar_mul(a[i], b[i])) and
                np.array_equal(gs[i], oracle.scalar_sqr(a[i]))):
            fail("scalar", kw, f"element {i}")
            break


def c_transpose():
    rows = rng.randint(1, big(3000, 1 << 17))
    cols = rng.randint(1, big(3000, 1 << 17))
    seed = rng.randrange(1 << 30)
    nrng = np.random.default_rng(seed)
    per = nrng.integers(0, rng.choice([2, 6, 50]), rows)
    rs = np.concatenate([[0], np.cumsum(per)[:-1]]).astype(np.uint32)
    nnz = int(per.sum())
    cs = nrng.integers(0, cols, nnz).astype(np.uint32)
    vs = random_fr(nnz, seed=seed, stream="fuzz-tv") if nnz else np.zeros((0, 4), np.uint64)
    kw = dict(rows=rows, cols=cols, nnz=nnz, seed=seed)
    case("transpose", **kw)
    t = gpu.SparseMatrix(rows, cols, rs, cs, vs).transpose()
    trs, tcs, tvs = oracle.sparse_transpose(rows, cols, rs, cs, vs)
    if not (np.array_equal(t.row_starts, trs) and np.array_equal(t.cols, tcs) and np.array_equal(t.vals, tvs)):
        fail("transpose", kw, "csr")


def c_eq():
    m = rng.randint(0, big(15, 20))
    ne = rng.randint(1, 1 << m)
    seed = rng.randrange(1 << 30)
    kw = dict(m=m, num_entries=ne, seed=seed)
    case("eq", **kw)
    r = random_fr(m, seed=seed, stream="fuzz-eq") if m else np.zeros((0, 4), np.uint64)
    if not np.array_equal(gpu.eq_evaluations(r, ne), oracle.eq_evals(r, ne)):
        fail("eq", kw, "table")


def c_pow():
    bits = rng.uniform(0.0, 13.0)
    chb = bytes(rng.randrange(256) for _ in range(32))
    ch = np.frombuffer(chb, dtype=np.uint64).copy()
    kw = dict(bits=round(bits, 3), challenge=chb.hex())
    case("pow", **kw)
    got = gpu.pow.solve(ch, bits)
    want = oracle.pow_solve(ch, bits)
    if got != want:
        fail("pow", kw, f"gpu {got} oracle {want}")


from tests.test_serve import make_job, commit_vars  # noqa: E402


def c_serve():
    slots = rng.randint(1, 8)
    njobs = rng.randint(1, 6)
    blow = rng.choice([2, 4])
    coset = rng.random() < 0.5
    leaf = rng.choice([8, 16])
    specs = [(rng.randint(1, big(11, 15)), rng.randrange(1 << 20), rng.randrange(1 << 20)) for _ in range(njobs)]
    # a tight budget admits one or two jobs at a time: later ones wait for
    # earlier releases
    budget = rng.choice([2 << 30, 2 << 30, 3 << 20, 6 << 20])
    kw = dict(slots=slots, jobs=specs, blowup=blow, coset=coset, leaf=leaf, budget=budget)
    case("serve", **kw)
    srv = gpu.Server(hbm_budget_bytes=budget, slots=slots)
    subs = []
    for (m, jseed, tseed) in specs:
        mats, z = make_job(m, jseed)
        try:
            jid = srv.submit(m, mats, z, seed=tseed, blowup=blow, coset=coset, leaf_elems=leaf)
        except gpu.ProvekitGpuError as e:
            if e.code == gpu.ERR_BUDGET:  # this job alone exceeds the budget
                continue
            raise
        subs.append((m, mats, z, tseed, jid))
    for (m, mats, z, tseed, jid) in subs:
        res = srv.wait(jid, timeout_ms=120000)
        n = 1 << m
        abc = [oracle.sparse_mul(n, n, rs, cs, vs, z) for rs, cs, vs in mats]
        sc = oracle.sumcheck_prove(abc[0], abc[1], abc[2], m, None, seed=tseed)
        root, cw, nodes = oracle.commit(oracle.whir_r1cs_shape(z, commit_vars(m)), blow, coset, leaf)
        ok = res["root"] == root and all(np.array_equal(res[k], sc[k][: len(res[k])] if k != "final" else sc[k])
                                         for k in ("coeffs", "alphas", "r", "final"))
        leaves = (blow << commit_vars(m)) // leaf
        depth = leaves.bit_length() - 1
        idx = [rng.randrange(leaves) for _ in range(rng.randint(1, 8))]
        vals, paths = srv.open(jid, idx, leaf, depth)
        for q, i in enumerate(idx):
            if not np.array_equal(vals[q], cw[leaf * i:leaf * (i + 1)]):
                ok = False
            off, n_l = 0, leaves
            for lv in range(depth):
                if paths[q, lv].tobytes() != nodes[off + ((i >> lv) ^ 1)].tobytes():
                    ok = False
                off += n_l
                n_l //= 2
        srv.release(jid)
        if not ok:
            fail("serve", kw, f"job m={m}")
    srv.destroy()


def c_open():
    leaf = rng.choice([2, 4, 8, 16, 32])
    while True:
        k = rng.randint(0, big(15, 19))
        b = rng.choice([1, 2, 4, 8])
        if (1 << k) * b >= leaf:
            break
    coset = rng.random() < 0.5
    seed = rng.randrange(1 << 30)
    n, N = 1 << k, (1 << k) * b
    leaves = N // leaf
    kw = dict(k=k, blowup=b, leaf=leaf, coset=coset, seed=seed)
    case("open", **kw)
    c = random_fr(n, seed=seed, stream="fuzz-open")
    dc = gpu.DeviceBuffer.from_array(c)
    dcw, dn = gpu.DeviceBuffer(32 * N), gpu.DeviceBuffer(32 * (2 * leaves - 1))
    gpu.commit_device(dc, n, b, coset, leaf, dcw, dn)
    gpu.synchronize()
    root, ecw, en = oracle.commit(c, b, coset, leaf)
    idx = np.array([rng.randrange(leaves) for _ in range(rng.randint(0, 40))], np.uint64)
    values, paths = gpu.merkle_open_device(dcw, dn, leaves, leaf, idx)
    depth = leaves.bit_length() - 1
    for q, i in enumerate(idx.tolist()):
        if not np.array_equal(values[q], ecw[i * leaf:(i + 1) * leaf]):
            fail("open", kw, f"values {i}")
            return
        off, n_l = 0, leaves
        for lv in range(depth):
            if paths[q, lv].tobytes() != en[off + ((i >> lv) ^ 1)].tobytes():
                fail("open", kw, f"path {i} level {lv}")
                return
            off += n_l
            n_l //= 2


def rand_csr(nrng, rows, cols, per_max):
    per = nrng.integers(0, per_max, rows)
    rs = np.concatenate([[0], np.cumsum(per)[:-1]]).astype(np.uint32) if rows else np.zeros(0, np.uint32)
    nnz = int(per.sum())
    cs = nrng.integers(0, cols, nnz).astype(np.uint32)
    return rs, cs, nnz


def c_interned():
    rows, cols = rng.randint(1, big(4000, 1 << 17)), rng.randint(1, big(4000, 1 << 17))
    ni = rng.randint(1, 300)
    seed = rng.randrange(1 << 30)
    nrng = np.random.default_rng(seed)
    rs, cs, nnz = rand_csr(nrng, rows, cols, rng.choice([2, 6, 60]))
    ids = nrng.integers(0, ni, nnz).astype(np.uint32)
    tab = random_fr(ni, seed=seed, stream="fuzz-tab")
    z = random_fr(cols, seed=seed, stream="fuzz-iz")
    kw = dict(rows=rows, cols=cols, nnz=nnz, interner=ni, seed=seed)
    case("interned", **kw)
    got = gpu.sparse_mul_interned(rows, cols, rs, cs, ids, tab, z)
    vs = tab[ids] if nnz else np.zeros((0, 4), np.uint64)
    exp = oracle.sparse_mul(rows, cols, rs, cs, vs, z)
    if not np.array_equal(got, exp):
        fail("interned", kw, "product")
        return
    trs, tcs, tids = gpu.sparse_transpose_interned(rows, cols, rs, cs, ids)
    ers, ecs, evs = oracle.sparse_transpose(rows, cols, rs, cs, vs)
    tv = tab[np.asarray(tids, dtype=np.int64)] if nnz else np.zeros((0, 4), np.uint64)
    if not (np.array_equal(trs, ers) and np.array_equal(tcs, ecs) and np.array_equal(tv, evs)):
        fail("interned", kw, "transpose")


def c_transposed_eq():
    n = rng.randint(1, big(3000, 1 << 16))
    cols = rng.randint(1, big(3000, 1 << 16))
    seed = rng.randrange(1 << 30)
    nrng = np.random.default_rng(seed)
    mats = []
    for k in range(3):
        rs, cs, nnz = rand_csr(nrng, n, cols, rng.choice([2, 5, 30]))
        vs = random_fr(nnz, seed=seed + k, stream="fuzz-te") if nnz else np.zeros((0, 4), np.uint64)
        mats.append((rs, cs, vs))
    m = max((n - 1).bit_length(), 0)
    alpha = random_fr(m, seed=seed, stream="fuzz-alpha") if m else np.zeros((0, 4), np.uint64)
    kw = dict(constraints=n, witnesses=cols, m=m, seed=seed)
    case("transposed_eq", **kw)
    sm = [gpu.SparseMatrix(n, cols, *x) for x in mats]
    ts = gpu.transpose_r1cs_matrices(*sm)
    got = gpu.multiply_transposed_by_eq_alpha(*ts, alpha, n)
    e = oracle.eq_evals(alpha, n)
    for k, (rs, cs, vs) in enumerate(mats):
        t = oracle.sparse_transpose(n, cols, rs, cs, vs)
        if not np.array_equal(got[k], oracle.sparse_mul(cols, n, *t, e)):
            fail("transposed_eq", kw, f"matrix {k}")
            return


def c_p2hash():
    count = rng.randint(0, 2000)
    ln = rng.randint(1, 12)
    seed = rng.randrange(1 << 30)
    kw = dict(count=count, len=ln, seed=seed)
    case("p2hash", **kw)
    msg = random_fr(count * ln, seed=seed, stream="fuzz-p2h").reshape(count, ln, 4) if count else \
        np.zeros((0, ln, 4), np.uint64)
    got = gpu.poseidon2_hash_many(msg)
    for i in range(0, count, max(1, count // 32)):
        if not np.array_equal(got[i], oracle.p2_hash(msg[i])):
            fail("p2hash", kw, f"message {i}")
            return


def c_bounds():
    rows = rng.randint(1, big(3000, 1 << 16))
    cols = rng.randint(1, big(3000, 1 << 16))
    seed = rng.randrange(1 << 30)
    nrng = np.random.default_rng(seed)
    mats = []
    for k in range(2):
        rs, cs, nnz = rand_csr(nrng, rows, cols, rng.choice([2, 5, 20]))
        vs = random_fr(nnz, seed=seed + k, stream="fuzz-wb") if nnz else np.zeros((0, 4), np.uint64)
        mats.append((rs, cs, vs))
    w = random_fr(cols, seed=seed, stream="fuzz-ww")
    kw = dict(rows=rows, cols=cols, seed=seed)
    case("bounds", **kw)
    a, b, c = gpu.calculate_witness_bounds(*[gpu.SparseMatrix(rows, cols, *m) for m in mats], w)
    n = 1 << max((rows - 1).bit_length(), 0)
    ea, eb = (oracle.sparse_mul(rows, cols, *m, w) for m in mats)
    ec = np.stack([oracle.fr_mul(ea[i], eb[i]) for i in range(rows)])
    ok = a.shape[0] == n and all(np.array_equal(x[:rows], e) and not x[rows:].any()
                                 for x, e in ((a, ea), (b, eb), (c, ec)))
    if not ok:
        fail("bounds", kw, "vectors")


_hip = None


def user_stream():
    """A non-blocking stream of this thread, made with the HIP runtime (as a
    caller keeping data resident would), and a sync for it."""
    global _hip
    import ctypes
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    if getattr(_local, "stream", None) is None:
        st = ctypes.c_void_p()
        assert _hip.hipStreamCreateWithFlags(ctypes.byref(st), 1) == 0
        _local.stream = st.value
    return _local.stream, lambda: _hip.hipStreamSynchronize(_local.stream)


def c_device():
    """Device entry points queued back to back on a caller's non-blocking
    stream, one host sync at the end."""
    s, sync = user_stream()
    log = rng.randint(0, big(14, 20))
    n = 1 << log
    ops = []
    for _ in range(rng.randint(1, 4)):
        op = rng.choice(["ntt", "intt_rn", "intt_nr", "reverse"])
        arg = 1 << rng.randint(0, log) if op in ("ntt", "reverse") else 0
        ops.append((op, arg))
    leaf = rng.choice([2, 4, 8, 16])
    blow = rng.choice([1, 2, 4])
    coset = rng.random() < 0.5
    seed = rng.randrange(1 << 30)
    kw = dict(n=n, ops=ops, leaf=leaf, blowup=blow, coset=coset, seed=seed)
    case("device", **kw)
    x = random_fr(n, seed=seed, stream="fuzz-dev")
    d = gpu.DeviceBuffer.from_array(x)
    want = x.copy()
    for op, arg in ops:
        if op == "ntt":
            gpu.ntt_nr_device(d, n, n, arg, stream=s)
            oracle.ntt_nr(want, n, arg)
        elif op == "intt_rn":
            gpu.intt_rn_device(d, n, stream=s)
            oracle.intt_rn(want)
        elif op == "intt_nr":
            gpu.intt_nr_device(d, n, stream=s)
            oracle.intt_nr(want)
        else:
            gpu.reverse_order_device(d, n, arg, stream=s)
            oracle.reverse_order(want, arg)
    # then commit the result from the device and rehash its codeword
    N = n * blow
    if N >= leaf:
        leaves = N // leaf
        dcw, dn = gpu.DeviceBuffer(32 * N), gpu.DeviceBuffer(32 * (2 * leaves - 1))
        gpu.commit_device(d, n, blow, coset, leaf, dcw, dn, stream=s)
        dh, dt = gpu.DeviceBuffer(32 * leaves), gpu.DeviceBuffer(32 * (2 * leaves - 1))
        gpu.hash_many_device(32 * leaf, dcw, leaves, dh, stream=s)
        gpu.merkle_device(dh, leaves, dt, stream=s)
    assert sync() == 0
    if not np.array_equal(d.download(shape=(n, 4)), want):
        fail("device", kw, "transform chain")
        return
    if N >= leaf:
        root, cw, nodes = oracle.commit(want, blow, coset, leaf)
        got_nodes = dn.download(np.uint8).reshape(-1, 32)
        got_tree = dt.download(np.uint8).reshape(-1, 32)
        if not (np.array_equal(dcw.download(shape=(N, 4)), cw) and np.array_equal(got_nodes, nodes)
                and np.array_equal(got_tree, nodes)):
            fail("device", kw, "commit / rehash")


CASES = [c_batch, c_commit, c_ntt, c_inverse, c_sumcheck, c_hash_many, c_poseidon2, c_spmv,
         c_wire, c_whir_commit, c_rs_encode, c_scalar, c_transpose, c_eq, c_pow,
         c_serve, c_open, c_interned, c_transposed_eq, c_p2hash, c_bounds, c_device]
if len(sys.argv) > 3:
    CASES = [c for c in CASES if c.__name__[2:] in sys.argv[3].split(",")]
gpu.init(0)
oracle.lib()
total = [0]


def worker(t):
    # PK_SWEEP_THREADS > 1: host threads call the library at once (each has
    # its own stream), as rayon workers calling the plugin would
    _local.r = random.Random(SEED + 7919 * t)
    i = 0
    while time.time() < t_end:
        fn = CASES[(i + t) % len(CASES)] if i < len(CASES) else rng.choice(CASES)
        i += 1
        try:
            fn()
        except gpu.ProvekitGpuError as e:
            fail(fn.__name__, {}, str(e))
    with _lock:
        total[0] += i


if THREADS == 1:
    worker(0)
else:
    ths = [threading.Thread(target=worker, args=(t,)) for t in range(THREADS)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
print({"cases": counts, "total": total[0], "threads": THREADS, "failures": len(fails)}, flush=True)
for f in fails:
    print("FAILURE", f, flush=True)
sys.exit(1 if fails else 0)
