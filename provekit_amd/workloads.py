"""Synthetic inputs of the benchmark workloads, shared by bench.py and the GPU
tests that check those workloads at full size.  Field elements come from
the device generator (pkgpu_random_fr_device: one ChaCha20 block per element,
independent per named stream); index structure from numpy's seeded PCG64."""
from __future__ import annotations

import numpy as np

# Montgomery form of 1 (R mod p)
ONE = np.array([0xAC96341C4FFFFFFB, 0x36FC76959F60CD29, 0x666EA36F7879462E, 0x0E0A77C19A07DF2F],
               dtype=np.uint64)


def batch_shapes(seed: int):
    """Uneven batch: 64 polynomials, log-lengths k ~ U{12..22}, blowup 2 or 4
    by coin flip."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(12, 23, size=64)
    blowups = rng.choice([2, 4], size=64)
    return [int(k) for k in ks], [int(b) for b in blowups]


def batch_polys(pk, seed: int):
    """The batch's polynomials on the device: polynomial i holds 2^k_i
    independent uniform coefficients (stream "batch-i").  Returns
    [(DeviceBuffer, n, blowup, coset=True)]."""
    ks, blowups = batch_shapes(seed)
    return [(pk.random_fr_device(1 << k, seed, f"batch-{i}"), 1 << k, b, True)
            for i, (k, b) in enumerate(zip(ks, blowups))]


def csr_structure(rng, n: int):
    """Row starts and columns of an n x n R1CS matrix: 1..4 entries per row
    at uniform columns, sorted and without repeats within a row (as
    SparseMatrix requires)."""
    k = rng.integers(1, 5, size=n)
    cand = np.sort(rng.integers(0, n, size=(n, 4)), axis=1)
    keep = np.ones((n, 4), bool)
    keep[:, 1:] = cand[:, 1:] != cand[:, :-1]
    keep &= np.arange(4)[None, :] < k[:, None]
    rs = np.zeros(n, np.uint32)
    rs[1:] = np.cumsum(keep.sum(axis=1))[:-1]
    return rs, cand[keep].astype(np.uint32)


def r1cs_job(pk, m: int, seed: int):
    """One serving job's inputs on the device: A, B, C with 2^m rows and
    columns (uniform Fr values) and the witness z (uniform, z[0] = 1).
    Returns (mats, z): mats = three (row_starts, cols, vals) DeviceBuffer
    triples."""
    n = 1 << m
    rng = np.random.default_rng([seed, m])
    mats = []
    for k in range(3):
        rs, cols = csr_structure(rng, n)
        vals = pk.random_fr_device(cols.shape[0], seed, f"r1cs-{m}-{k}")
        mats.append((pk.DeviceBuffer.from_array(rs), pk.DeviceBuffer.from_array(cols), vals))
    z = pk.random_fr_device(n, seed, f"witness-{m}")
    z.upload(ONE.reshape(1, 4))  # z[0] = 1
    return mats, z


def download_job(mats, z, m: int):
    """Host copies of a job's inputs (for the oracle)."""
    n = 1 << m
    host = []
    for rs, cols, vals in mats:
        nnz = cols.nbytes // 4
        host.append((rs.download(np.uint32), cols.download(np.uint32),
                     vals.download(shape=(nnz, 4))))
    return host, z.download(shape=(n, 4))


def p2_leaves(pk, seed: int, log_n: int = 24):
    """2^log_n uniform Fr leaves of the Poseidon2 tree, on the device."""
    return pk.random_fr_device(1 << log_n, seed, "p2-leaves")
