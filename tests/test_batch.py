"""Batched uneven commitment (one persistent hashing launch from a device work
queue): every polynomial's codeword, tree and root equal the oracle's plain
commit of that polynomial, on plain runs and on hipGraph replays."""
import numpy as np
import pytest

from provekit_amd.inputs import random_fr

pytestmark = pytest.mark.gpu


def make_batch(gpu, ks, blowups, cosets, seed):
    polys, host = [], []
    for i, (k, b, c) in enumerate(zip(ks, blowups, cosets)):
        coeffs = random_fr(1 << k, seed=seed + i, stream="batch")
        host.append((coeffs, b, c))
        polys.append((gpu.DeviceBuffer.from_array(coeffs), 1 << k, b, c))
    return polys, host


def check(gpu, oracle, batch, host, roots):
    for i, (coeffs, b, c) in enumerate(host):
        eroot, ecw, enodes = oracle.commit(coeffs, b, c, batch.leaf_elems)
        N = coeffs.shape[0] * b
        cw = batch.codewords[i].download(shape=(N, 4))
        nodes = batch.nodes[i].download(np.uint8).reshape(-1, 32)
        assert np.array_equal(cw, ecw), f"codeword {i}"
        assert np.array_equal(nodes, enodes), f"tree {i}"
        assert roots[i] == eroot, f"root {i}"


def test_batch_uneven_small(gpu, oracle):
    ks = [4, 9, 10, 11, 12, 13, 14, 9, 16, 12, 10, 3]
    blowups = [2, 4, 2, 4, 2, 4, 2, 4, 2, 4, 2, 4]
    cosets = [True, False] * 6
    polys, host = make_batch(gpu, ks, blowups, cosets, seed=100)
    print(f"[case] batch of {len(ks)} polys, k={ks}", flush=True)
    batch = gpu.BatchCommit(polys, leaf_elems=16)
    batch.run()
    check(gpu, oracle, batch, host, batch.roots())
    # a second run on the same plan gives the same result
    for b in batch.nodes:
        b.upload(np.zeros(b.nbytes // 8, np.uint64))
    batch.run()
    check(gpu, oracle, batch, host, batch.roots())


@pytest.mark.parametrize("leaf_elems", [2, 4, 8, 32, 64])
def test_batch_leaf_sizes(gpu, oracle, leaf_elems):
    # leaf tasks fold 1..63 words through the four-word window: sizes below,
    # at and above the window, and trees from 2 leaves up
    ks = [6, 9, 12, 14, 10]
    blowups = [2, 4, 4, 2, 2]
    cosets = [True, False, True, True, False]
    polys, host = make_batch(gpu, ks, blowups, cosets, seed=500 + leaf_elems)
    print(f"[case] batch leaf_elems={leaf_elems} k={ks}", flush=True)
    batch = gpu.BatchCommit(polys, leaf_elems=leaf_elems)
    batch.run()
    check(gpu, oracle, batch, host, batch.roots())


def test_batch_single_coefficient_polynomials(gpu, oracle):
    # k = 0 has no NTT stage: the codeword is the coefficient in every
    # position, with and without the coset, beside transforms that have
    # passes, on plain runs and graph replays; a one-leaf tree is its digest
    ks = [0, 5, 0, 0, 9, 0]
    blowups = [2, 4, 8, 4, 2, 2]
    cosets = [True, False, False, True, True, False]
    polys, host = make_batch(gpu, ks, blowups, cosets, seed=900)
    print(f"[case] batch with single-coefficient polys k={ks}", flush=True)
    batch = gpu.BatchCommit(polys, leaf_elems=2)
    batch.run()
    check(gpu, oracle, batch, host, batch.roots())
    for b in batch.codewords:
        b.upload(np.zeros(b.nbytes // 8, np.uint64))
    batch.graph_launch()
    check(gpu, oracle, batch, host, batch.roots())


def test_batch_graph_replays(gpu, oracle):
    ks = [12, 15, 11, 17, 13]
    polys, host = make_batch(gpu, ks, [4, 2, 4, 2, 4], [True] * 5, seed=7)
    batch = gpu.BatchCommit(polys, leaf_elems=16)
    for rep in range(3):
        gpu.synchronize()
        for b in batch.nodes:
            b.upload(np.zeros(b.nbytes // 8, np.uint64))
        for c in batch.codewords:
            c.upload(np.zeros(c.nbytes // 8, np.uint64))
        batch.graph_launch()
        roots = batch.roots()
        check(gpu, oracle, batch, host, roots)
        gpu.poseidon2_hash(np.zeros((1, 4), np.uint64))  # eager work between replays


def test_batch_matches_plain_device_commits(gpu):
    ks = [14, 18, 12, 20]
    polys, host = make_batch(gpu, ks, [4, 2, 2, 4], [True, True, False, True], seed=3)
    batch = gpu.BatchCommit(polys, leaf_elems=16)
    batch.run()
    roots = batch.roots()
    for i, (coeffs, b, c) in enumerate(host):
        root, _, _ = gpu.commit(coeffs, b, c, 16, want_codeword=False, want_nodes=False)
        assert root == roots[i], i


def test_bench_batch_full_size(gpu, oracle):
    """The benchmark's own uneven batch at full size (64 polynomials, k in
    [12, 22], blowup 2 or 4, independent coefficients per polynomial): every
    root against the oracle's plain commit of that polynomial."""
    from provekit_amd import workloads
    seed = 3
    polys = workloads.batch_polys(gpu, seed)
    ks, blowups = workloads.batch_shapes(seed)
    print(f"[case] bench batch ks={ks}", flush=True)
    batch = gpu.BatchCommit(polys, leaf_elems=16)
    batch.run()
    roots = batch.roots()
    for i, (buf, n, b, c) in enumerate(polys):
        coeffs = buf.download(shape=(n, 4))
        eroot, _, _ = oracle.commit(coeffs, b, c, 16)
        assert roots[i] == eroot, (i, ks[i], b)
