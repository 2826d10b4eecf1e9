"""GPU parity: Skyscraper compression, hash_many, Merkle trees and the
one-polynomial commit, bit-exact against the oracle."""
import numpy as np
import pytest

from provekit_amd.inputs import random_fr
from tests.util import P, limbs

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


EDGE = [0, 1, P - 1, P, P + 1, 2 * P - 1, 2 * P, 4 * P, 5 * P, (1 << 256) - 1, 1 << 255]


def test_compress_many_random_and_edges(gpu, oracle):
    rng = np.random.default_rng(5)
    rand = rng.integers(0, 2**64, size=(4096, 8), dtype=np.uint64)
    edge = np.array([np.concatenate([limbs(a), limbs(b)]) for a in EDGE for b in EDGE], dtype=np.uint64)
    msgs = np.concatenate([rand, edge]).astype(np.uint64)
    say(f"compress_many n={msgs.shape[0]}")
    got = gpu.skyscraper.compress_many(msgs.tobytes())
    exp = oracle.sky_compress_many(msgs.tobytes())
    assert got == exp


def test_compress_single_and_kats(gpu, oracle, kats):
    for name in ("test_zero", "test_random"):
        l, r = limbs(int(kats[name]["l"])), limbs(int(kats[name]["r"]))
        assert np.array_equal(gpu.skyscraper.compress(l, r), oracle.sky_compress_spec(l, r))
    assert np.array_equal(gpu.skyscraper.compress([1, 2, 3, 4], [5, 6, 7, 8]),
                          oracle.sky_compress([1, 2, 3, 4], [5, 6, 7, 8]))


@pytest.mark.parametrize("size,count", [(32, 5), (64, 1000), (96, 77), (512, 3), (512, 4096), (1024, 300)])
def test_hash_many(gpu, oracle, size, count):
    rng = np.random.default_rng(size + count)
    data = rng.integers(0, 2**64, size=(count * size // 8,), dtype=np.uint64).tobytes()
    say(f"hash_many size={size} count={count}")
    assert gpu.SkyscraperHashEngine().hash_many(size, data, count) == oracle.sky_hash_many(size, data, count)


def test_hash_many_errors(gpu):
    import provekit_amd as pk
    eng = gpu.SkyscraperHashEngine()
    for bad in (0, 1, 31, 33):
        assert not eng.supports_size(bad)
    with pytest.raises(pk.ProvekitGpuError) as e:
        eng.hash_many(31, b"\0" * 31, 1)
    assert "unsupported message size 31" in str(e.value)
    with pytest.raises(pk.ProvekitGpuError) as e:
        eng.hash_many(64, b"\0" * 100, 2)
    assert "input length 100 != size 64 * count 2" in str(e.value)
    assert eng.hash_many(64, b"", 0) == b""


@pytest.mark.parametrize("leaves_log", [0, 1, 5, 11, 12, 13])
def test_merkle_device(gpu, oracle, leaves_log):
    n = 1 << leaves_log
    rng = np.random.default_rng(leaves_log)
    leaves = rng.integers(0, 2**64, size=(n, 4), dtype=np.uint64)
    say(f"merkle n=2^{leaves_log}")
    d_nodes = gpu.DeviceBuffer(32 * (2 * n - 1))
    d_leaves = gpu.DeviceBuffer.from_array(leaves)
    gpu.merkle_device(d_leaves, n, d_nodes)
    got = d_nodes.download(np.uint8).reshape(2 * n - 1, 32)
    exp = oracle.merkle_tree(leaves.view(np.uint8).reshape(n, 32))
    assert np.array_equal(got, exp)


def test_merkle_device_2p26_sampled_paths(gpu, oracle):
    """A 2^26-leaf tree (4 GiB of nodes, 64-bit node offsets): the leaves are
    copied through, and along the paths of sampled leaves every node is the
    oracle's compression of its two children, up to the root."""
    k = 26
    n = 1 << k
    rng = np.random.default_rng(26)
    leaves = rng.integers(0, 2**64, size=(n, 4), dtype=np.uint64)
    say(f"merkle n=2^{k}")
    d_nodes = gpu.DeviceBuffer(32 * (2 * n - 1))
    d_leaves = gpu.DeviceBuffer.from_array(leaves)
    gpu.merkle_device(d_leaves, n, d_nodes)
    nodes = d_nodes.download(np.uint64).reshape(2 * n - 1, 4)
    d_nodes.free()
    d_leaves.free()
    assert np.array_equal(nodes[:n], leaves)
    picks = [0, n - 1] + [int(x) for x in rng.integers(0, n, 14)]
    for leaf in picks:
        i, off, width = leaf, 0, n
        while width > 1:
            i >>= 1
            up = off + width
            exp = oracle.sky_compress(nodes[off + 2 * i], nodes[off + 2 * i + 1])
            assert np.array_equal(nodes[up + i], exp), (leaf, width)
            off, width = up, width >> 1
    assert off == 2 * n - 2


@pytest.mark.parametrize("log_n,blowup,coset", [(4, 4, True), (8, 2, False), (12, 4, True), (12, 4, False),
                                                 (14, 2, True)])
def test_commit_small(gpu, oracle, log_n, blowup, coset):
    c = random_fr(1 << log_n, seed=log_n, stream="commit")
    say(f"commit 2^{log_n} blowup={blowup} coset={coset}")
    root, cw, nodes = gpu.commit(c, blowup, coset, 16)
    eroot, ecw, enodes = oracle.commit(c, blowup, coset, 16)
    assert np.array_equal(cw, ecw)
    assert np.array_equal(nodes, enodes)
    assert root == eroot


@pytest.mark.parametrize("log_n,blowup,leaf", [(2, 4, 16), (6, 2, 2), (8, 4, 8), (10, 2, 64), (9, 4, 1),
                                                (11, 2, 4), (13, 2, 2), (12, 4, 32)])
def test_commit_leaf_shapes(gpu, oracle, log_n, blowup, leaf):
    # one leaf, tails of every size up to 64 nodes, messages that are not a
    # multiple of the four-word read group, single-word leaves
    c = random_fr(1 << log_n, seed=100 + log_n, stream="commit-leaves")
    say(f"commit 2^{log_n} blowup={blowup} leaf={leaf}")
    root, cw, nodes = gpu.commit(c, blowup, True, leaf)
    eroot, ecw, enodes = oracle.commit(c, blowup, True, leaf)
    assert np.array_equal(cw, ecw)
    assert np.array_equal(nodes, enodes)
    assert root == eroot


def test_commit_2p20_blowup4_full_size(gpu, oracle):
    # 2^20 coefficients, blowup 4, coset NTT to 2^22, leaves of 16, Skyscraper tree
    c = random_fr(1 << 20, seed=2, stream="commit-2^20")
    say("commit 2^20 -> 2^22 (full size)")
    root, cw, nodes = gpu.commit(c, 4, True, 16)
    eroot, ecw, enodes = oracle.commit(c, 4, True, 16)
    assert np.array_equal(cw, ecw), "codeword"
    assert np.array_equal(nodes[: 1 << 18], enodes[: 1 << 18]), "leaf digests"
    assert np.array_equal(nodes, enodes), "tree"
    assert root == eroot


@pytest.mark.parametrize("log_n,blowup,leaf", [(3, 2, 16), (10, 4, 16), (12, 2, 4), (14, 4, 16)])
def test_merkle_open_from_resident_commit(gpu, oracle, log_n, blowup, leaf):
    # the tree stays on the device; openings come back against the oracle's
    # codeword and nodes, and each path hashes back to the root
    n = 1 << log_n
    c = random_fr(n, seed=300 + log_n, stream="open")
    N = n * blowup
    leaves = N // leaf
    dc = gpu.DeviceBuffer.from_array(c)
    dcw, dnodes = gpu.DeviceBuffer(32 * N), gpu.DeviceBuffer(32 * (2 * leaves - 1))
    say(f"open from resident commit 2^{log_n} x{blowup} leaf={leaf}")
    gpu.commit_device(dc, n, blowup, True, leaf, dcw, dnodes)
    gpu.synchronize()
    root, ecw, enodes = oracle.commit(c, blowup, True, leaf)
    rng = np.random.default_rng(log_n)
    idx = np.concatenate([[0, leaves - 1], rng.integers(0, leaves, size=64)]).astype(np.uint64)
    values, paths = gpu.merkle_open_device(dcw, dnodes, leaves, leaf, idx)
    depth = leaves.bit_length() - 1
    for q, i in enumerate(idx.tolist()):
        assert np.array_equal(values[q], ecw[i * leaf:(i + 1) * leaf])
        # Merkle verification from the definition: leaf digest, then siblings
        h = np.frombuffer(enodes[i].tobytes(), np.uint64)
        for l in range(depth):
            sib = np.frombuffer(paths[q, l].tobytes(), np.uint64)
            h = oracle.sky_compress(sib, h) if (i >> l) & 1 else oracle.sky_compress(h, sib)
        assert h.tobytes() == root
    with pytest.raises(gpu.ProvekitGpuError):
        gpu.merkle_open_device(dcw, dnodes, leaves, leaf, [leaves])


def test_merkle_open_poseidon2_tree(gpu, oracle):
    # a Poseidon2 tree has the same node layout: leaves of one element
    n = 1 << 12
    x = random_fr(n, seed=41, stream="p2-open")
    dx = gpu.DeviceBuffer.from_array(x)
    dn = gpu.DeviceBuffer(32 * (2 * n - 1))
    g = gpu.poseidon2_merkle_graph(dx, n, dn)
    g.launch()
    gpu.synchronize()
    enodes = oracle.p2_merkle_levels(x)
    idx = np.array([0, 1, 2, 777, n - 1], np.uint64)
    values, paths = gpu.merkle_open_device(dx, dn, n, 1, idx)
    for q, i in enumerate(idx.tolist()):
        assert np.array_equal(values[q, 0], x[i])
        for l in range(12):
            off = 2 * n - 2 * (n >> l)
            assert paths[q, l].tobytes() == enodes[off + ((i >> l) ^ 1)].tobytes()
    g.destroy()
