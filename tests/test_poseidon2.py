"""Poseidon2: the oracle pinned by the reference's KATs
(poseidon2/src/permutation.rs:183-194, hash.rs:61-78, constants.rs:969-1019),
and the GPU kernels bit-exact against the oracle."""
import os
import re

import numpy as np
import pytest

from provekit_amd.inputs import random_fr
from tests.util import P, mont, unmont

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def fe(h):
    return int(h, 16) % P


# ---------------------------------------------------------------- CPU: oracle
def test_permutation_zero_kat(oracle, kats):
    out = oracle.p2_permutation(np.zeros((4, 4), np.uint64))
    assert [unmont(x) for x in out] == [fe(h) for h in kats["p2_perm_zero"]]


def test_hash_kats(oracle, kats):
    assert unmont(oracle.p2_hash(np.array([mont(1)]))) == fe(kats["p2_hash_one"])
    e = mont(fe(kats["p2_hash_ee"]["e"]))
    assert unmont(oracle.p2_hash(np.array([e, e]))) == fe(kats["p2_hash_ee"]["out"])


def test_parameter_tables_match_samples(kats):
    # both copies of the parameter table (oracle and product) carry the
    # reference's sampled constants and diagonal
    for path, pat in ((os.path.join(ROOT, "oracle", "poseidon2_params.h"), r"OR_P2_(\w+)"),
                      (os.path.join(ROOT, "provekit_amd", "csrc", "poseidon2_params.hpp"), r"kP2(\w+)")):
        src = open(path).read()
        tables = {}
        for m in re.finditer(pat + r"\[[^=]*= \{(.*?)\n\};", src, re.S):
            tables[m.group(1)] = re.findall(r'"([0-9a-f]{64})"', m.group(2))
        names = {k.lower(): v for k, v in tables.items()}
        full1, full2 = names["full1"], names["full2"]
        partial, diag = names["partial"], names["diag"]
        assert len(full1) == 16 and len(full2) == 16 and len(partial) == 56 and len(diag) == 4
        s = kats["p2_constant_samples"]
        assert full1[0] == s["rc1[0][0]"] and full1[1 * 4 + 3] == s["rc1[1][3]"]
        assert full2[3 * 4 + 2] == s["rc2[3][2]"]
        assert partial[0] == s["rcp[0]"] and partial[-1] == s["rcp[rcp.len() - 1]"]
        assert diag == kats["p2_diag"]


def test_sponge_shapes(oracle):
    # rate boundaries (hash.rs:151-228): lengths 0..7 give distinct digests and
    # the length enters through the IV
    seen = set()
    for n in range(8):
        d = unmont(oracle.p2_hash(np.array([mont(i + 1) for i in range(n)], dtype=np.uint64).reshape(-1, 4)))
        assert d not in seen
        seen.add(d)


def test_merkle_levels_definition(oracle):
    leaves = np.array([mont(i * 7 + 1) for i in range(8)], dtype=np.uint64)
    nodes = oracle.p2_merkle_levels(leaves)
    k = 8
    lvl = list(leaves)
    while len(lvl) > 1:
        lvl = [oracle.p2_hash(np.array([lvl[2 * i], lvl[2 * i + 1]])) for i in range(len(lvl) // 2)]
        for x in lvl:
            assert np.array_equal(nodes[k], x)
            k += 1


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_permutation_many(gpu, oracle, kats):
    rng = np.random.default_rng(1)
    states = random_fr(4 * 1000, seed=3, stream="p2").reshape(1000, 4, 4)
    got = gpu.poseidon2_permutation(states)
    for i in range(0, 1000, 97):
        assert np.array_equal(got[i], oracle.p2_permutation(states[i]))
    zero = gpu.poseidon2_permutation(np.zeros((4, 4), np.uint64))
    assert [unmont(x) for x in zero] == [fe(h) for h in kats["p2_perm_zero"]]
    del rng


@pytest.mark.gpu
@pytest.mark.parametrize("ln", [0, 1, 2, 3, 4, 5, 6, 7, 16])
def test_gpu_hash_many(gpu, oracle, ln):
    msgs = random_fr(max(1, 50 * ln), seed=ln, stream="p2h")[: 50 * ln].reshape(50, ln, 4)
    got = gpu.poseidon2_hash_many(msgs)
    for i in range(50):
        assert np.array_equal(got[i], oracle.p2_hash(msgs[i])), (ln, i)


@pytest.mark.gpu
def test_gpu_hash_kats(gpu, kats):
    assert unmont(gpu.poseidon2_hash(np.array([mont(1)]))) == fe(kats["p2_hash_one"])
    e = mont(fe(kats["p2_hash_ee"]["e"]))
    assert unmont(gpu.poseidon2_hash(np.array([e, e]))) == fe(kats["p2_hash_ee"]["out"])


@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [0, 1, 3, 6, 7, 12, 16])
def test_gpu_merkle(gpu, oracle, log_n):
    leaves = random_fr(1 << log_n, seed=log_n, stream="p2m")
    print(f"[case] poseidon2 merkle 2^{log_n}", flush=True)
    root, nodes = gpu.poseidon2_merkle(leaves)
    exp = oracle.p2_merkle_levels(leaves)
    assert np.array_equal(nodes, exp)
    assert np.array_equal(root, exp[-1])


@pytest.mark.gpu
def test_gpu_merkle_2p24_sampled_paths(gpu, oracle):
    """The 2^24-leaf Poseidon2 tree of the benchmark workload: along the
    paths of sampled leaves every node is the oracle's two-element hash of its
    children, up to the returned root."""
    import numpy as np
    k = 24
    n = 1 << k
    rng = np.random.default_rng(24)
    leaves = rng.integers(0, 1 << 63, size=(n, 4), dtype=np.uint64)
    leaves[:, 3] &= np.uint64((1 << 61) - 1)  # < 2^253 < p: canonical
    print(f"[case] poseidon2 merkle 2^{k}", flush=True)
    root, nodes = gpu.poseidon2_merkle(leaves)
    assert np.array_equal(nodes[:n], leaves)
    picks = [0, n - 1] + [int(x) for x in rng.integers(0, n, 14)]
    for leaf in picks:
        i, off, width = leaf, 0, n
        while width > 1:
            i >>= 1
            up = off + width
            exp = oracle.p2_hash(nodes[off + 2 * i:off + 2 * i + 2])
            assert np.array_equal(nodes[up + i], exp), (leaf, width)
            off, width = up, width >> 1
    assert np.array_equal(root, nodes[-1])


@pytest.mark.gpu
def test_gpu_merkle_2p24_full_root(gpu, oracle):
    """The benchmark's Poseidon2 tree at full size (2^24 uniform leaves from
    the device generator, replayed from the captured graph): every node,
    the root included, against the oracle's whole tree."""
    from provekit_amd import workloads
    n = 1 << 24
    d_leaves = workloads.p2_leaves(gpu, seed=2)
    d_nodes = gpu.DeviceBuffer(32 * (2 * n - 1))
    print("[case] poseidon2 merkle 2^24 full tree vs oracle", flush=True)
    g = gpu.poseidon2_merkle_graph(d_leaves, n, d_nodes)
    g.launch()
    gpu.synchronize()
    leaves = d_leaves.download(shape=(n, 4))
    exp = oracle.p2_merkle_levels(leaves)
    got = d_nodes.download(shape=(2 * n - 1, 4))
    assert np.array_equal(got[-1], exp[-1])  # the root
    assert np.array_equal(got, exp)
    g.destroy()


@pytest.mark.gpu
def test_gpu_merkle_graph_replays(gpu, oracle):
    # the captured graph, replayed several times with eager work in between,
    # gives the eager result every time
    n = 1 << 12
    leaves = random_fr(n, seed=9, stream="p2g")
    d_leaves = gpu.DeviceBuffer.from_array(leaves)
    d_nodes = gpu.DeviceBuffer(32 * (2 * n - 1))
    g = gpu.poseidon2_merkle_graph(d_leaves, n, d_nodes)
    exp = oracle.p2_merkle_levels(leaves)
    for rep in range(3):
        gpu.synchronize()
        d_nodes.upload(np.zeros((2 * n - 1, 4), np.uint64))
        g.launch()
        gpu.synchronize()
        assert np.array_equal(d_nodes.download(shape=(2 * n - 1, 4)), exp), rep
        gpu.poseidon2_hash(np.array([mont(rep)]))  # eager work between replays
    g.destroy()
