"""R1CS wire format: postcard encodings of R1CS (provekit/common/src/r1cs.rs:59-66),
Interner (interner.rs:6-10 + utils/serde_ark.rs:11-30) and SparseMatrix
(sparse_matrix.rs:73-160, 149-290) decoded to hydrated CSR.

The encoder below is written from postcard's published wire format (LEB128
varints for integers wider than a byte, a sequence as its varint length then
its elements, a struct as its fields in order, bytes as a varint length then
the raw bytes) and ark's compressed Vec<Fr> (u64 LE count, then 32-byte LE
canonical integers).  CPU: textbook LEB128 values, the reference's own delta
coding vectors (sparse_matrix.rs:651-787), the bytes of the reference's serde
round-trip matrix (sparse_matrix.rs:789-808) spelled out from the format,
encode -> decode round trips and rejection of malformed input by the oracle.
GPU: the device decoder bit-exact against the oracle, the same rejections,
and a full-size R1CS (2^22 constraints) round trip."""
import numpy as np
import pytest

from provekit_amd.inputs import random_fr
from tests.util import P, mont, unmont

M64 = (1 << 64) - 1


def varint(v: int) -> bytes:
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def seq(items, enc=varint) -> bytes:
    return varint(len(items)) + b"".join(enc(x) for x in items)


def deltas_of(row_starts, cols, nnz):
    """encode_col_deltas (sparse_matrix.rs:73-110), from its definition."""
    out = []
    for r, s in enumerate(row_starts):
        e = row_starts[r + 1] if r + 1 < len(row_starts) else nnz
        for k in range(s, e):
            out.append(cols[k] if k == s else cols[k] - cols[k - 1])
    return out


def encode_matrix(num_rows, num_cols, row_starts, cols, ids) -> bytes:
    return (varint(num_rows) + varint(num_cols) + seq(list(row_starts)) +
            seq(deltas_of(list(row_starts), list(cols), len(ids))) + seq(list(ids)))


def encode_interner(values) -> bytes:
    payload = len(values).to_bytes(8, "little") + b"".join(int(v).to_bytes(32, "little") for v in values)
    return varint(len(payload)) + payload


def encode_r1cs(npub, values, mats) -> bytes:
    return varint(npub) + encode_interner(values) + b"".join(encode_matrix(*m) for m in mats)


def random_matrix(rng, rows, cols, ni, per_row=(0, 5)):
    k = rng.integers(per_row[0], per_row[1], size=rows)
    rs, cs = [], []
    for r in range(rows):
        rs.append(len(cs))
        cs.extend(sorted(rng.choice(cols, size=min(int(k[r]), cols), replace=False).tolist()))
    ids = rng.integers(0, ni, size=len(cs)).tolist() if ni else []
    if not ni:
        rs, cs = [0] * rows, []
    return rows, cols, rs, cs, ids


def hydrated(values, ids):
    return np.array([mont(values[i]) for i in ids], dtype=np.uint64).reshape(-1, 4)


# ---------------------------------------------------------------- CPU: oracle
def test_varint_textbook_values(oracle):
    # LEB128 examples and the postcard limits (5 bytes for u32, 10 for u64)
    cases = [(0, b"\x00"), (1, b"\x01"), (127, b"\x7f"), (128, b"\x80\x01"), (300, b"\xac\x02"),
             (16384, b"\x80\x80\x01"), (2**32 - 1, b"\xff\xff\xff\xff\x0f"),
             (M64, b"\xff" * 9 + b"\x01")]
    for v, enc in cases:
        assert varint(v) == enc
        bits, mb = (32, 5) if v < 2**32 else (64, 10)
        assert oracle.varint_decode(enc, mb, bits) == (v, len(enc))
    assert oracle.varint_decode(b"\x80\x80\x80\x80\x00", 5, 32) == (0, 5)  # long form accepted
    for bad, mb, bits in [(b"\x80\x80\x80\x80\x10", 5, 32),   # past 32 bits
                          (b"\x80\x80\x80\x80\x80\x00", 5, 32),  # six bytes
                          (b"\xff" * 9 + b"\x02", 10, 64),    # past 64 bits
                          (b"\x80", 5, 32), (b"", 10, 64)]:   # truncated
        with pytest.raises(oracle.WireError):
            oracle.varint_decode(bad, mb, bits)


def test_reference_delta_vectors():
    # sparse_matrix.rs:663-787: (col_indices, new_row_indices, total, expected deltas)
    cases = [([3, 15, 100], [0], 3, [3, 12, 85]),
             ([0, 10, 20, 5, 15], [0, 3], 5, [0, 10, 10, 5, 10]),
             ([42], [0], 1, [42]),
             ([0, 5, 100], [0, 1, 2], 3, [0, 5, 100]),
             ([10, 11, 12, 13], [0], 4, [10, 1, 1, 1]),
             ([], [0, 0, 0], 0, []),
             ([3, 8], [0, 0, 0, 2], 2, [3, 5]),
             ([1_000_000, 1_000_001, 2_000_000], [0], 3, [1_000_000, 1, 999_999])]
    for cols, rows, total, want in cases:
        assert deltas_of(rows, cols, total) == want


def test_oracle_reference_serde_matrix_bytes(oracle):
    # sparse_matrix.rs:789-808: 3 x 100, (0,5)=id0, (0,20)=id1, (1,50)=id2;
    # new_row_indices [0, 2, 3], deltas [5, 15, 50], values [0, 1, 2]
    blob = bytes([3, 100, 3, 0, 2, 3, 3, 5, 15, 50, 3, 0, 1, 2])
    assert encode_matrix(3, 100, [0, 2, 3], [5, 20, 50], [0, 1, 2]) == blob
    vals = [1, 2, 3]
    nr, nc, rs, cs, vs = oracle.sparse_matrix_decode(blob, np.array([mont(v) for v in vals]))
    assert (nr, nc) == (3, 100)
    assert rs.tolist() == [0, 2, 3] and cs.tolist() == [5, 20, 50]
    assert [unmont(x) for x in vs] == [1, 2, 3]


@pytest.mark.parametrize("seed", range(4))
def test_oracle_r1cs_round_trip(oracle, seed):
    rng = np.random.default_rng(seed)
    ni = int(rng.integers(1, 40))
    values = [int(x) for x in rng.integers(0, 2**62, size=ni)] + [P - 1, 0][: max(0, 2 - seed)]
    ni = len(values)
    mats = [random_matrix(rng, int(rng.integers(0, 300)), int(rng.integers(1, 2000)), ni) for _ in range(3)]
    blob = encode_r1cs(7 + seed, values, mats)
    got = oracle.r1cs_decode(blob)
    assert got["num_public_inputs"] == 7 + seed and got["interner_len"] == ni
    for key, (rows, cols, rs, cs, ids) in zip("abc", mats):
        nr, nc, grs, gcs, gvs = got[key]
        assert (nr, nc) == (rows, cols)
        assert grs.tolist() == list(rs) and gcs.tolist() == list(cs)
        assert np.array_equal(gvs, hydrated(values, ids))
    assert [unmont(x) for x in oracle.interner_decode(encode_interner(values))] == values


def malformed_cases():
    ok = (4, 10, [0, 1, 1, 3], [2, 0, 9], [0, 1, 0])
    cases = {
        "truncated": encode_matrix(*ok)[:-1],
        "num-rows-mismatch": encode_matrix(5, 10, [0, 1, 1, 3], [2, 0, 9], [0, 1, 0]),
        "first-row-start": varint(2) + varint(10) + seq([1, 1]) + seq([2]) + seq([0]),
        "decreasing-rows": varint(3) + varint(10) + seq([0, 2, 1]) + seq([1, 1]) + seq([0, 0]),
        "row-past-values": varint(2) + varint(10) + seq([0, 3]) + seq([1, 1]) + seq([0, 0]),
        "delta-count": varint(1) + varint(10) + seq([0]) + seq([1, 1, 1]) + seq([0, 0]),
        "column-range": encode_matrix(4, 9, [0, 1, 1, 3], [2, 0, 9], [0, 1, 0]),
        "id-range": encode_matrix(4, 10, [0, 1, 1, 3], [2, 0, 9], [0, 1, 2]),
        "u32-overflow": varint(1) + varint(10) + seq([0]) + varint(1) + b"\xff\xff\xff\xff\x1f" + seq([0]),
        "long-varint": varint(1) + varint(10) + seq([0]) + varint(1) + b"\x81\x80\x80\x80\x80\x00" + seq([0]),
    }
    return cases


@pytest.mark.parametrize("name", list(malformed_cases()))
def test_oracle_rejects_malformed(oracle, name):
    blob = malformed_cases()[name]
    with pytest.raises(oracle.WireError):
        oracle.sparse_matrix_decode(blob, np.array([mont(5), mont(6)]))


def test_oracle_rejects_bad_interner(oracle):
    with pytest.raises(oracle.WireError):
        oracle.interner_decode(encode_interner([P]))  # not below the modulus
    with pytest.raises(oracle.WireError):
        oracle.interner_decode(encode_interner([1, 2])[:-1])


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_reference_serde_matrix(gpu):
    blob = bytes([3, 100, 3, 0, 2, 3, 3, 5, 15, 50, 3, 0, 1, 2])
    A = gpu.decode_sparse_matrix(blob, np.array([mont(v) for v in (1, 2, 3)]))
    assert (A.num_rows, A.num_cols) == (3, 100)
    assert A.row_starts.tolist() == [0, 2, 3] and A.cols.tolist() == [5, 20, 50]
    assert [unmont(x) for x in A.vals] == [1, 2, 3]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_gpu_r1cs_matches_oracle(gpu, oracle, seed):
    rng = np.random.default_rng(100 + seed)
    ni = int(rng.integers(1, 60))
    values = [int(x) for x in rng.integers(0, 2**62, size=ni - 1)] + [P - 1]
    mats = [random_matrix(rng, int(rng.integers(0, 3000)), int(rng.integers(1, 5000)), ni) for _ in range(3)]
    blob = encode_r1cs(seed, values, mats)
    print(f"[wire] seed={seed} bytes={len(blob)}", flush=True)
    got = gpu.decode_r1cs(blob)
    want = oracle.r1cs_decode(blob)
    assert got["num_public_inputs"] == want["num_public_inputs"]
    assert np.array_equal(got["interner"], oracle.interner_decode(encode_interner(values)))
    for key in "abc":
        g, (nr, nc, rs, cs, vs) = got[key], want[key]
        assert (g.num_rows, g.num_cols) == (nr, nc)
        assert np.array_equal(g.row_starts, rs) and np.array_equal(g.cols, cs)
        assert np.array_equal(g.vals, vs)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(malformed_cases()))
def test_gpu_rejects_malformed(gpu, name):
    blob = malformed_cases()[name]
    print(f"[wire] malformed {name}", flush=True)
    with pytest.raises(gpu.ProvekitGpuError):
        gpu.decode_sparse_matrix(blob, np.array([mont(5), mont(6)]))


@pytest.mark.gpu
def test_gpu_rejects_bad_interner(gpu):
    with pytest.raises(gpu.ProvekitGpuError):
        gpu.decode_interner(encode_interner([P]))
    # the largest 256-bit value among valid ones: rejected, and the product
    # that converts it is fed a zero in its place
    with pytest.raises(gpu.ProvekitGpuError):
        gpu.decode_interner(encode_interner([1, (1 << 256) - 1, P - 1]))
    assert [unmont(x) for x in gpu.decode_interner(encode_interner([1, P - 1]))] == [1, P - 1]


def encode_matrix_np(rs, cs, ids, rows, cols) -> bytes:
    """Vectorised encode_matrix for large matrices (same bytes)."""
    def varints(a):
        a = np.asarray(a, dtype=np.uint64)
        n = np.ones(a.shape, np.int64)
        for k in range(1, 10):
            n += (a >= (np.uint64(1) << np.uint64(7 * k))).astype(np.int64)
        out = np.zeros(int(n.sum()), np.uint8)
        pos = np.concatenate([[0], np.cumsum(n)[:-1]])
        for k in range(10):
            m = n > k
            byte = ((a[m] >> np.uint64(7 * k)) & np.uint64(0x7F)).astype(np.uint8)
            byte |= np.where(n[m] > k + 1, 0x80, 0).astype(np.uint8)
            out[pos[m] + k] = byte
        return out.tobytes()
    nnz = len(cs)
    d = np.asarray(cs, dtype=np.int64).copy()
    heads = np.zeros(nnz, bool)
    rsa = np.asarray(rs, dtype=np.int64)
    ends = np.append(rsa[1:], nnz)
    heads[rsa[rsa < ends]] = True
    d[1:] = d[1:] - np.asarray(cs, dtype=np.int64)[:-1]
    d[heads] = np.asarray(cs, dtype=np.int64)[heads]
    return (varint(rows) + varint(cols) + varint(len(rs)) + varints(rs) + varint(nnz) + varints(d) +
            varint(nnz) + varints(ids))


@pytest.mark.gpu
def test_gpu_r1cs_full_size(gpu, oracle):
    # 2^22 constraints, 1-4 entries per row, 16 interned values
    n = 1 << 22
    rng = np.random.default_rng(5)
    values = [int(x) for x in rng.integers(0, 2**62, size=16)]
    mats = []
    for _ in range(3):
        k = rng.integers(1, 5, size=n)
        cand = np.sort(rng.integers(0, n, size=(n, 4)), axis=1)
        keep = np.ones((n, 4), bool)
        keep[:, 1:] = cand[:, 1:] != cand[:, :-1]
        keep &= np.arange(4)[None, :] < k[:, None]
        rs = np.zeros(n, np.int64)
        rs[1:] = np.cumsum(keep.sum(axis=1))[:-1]
        cs = cand[keep]
        ids = rng.integers(0, 16, size=cs.shape[0])
        mats.append((rs, cs, ids))
    blob = varint(3) + encode_interner(values) + b"".join(encode_matrix_np(rs, cs, ids, n, n) for rs, cs, ids in mats)
    print(f"[wire] full size bytes={len(blob)}", flush=True)
    got = gpu.decode_r1cs(blob)
    inter = np.array([mont(v) for v in values], dtype=np.uint64)
    for key, (rs, cs, ids) in zip("abc", mats):
        g = got[key]
        assert (g.num_rows, g.num_cols, g.nnz) == (n, n, cs.shape[0])
        assert np.array_equal(g.row_starts, rs.astype(np.uint32))
        assert np.array_equal(g.cols, cs.astype(np.uint32))
        assert np.array_equal(g.vals, inter[ids])
    # the oracle decodes the same bytes to the same matrices
    want = oracle.r1cs_decode(blob)
    assert np.array_equal(want["a"][3], got["a"].cols) and np.array_equal(want["c"][4], got["c"].vals)
