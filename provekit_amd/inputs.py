"""Seeded synthetic inputs shared by the GPU path, the CPU baseline and the
tests.

The reference draws inputs with `Fr::rand` over rand_chacha / StdRng
(third-party streams that are not vendored), so this project defines its own
generator:
  keystream = ChaCha20 block function (RFC 8439), key = SHA-256(seed_le64 ||
              stream_id), nonce = 0, block counter 0, 1, 2, ...
  sample    = 32 keystream bytes -> 4 LE u64 limbs, top 2 bits cleared,
              rejected if >= p
The accepted 256-bit value v (< p) is used directly as the Montgomery limbs,
i.e. the field element v * 2^-256 mod p; this is uniform over Fr.
"""
from __future__ import annotations

import hashlib

import numpy as np

P = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
P_LIMBS = np.array([(P >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)


def _rotl(x: np.ndarray, n: int) -> np.ndarray:
    return (x << np.uint32(n)) | (x >> np.uint32(32 - n))


def chacha20_blocks(key: bytes, first_block: int, nblocks: int) -> np.ndarray:
    """nblocks 64-byte ChaCha20 blocks (RFC 8439, nonce 0) as uint8."""
    k = np.frombuffer(key, dtype="<u4")
    st = np.zeros((16, nblocks), dtype=np.uint32)
    st[0], st[1], st[2], st[3] = 0x61707865, 0x3320646E, 0x79622D32, 0x6B206574
    for i in range(8):
        st[4 + i] = k[i]
    st[12] = (np.arange(nblocks, dtype=np.uint64) + first_block).astype(np.uint32)
    x = st.copy()

    def qr(a, b, c, d):
        x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 16)
        x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 12)
        x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 8)
        x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 7)

    for _ in range(10):
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
    x += st
    return np.ascontiguousarray(x.T).view(np.uint8).reshape(nblocks * 64)


def _lt_p(v: np.ndarray) -> np.ndarray:
    """v: (n, 4) uint64 -> v < p (lexicographic from the top limb)."""
    lt = np.zeros(v.shape[0], dtype=bool)
    eq = np.ones(v.shape[0], dtype=bool)
    for i in (3, 2, 1, 0):
        lt |= eq & (v[:, i] < P_LIMBS[i])
        eq &= v[:, i] == P_LIMBS[i]
    return lt


def random_fr(n: int, seed: int, stream: str = "") -> np.ndarray:
    """n uniform Fr elements (Montgomery limbs), shape (n, 4) uint64."""
    key = hashlib.sha256(int(seed).to_bytes(8, "little") + stream.encode()).digest()
    out = np.empty((n, 4), dtype=np.uint64)
    have = 0
    block = 0
    while have < n:
        want = n - have
        nb = max(1, (want * 4 + 2) // 5)  # 2 samples per block, ~76% accepted
        ks = chacha20_blocks(key, block, nb).view("<u8").reshape(-1, 4).copy()
        block += nb
        ks[:, 3] &= np.uint64(0x3FFFFFFFFFFFFFFF)
        ok = ks[_lt_p(ks)]
        take = min(want, ok.shape[0])
        out[have:have + take] = ok[:take]
        have += take
    return out


def wide_key(seed: int, stream: str = "") -> bytes:
    """The ChaCha20 key of the device generator (pkgpu_random_fr_device)."""
    return hashlib.sha256(int(seed).to_bytes(8, "little") + stream.encode() + b"/wide").digest()


def random_fr_wide(n: int, seed: int, stream: str = "", first: int = 0) -> np.ndarray:
    """The device generator's elements first .. first + n - 1 on the host:
    element i is ChaCha20 block first + i (nonce 0) as a 512-bit LE integer
    mod p, used as the Montgomery limbs.  Python integers: for checks of
    the device generator, not for bulk data."""
    ks = chacha20_blocks(wide_key(seed, stream), first, n)
    out = np.empty((n, 4), dtype=np.uint64)
    for i in range(n):
        v = int.from_bytes(ks[64 * i: 64 * i + 64].tobytes(), "little") % P
        out[i] = [(v >> (64 * j)) & 0xFFFFFFFFFFFFFFFF for j in range(4)]
    return out


def to_int(limbs) -> int:
    a = np.asarray(limbs, dtype=np.uint64).reshape(4)
    return sum(int(a[i]) << (64 * i) for i in range(4))


def from_int(v: int) -> np.ndarray:
    return np.array([(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)


R = 1 << 256


def mont(v: int) -> np.ndarray:
    """Montgomery limbs of the field element v."""
    return from_int((v % P) * R % P)


def unmont(limbs) -> int:
    return to_int(limbs) * pow(R, -1, P) % P
