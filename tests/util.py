"""Big-integer helpers: closed forms the oracle is pinned against."""
import numpy as np

P = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
R = 1 << 256
RINV = pow(R, -1, P)


# ---- big-integer helpers (closed forms for pinning the oracle)
def limbs(v: int) -> np.ndarray:
    return np.array([(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)


def to_int(a) -> int:
    a = np.asarray(a, dtype=np.uint64).reshape(4)
    return sum(int(a[i]) << (64 * i) for i in range(4))


def mont(v: int) -> np.ndarray:
    return limbs((v % P) * R % P)


def unmont(a) -> int:
    return to_int(a) * RINV % P


def fr_vec_ints(arr) -> list:
    arr = np.asarray(arr, dtype=np.uint64).reshape(-1, 4)
    return [unmont(x) for x in arr]
