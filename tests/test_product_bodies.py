"""The generated Montgomery product bodies (provekit_amd/csrc/gen_fr_mul.py),
checked on the CPU as integer identities.

fr_mul and fr_sqr run the same finely integrated product scanning: column k
sums its digit products, the quotient digit cancels the low word for k < 8,
and the low word is an output digit for k >= 8.  These tests evaluate the
generator's term lists with Python integers:

- the a*b terms of every column sum to a*b, and the m*p terms to m*p;
- the squaring's terms (each cross product once, against the doubled limbs)
  sum to a*a for every a < 2^255, including limbs with their top bit set;
- running the column schedule with those terms gives a*b*2^-256 mod p
  (closed form), lazy (< 2p) for lazy inputs, and the squaring schedule
  gives the same 256-bit result as the product schedule with b = a.

A wrong index or a dropped doubled limb in the generator fails here without
a GPU; the GPU tests then compare the kernels that use them with the oracle.
"""
import importlib.util
import os
import random

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
M32 = (1 << 32) - 1


@pytest.fixture(scope="module")
def gen():
    spec = importlib.util.spec_from_file_location(
        "gen_fr_mul", os.path.join(ROOT, "provekit_amd", "csrc", "gen_fr_mul.py"))
    g = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(g)
    return g


def limbs(x):
    return [(x >> (32 * i)) & M32 for i in range(8)]


def operand_env(a, b):
    la, lb = limbs(a), limbs(b)
    u = [(x << 1) & M32 for x in la]
    d = [0] + [((la[j] << 1) | (la[j - 1] >> 31)) & M32 for j in range(1, 8)]
    return {"a.l": la, "b.l": lb, "u": u, "d": d, "p_limb": limbs(P)}


def value(env, expr, m):
    name, idx = expr[:-1].split("(" if "(" in expr else "[")
    idx = int(idx)
    if name == "m":
        return m[idx]
    return env[name][idx]


def run_schedule(gen, a, b, square):
    """The column schedule of the generated body, on Python integers."""
    env = operand_env(a, b)
    p0inv = (-pow(P, -1, 1 << 32)) % (1 << 32)
    m = [0] * 8
    acc = 0
    out = []
    for k in range(15):
        for x, y, _ in gen.column_terms(k, square):
            acc += value(env, x, m) * value(env, y, m)
        if k < 8:
            m[k] = ((acc & M32) * p0inv) & M32
            acc += m[k] * limbs(P)[0]
            assert acc & M32 == 0
        else:
            out.append(acc & M32)
        acc >>= 32
    out.append(acc & M32)
    assert acc >> 32 == 0
    return sum(w << (32 * i) for i, w in enumerate(out))


def edge_values(rnd):
    vals = [0, 1, 2, P - 1, P, P + 1, 2 * P - 1, (1 << 254) - 1, (1 << 255) - 1,
            int("7fffffff" + "80000000" * 7, 16), int("7fffffff" + "ffffffff" * 7, 16)]
    for _ in range(300):  # random limbs with the top bit of each set or clear
        x = sum(((0x80000000 | rnd.getrandbits(31)) if rnd.random() < 0.5
                 else rnd.getrandbits(31)) << (32 * i) for i in range(7))
        vals.append(x + (rnd.getrandbits(31) << 224))
    vals += [rnd.getrandbits(255) for _ in range(700)]
    return vals


def test_square_terms_sum_to_a_squared(gen):
    rnd = random.Random(7)
    for a in edge_values(rnd):
        env = operand_env(a, a)
        tot = 0
        for k in range(15):
            for x, y, _ in gen.square_terms(k):
                tot += value(env, x, None) * value(env, y, None) << (32 * k)
        assert tot == a * a, hex(a)


def test_square_terms_count(gen):
    # each cross product once plus the diagonal: 28 + 8
    assert sum(len(gen.square_terms(k)) for k in range(15)) == 36


def test_product_and_square_schedules(gen):
    rnd = random.Random(11)
    rinv = pow(1 << 256, -1, P)
    lazy = [x % (2 * P) for x in edge_values(rnd)]
    for a in lazy[:400]:
        b = lazy[rnd.randrange(len(lazy))]
        r = run_schedule(gen, a, b, False)
        assert r % P == a * b * rinv % P
        assert r < 2 * P
        s = run_schedule(gen, a, a, True)
        assert s == run_schedule(gen, a, a, False)
        assert s < 2 * P


def test_generated_files_are_current(gen, tmp_path, monkeypatch):
    # the committed .inc files are what the generator writes today
    monkeypatch.setattr(gen, "HERE", str(tmp_path))
    gen.main()
    for name in ("fr_mul_asm.inc", "fr_sqr_asm.inc"):
        with open(os.path.join(ROOT, "provekit_amd", "csrc", name)) as f:
            committed = f.read()
        assert (tmp_path / name).read_text() == committed, name
