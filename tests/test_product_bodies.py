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
  gives the same 256-bit result as the product schedule with b = a;
- the terms each column sums without a carry add stay below 2^64 for every
  operand pair within the stated bound (top limbs summing to at most
  2^32 - 18), and operands past it are caught dropping a carry.

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
M64 = (1 << 64) - 1


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


def run_schedule(gen, a, b, square, overflow=None):
    """The column schedule of the generated body on 32/64-bit registers: c01
    (64 bits) and c2.  Each column opens with its no-carry terms, whose
    carry-out the asm drops (a drop is recorded in `overflow`), then the
    carrying terms count their carries into c2."""
    env = operand_env(a, b)
    p0inv = (-pow(P, -1, 1 << 32)) % (1 << 32)
    m = [0] * 8
    c01 = 0
    out = []
    for k in range(15):
        nc = gen.no_carry_terms(k, square)
        carrying = [t for t in gen.column_terms(k, square) if t not in nc]
        for x, y, _ in nc:
            c01 += value(env, x, m) * value(env, y, m)
            if c01 >> 64:
                if overflow is None:
                    raise AssertionError(f"no-carry term overflows in column {k}")
                overflow.append(k)
                c01 &= M64
        c2 = 0
        for x, y, _ in carrying:
            c01 += value(env, x, m) * value(env, y, m)
            c2 += c01 >> 64
            c01 &= M64
        if k < 8:
            m[k] = ((c01 & M32) * p0inv) & M32
            c01 += m[k] * limbs(P)[0]
            c2 += c01 >> 64
            c01 &= M64
            assert c01 & M32 == 0
        else:
            out.append(c01 & M32)
        assert c2 <= M32
        c01 = (c01 >> 32) | (c2 << 32)
    out.append(c01 & M32)
    assert c01 >> 32 == 0
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


def top_heavy(top):
    # the largest value with this top limb: every lower limb all ones
    return (top << 224) | ((1 << 224) - 1)


def test_no_carry_terms_within_bound(gen):
    """Operands at the edge of each case the callers pass: lazy x lazy, a
    canonical value times one below 4p, and top limbs summing to exactly
    2^32 - 18, with every lower limb all ones; squarings up to 2p - 1."""
    rinv = pow(1 << 256, -1, P)
    top = lambda x: x >> 224
    cases = [(2 * P - 1, 2 * P - 1), (P - 1, 4 * P - 1), (4 * P - 1, P - 1),
             (top_heavy(top(2 * P) - 1), top_heavy(top(2 * P) - 1)),
             (top_heavy(top(P) - 1), top_heavy(top(4 * P) - 1)),
             (top_heavy((1 << 31) - 9), top_heavy((1 << 31) - 9)),
             (top_heavy((1 << 32) - 18), top_heavy(0))]
    for a, b in cases:
        assert top(a) + top(b) <= (1 << 32) - 18
        r = run_schedule(gen, a, b, False)
        assert r % P == a * b * rinv % P and r < 1 << 256
    for a in (2 * P - 1, top_heavy(top(2 * P) - 1), P - 1):
        s = run_schedule(gen, a, a, True)
        assert s == run_schedule(gen, a, a, False) and s % P == a * a * rinv % P


def test_no_carry_overflow_is_detected(gen):
    # operands past the bound (top limbs summing to 2^32 and more) drop a
    # carry in the no-carry terms: the schedule reports it and the result
    # is wrong, so the bound above is what keeps the products exact
    rinv = pow(1 << 256, -1, P)
    for a, b in [((1 << 256) - 1, (1 << 256) - 1), (top_heavy(1 << 31), top_heavy(1 << 31))]:
        dropped = []
        r = run_schedule(gen, a, b, False, overflow=dropped)
        assert dropped, hex(a)
        assert r % P != a * b * rinv % P


def test_generated_files_are_current(gen, tmp_path, monkeypatch):
    # the committed .inc files are what the generator writes today
    monkeypatch.setattr(gen, "HERE", str(tmp_path))
    gen.main()
    for name in ("fr_mul_asm.inc", "fr_sqr_asm.inc"):
        with open(os.path.join(ROOT, "provekit_amd", "csrc", name)) as f:
            committed = f.read()
        assert (tmp_path / name).read_text() == committed, name


# ---- wait states: gfx950 needs two between a VALU write of an SGPR (or vcc)
# and a VALU read of it, a carry-in included (hipcc puts `s_nop 1` between
# the links of its own add-with-carry chains for gfx942/gfx950)
def asm_blocks(path):
    """The instruction lists of every asm statement in a source file, with the
    operand numbers the statement binds to SGPR outputs ("=&s" / "=s")."""
    import re
    src = open(path).read()
    out = []
    for m in re.finditer(r'asm\(\s*((?:"[^"]*"\s*)+):([^:;]*)(?::([^:;]*))?', src):
        text = "".join(re.findall(r'"([^"]*)"', m.group(1)))
        outs = re.findall(r'"([^"]*)"\s*\(', m.group(2))
        sgpr = {f"%{i}" for i, c in enumerate(outs) if "s" in c}
        insts = [t.strip() for t in text.replace("\\n", "\n").replace("\\t", "").split("\n") if t.strip()]
        out.append((insts, sgpr))
    return out


def sgpr_hazards(insts, sgpr):
    """(instruction, register, wait states seen) for every read of an SGPR
    or vcc fewer than two wait states after the VALU that wrote it."""
    last_write, pos, bad = {}, 0, []
    for ins in insts:
        op, _, rest = ins.partition(" ")
        args = [a.strip() for a in rest.split(",")] if rest else []
        if op == "s_nop":
            pos += int(args[0]) + 1
            continue
        reads, writes = [], []
        if op in ("v_addc_co_u32_e64", "v_subb_co_u32_e64"):
            reads.append(args[4])
            writes.append(args[1])
        elif op in ("v_add_co_u32_e64", "v_sub_co_u32_e64", "v_mad_u64_u32"):
            writes.append(args[1])
        elif op == "v_cndmask_b32_e64":
            reads.append(args[3])
        elif op == "v_cndmask_b32_e32":
            reads.append("vcc")
        elif op.endswith("_e32") and ("co_u32" in op or "cndmask" in op):
            if args[-1] == "vcc" and op not in ("v_add_co_u32_e32", "v_sub_co_u32_e32", "v_subrev_co_u32_e32"):
                reads.append("vcc")
            writes.append("vcc")
        for r in reads:
            if (r in sgpr or r == "vcc") and r in last_write and pos - last_write[r] - 1 < 2:
                bad.append((ins, r, pos - last_write[r] - 1))
        for w in writes:
            if w in sgpr or w == "vcc":
                last_write[w] = pos
        pos += 1
    return bad


@pytest.mark.parametrize("name", ["fr_mul_asm.inc", "fr_sqr_asm.inc", "fr.hpp"])
def test_carry_reads_wait_two_states(name):
    blocks = asm_blocks(os.path.join(ROOT, "provekit_amd", "csrc", name))
    assert blocks
    for insts, sgpr in blocks:
        assert sgpr_hazards(insts, sgpr) == [], insts


def test_hazard_checker_catches_a_back_to_back_carry():
    insts = ["v_mad_u64_u32 %0, %2, %3, %4, %0", "v_addc_co_u32_e64 %1, %2, 0, 0, %2"]
    assert sgpr_hazards(insts, {"%2"}) == [(insts[1], "%2", 0)]
    assert sgpr_hazards([insts[0], "s_nop 1", insts[1]], {"%2"}) == []
    vcc = ["v_add_co_u32_e32 %0, vcc, %1, %2", "v_addc_co_u32_e32 %3, vcc, %4, %5, vcc"]
    assert sgpr_hazards(vcc, set()) == [(vcc[1], "vcc", 0)]
