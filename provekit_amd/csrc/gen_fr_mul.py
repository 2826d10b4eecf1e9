"""Writes fr_mul_asm.inc and fr_sqr_asm.inc, the bodies of fr_mul and fr_sqr
in fr.hpp: the 8 x 32-bit Montgomery product with one inline-asm statement
per column (run: python3 gen_fr_mul.py, from this folder).  The squaring
forms each cross product of a * a once against the doubled limbs of a (36
digit products instead of 64; see square_terms).

Why per column: the compiler cannot see into inline asm, so after every asm
statement it inserts a wait state before the next use of the statement's
outputs (the statement could hold an instruction with a hazard).  With one
statement per multiply-add pair that was ~130 s_nop per product, and a lone
wave (tree tops, transcript steps) spends a full issue slot on each: 1700
cycles per product.  One statement per column, plus one per quotient digit,
leaves ~23 and measures 1400 cycles; at full occupancy the rate is the same
(the nops never held the VALU).  A single statement for the whole product
(accumulator in fixed registers) measured the same latency and a slightly
lower rate.

Algorithm (finely integrated product scanning): column k of a*b and of m*p
accumulates in the 96-bit (c2 : c01); for k < 8 the quotient digit
m_k = lo * (-p^-1 mod 2^32) = lo * 0xefffffff mod 2^32 makes the low word
cancel: it is the low word of one v_mad_u64_u32 (the instruction the
products are measured full-rate with; written as a 32-bit product in C the
compiler emits v_mul_lo_u32 instead, and -(lo + (lo << 28)) took two
instructions); for k >= 8 the low word is output digit k - 8; then the
accumulator shifts down one word (in C: register renaming and one move).
The first add-with-carry of each column writes c2 from zero instead of
adding into it, which saves a zeroing move per column (15 per product: on a
lone wave the unrolled Skyscraper compression went 30.3k -> 28.5k cycles).
In columns 7-14 the products that hold a top limb come first and take no
carry add at all (see no_carry_terms for the operand bound that allows it)."""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def square_terms(k):
    """Column k of a*a with each cross product once: row i takes the doubled
    limbs of a above it, d[j] = (a_j << 1) | (a_{j-1} >> 31) for j >= i + 2
    and u[j] = a_j << 1 for j = i + 1 (a_i's own top bit belongs to row i's
    diagonal side, not to 2 * A_{>i}).  Then sum_i a_i * 2 * A_{>i} equals the
    cross-product sum exactly, so every column, quotient digit and output is
    the one fr_mul(a, a) forms.  Needs a < 2^255 (2a fits in eight limbs)."""
    terms = []
    for i in range(8):
        j = k - i
        if i < j <= 7:
            terms.append((f"a.l[{i}]", f"u[{j}]" if j == i + 1 else f"d[{j}]", "v"))
    if k % 2 == 0 and k // 2 <= 7:
        terms.append((f"a.l[{k // 2}]", f"a.l[{k // 2}]", "v"))
    return terms


def column_terms(k, square=False):
    """The digit products summed in column k before its quotient digit (for
    k < 8 the asm then adds m_k * p_0): a*b (or a*a) and m_j * p_{k-j}."""
    if square:
        terms = square_terms(k)
    else:
        terms = [(f"a.l[{i}]", f"b.l[{k - i}]", "v") for i in range(8) if 0 <= k - i < 8]
    terms += [(f"m[{j}]", f"p_limb({k - j})", "s") for j in range(8) if j < k and 1 <= k - j < 8]
    return terms


def no_carry_terms(k, square=False):
    """The terms of column k that are summed first and without a carry add.
    Column k >= 7 opens with the accumulator below 17 * 2^32 (the previous
    column's high word plus at most 16 carries).  Product: the two terms with
    a top limb, a_{k-7} b_7 and a_7 b_{k-7}; their sum stays below 2^64 - 17 *
    2^32 when top(a) + top(b) <= 2^32 - 18 (lazy times lazy: 0.76 * 2^32;
    canonical times a value below 4p: 0.95 * 2^32).  Squaring (a < 2p, so
    the doubled top limb d_7 or u_7 is below 0.76 * 2^32): a_{k-7} d_7 (u_7 in
    column 13) and m_{k-7} p_7 (p_7 < 0.19 * 2^32).  Column 14 holds only
    a_7 b_7 (or a_7^2) and m_7 p_7, both small."""
    if k < 7:
        return []
    if square:
        nc = [t for t in column_terms(k, True)
              if t[1] in ("d[7]", "u[7]") or (k == 14 and t[0] == "a.l[7]")]
    else:
        nc = [t for t in column_terms(k, False)
              if t[0] == "a.l[7]" or t[1] == "b.l[7]"]
    nc += [t for t in column_terms(k, square) if t[0] == f"m[{k - 7}]" and t[1] == "p_limb(7)"
           and (square or k == 14)]
    return nc


# gfx950 (like gfx942) needs two wait states between a VALU instruction that
# writes an SGPR and a VALU instruction that reads it, the carry-in of an
# add-with-carry included: hipcc's own code puts `s_nop 1` between the links
# of a v_add_co / v_addc_co chain for these targets (none for gfx90a).  The
# carry of a multiply-add is therefore added three instructions later: the
# carrying terms' carries rotate through three SGPR pairs, so multiply-add i
# writes pair i % 3 and its carry add follows the multiply-add of term i + 2;
# an s_nop fills a gap that has fewer than two instructions.
WAIT_STATES = 2
CARRY_REGS = 3


def column_asm(nc, carrying):
    """(instructions, operand count used by carry pairs) for one column:
    %0 is c01, %1 c2, %2.. the carry pairs, then the term operands."""
    ncar = len(carrying)
    nreg = CARRY_REGS if ncar else 1
    base = (2 if ncar else 1) + nreg  # first term operand
    body = []
    n = base
    mad_at = {}  # carrying index -> position of its multiply-add in body
    cy = lambda i: f"%{(2 if ncar else 1) + i % nreg}"

    def add_carry(i):
        gap = len(body) - mad_at[i] - 1
        if gap < WAIT_STATES:
            body.append(f"s_nop {WAIT_STATES - gap - 1}")
        if i == 0:
            # the column's first carry starts c2 (0 + 0 + carry): no zeroing move
            body.append(f"v_addc_co_u32_e64 %1, {cy(i)}, 0, 0, {cy(i)}")
        else:
            body.append(f"v_addc_co_u32_e64 %1, {cy(i)}, 0, %1, {cy(i)}")

    for t in range(len(nc)):
        # summed without a carry add: the carry-out is never set (see
        # no_carry_terms) and never read
        body.append(f"v_mad_u64_u32 %0, {cy(0)}, %{n}, %{n + 1}, %0")
        n += 2
    for i in range(ncar):
        mad_at[i] = len(body)
        body.append(f"v_mad_u64_u32 %0, {cy(i)}, %{n}, %{n + 1}, %0")
        n += 2
        if i >= 2:
            add_carry(i - 2)
    for i in range(max(0, ncar - 2), ncar):
        add_carry(i)
    return body, nreg


def column(k, square=False):
    nc = no_carry_terms(k, square)
    carrying = [t for t in column_terms(k, square) if t not in nc]
    body, nreg = column_asm(nc, carrying)
    ops = []
    for (x, y, cls) in nc + carrying:
        ops += [f'"v"({x})', f'"{cls}"({y})']
    if carrying:
        outs = '"+v"(c01), "=&v"(c2), ' + ", ".join(f'"=&s"(cy[{i}])' for i in range(nreg))
    else:
        outs = '"+v"(c01), "=&s"(cy[0])'
    out = []
    if body:
        text = "\\n\\t".join(body)
        out.append(f'  asm("{text}"\n               : {outs}\n'
                   f'               : {", ".join(ops)});')
    if k < 8:
        # m_k = the low word of lo * 0xefffffff (its carry-out is never set:
        # the 64-bit product plus 0 cannot pass 2^64), then m_k * p0 into the
        # column; the carry add waits two states for the multiply-add's carry
        out.append(f'  asm("v_mad_u64_u32 %0, %1, %2, %3, 0"\n'
                   f'               : "=v"(mq), "=s"(cy[0])\n'
                   f'               : "v"((uint32_t)c01), "s"(0xefffffffu));')
        out.append(f'  m[{k}] = (uint32_t)mq;')
        out.append(f'  asm("v_mad_u64_u32 %0, %2, %3, %4, %0\\n\\ts_nop 1\\n\\t'
                   f'v_addc_co_u32_e64 %1, %2, 0, %1, %2"\n'
                   f'               : "+v"(c01), "+v"(c2), "=&s"(cy[0])\n'
                   f'               : "v"(m[{k}]), "s"(p_limb(0)));')
    else:
        out.append(f"  r.l[{k - 8}] = (uint32_t)c01;")
    out.append("  c01 = (c01 >> 32) | ((uint64_t)c2 << 32);" if carrying else "  c01 >>= 32;")
    return out


def write(name, head, square):
    lines = [f"// Generated by gen_fr_mul.py; do not edit.  {head}"]
    for k in range(15):
        lines.append(f"  // column {k}")
        lines += column(k, square)
    lines.append("  r.l[7] = (uint32_t)c01;")
    with open(os.path.join(HERE, name), "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    write("fr_mul_asm.inc", "The body of fr_mul (fr.hpp):\n"
          "// in a, b (lazy, or one canonical and one < 4p: a_7 + b_7 <= 2^32 - 18);\n"
          "// out r = a * b * 2^-256 (lazy); m[8], mq, c01, c2, cy[3] declared there.",
          False)
    write("fr_sqr_asm.inc", "The body of fr_sqr (fr.hpp):\n"
          "// in a (< 2p) and its doubled limbs u, d; out r = a * a * 2^-256 (lazy);\n"
          "// m[8], mq, c01, c2, cy[3] declared there.", True)


if __name__ == "__main__":
    main()
