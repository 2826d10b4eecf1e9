"""Extract the known-answer vectors that the reference's own tests hold into
reference_kats.json (data only: inputs and expected outputs).

Run here (where /root/reference exists):  python tests/golden/make_reference_kats.py
Sources:
  skyscraper/core/src/reference.rs:104-188   Sage KATs (ss, bb, permute) + S-box Table 3
  poseidon2/src/permutation.rs:183-194        permutation([0;4])
  poseidon2/src/hash.rs:61-78                  hash([1]), hash([e, e])
  poseidon2/src/constants.rs:969-1019          parameter samples
  provekit/common/src/utils/sumcheck.rs:241-435  calculate_eq, eq table, 3x4 R1CS
  provekit/common/proptest-regressions/ntt.txt   regression inputs
"""
import json
import os
import re

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")


def read(p):
    with open(os.path.join(REF, p)) as f:
        return f.read()


def fn_body(src, name):
    i = src.index("fn " + name)
    j = src.index("\n    }\n", i)
    return src[i:j]


def decimals(s):
    return re.findall(r'"(\d{10,})"', s)


def main():
    kats = {}
    sky = read("skyscraper/core/src/reference.rs")
    for name, rnd in (("test_ss_2", 2), ("test_bb_6", 6)):
        l, r, el, er = decimals(fn_body(sky, name))
        kats[name] = {"round": rnd, "l": l, "r": r, "out_l": el, "out_r": er}
    el, er = decimals(fn_body(sky, "test_zero"))
    kats["test_zero"] = {"l": "0", "r": "0", "out_l": el, "out_r": er}
    l, r, el, er = decimals(fn_body(sky, "test_random"))
    kats["test_random"] = {"l": l, "r": r, "out_l": el, "out_r": er}
    kats["sbox_table3"] = [[int(a, 16), int(b, 16)] for a, b in
                           re.findall(r"sbox\(0x([0-9a-f]{2})\), 0x([0-9a-f]{2})", fn_body(sky, "test_sbox"))]
    kats["sigma_inv"] = decimals(sky[sky.index("SIGMA_INV"):sky.index("SIGMA_INV") + 400])[0]

    perm = read("poseidon2/src/permutation.rs")
    kats["p2_perm_zero"] = re.findall(r'fe\("([0-9A-Fa-f]{64})"\)', fn_body(perm, "smoke_test"))
    h = read("poseidon2/src/hash.rs")
    kats["p2_hash_one"] = re.findall(r'fe\("0x([0-9a-f]{64})"\)', fn_body(h, "test_poseidon2_hash_single_element"))[0]
    e, out = re.findall(r'fe\("0x([0-9a-f]{64})"\)', fn_body(h, "test_poseidon2_hash_two_elements"))
    kats["p2_hash_ee"] = {"e": e, "out": out}
    c = read("poseidon2/src/constants.rs")
    body = fn_body(c, "t4_round_constants_sample_match_reference")
    samples = re.findall(r'(rc1\[0\]\[0\]|rc1\[1\]\[3\]|rc2\[3\]\[2\]|rcp\[0\]|rcp\[rcp\.len\(\) - 1\]),\s*fe\("0x([0-9a-f]{64})"\)', body)
    kats["p2_constant_samples"] = {k: v for k, v in samples}
    kats["p2_diag"] = re.findall(r'fe\("0x([0-9a-f]{64})"\)', fn_body(c, "t4_internal_mds_diag_matches_reference"))

    sc = read("provekit/common/src/utils/sumcheck.rs")
    kats["calculate_eq"] = {"r": [2, 3, 4, 5], "alpha": [6, 7, 8, 9], "out": 2289441}
    assert "2_289_441" in sc
    kats["eval_eq_left_only"] = {"r": [2, 3, 5], "out": -8}
    kats["eq_single_entry"] = {"r": [2, 3], "out": 2}
    # 3x4 R1CS (sumcheck.rs:241-256) and multiply_transposed_by_eq_alpha (:417-435)
    kats["r1cs_3x4"] = {
        "A": [[[1, 0], [2, 1]], [[3, 2]], [[1, 0], [1, 3]]],
        "B": [[[1, 1]], [[2, 0], [1, 3]], [[4, 2]]],
        "C": [[[1, 2]], [[1, 1], [3, 3]], [[2, 0]]],
        "alpha": [2, 3],
        "expected_a": [-2, 4, -9, -4],
        "expected_b": [-6, 2, -16, -3],
        "expected_c": [-8, -3, 2, -9],
    }
    body = fn_body(sc, "test_multiply_transposed_by_eq_alpha")
    for key, name in (("expected_a", "expected_a"), ("expected_b", "expected_b"), ("expected_c", "expected_c")):
        got = [int(v) for v in re.findall(r"fe\((-?\d+)\)", body[body.index(name):body.index("];", body.index(name))])]
        assert got == kats["r1cs_3x4"][key], (key, got)
    reg = read("provekit/common/proptest-regressions/ntt.txt")
    kats["ntt_regressions_raw"] = [ln for ln in reg.splitlines() if ln.startswith("cc ")]
    with open(OUT, "w") as f:
        json.dump(kats, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
