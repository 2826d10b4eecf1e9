"""NTT schedule tuning: the tuning file format and lookup (host only), every
pass shape bit-exact against the oracle (each shape forced through a tuning
table), and an autotune run whose chosen shapes keep the results exact."""
import json
import os

import numpy as np
import pytest

from provekit_amd.inputs import random_fr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_TABLE = os.path.join(ROOT, "provekit_amd", "ntt_tuning.json")

# (tile_log, butterflies per thread) pairs with a pass kernel
SHAPES = [(tl, b) for tl in range(6, 12) for b in (1, 2, 4)]


def write_table(path, entries):
    with open(path, "w") as f:
        json.dump({"device": "test", "entries": entries}, f)


def restore_default(pk, tmp_path):
    if os.path.exists(DEFAULT_TABLE):
        pk.ntt_tuning_load(DEFAULT_TABLE)
    else:
        empty = str(tmp_path / "empty.json")
        write_table(empty, [])
        pk.ntt_tuning_load(empty)


def test_tuning_file_lookup(tmp_path):
    import provekit_amd as pk
    path = str(tmp_path / "t.json")
    write_table(path, [
        {"log": 20, "stages": 18, "tile_log": 10, "bpt": 2, "min_run": 1, "us": 5.0},
        {"log": 14, "stages": 14, "tile_log": 8, "bpt": 1, "min_run": 0, "us": 1.0},
        {"log": 16, "stages": 16, "tile_log": 12, "bpt": 4, "min_run": 0},  # unsupported: dropped
        {"log": 22, "stages": 20, "tile_log": 9, "bpt": 3, "min_run": 0},   # unsupported: dropped
    ])
    try:
        pk.ntt_tuning_load(path)
        assert pk.ntt_tuning_get(20, 18) == {"tile_log": 10, "bpt": 2, "min_run": 1}
        assert pk.ntt_tuning_get(14, 14) == {"tile_log": 8, "bpt": 1, "min_run": 0}
        # no entry: nearest length with the same number of skipped stages
        assert pk.ntt_tuning_get(24, 22) == {"tile_log": 10, "bpt": 2, "min_run": 1}
        assert pk.ntt_tuning_get(16, 16) == {"tile_log": 8, "bpt": 1, "min_run": 0}
        # no entry with that skip: nearest length overall
        assert pk.ntt_tuning_get(15, 12) == {"tile_log": 8, "bpt": 1, "min_run": 0}
        with pytest.raises(pk.ProvekitGpuError):
            pk.ntt_tuning_load(str(tmp_path / "missing.json"))
        bad = str(tmp_path / "bad.json")
        with open(bad, "w") as f:
            f.write("{}")
        with pytest.raises(pk.ProvekitGpuError):
            pk.ntt_tuning_load(bad)
        empty = str(tmp_path / "empty.json")
        write_table(empty, [])
        pk.ntt_tuning_load(empty)
        assert pk.ntt_tuning_get(20, 18) == {"tile_log": 11, "bpt": 4, "min_run": 2}
    finally:
        restore_default(pk, tmp_path)


def test_committed_table_is_valid():
    if not os.path.exists(DEFAULT_TABLE):
        pytest.skip("no tuning table committed")
    with open(DEFAULT_TABLE) as f:
        t = json.load(f)
    assert t["entries"]
    for e in t["entries"]:
        assert (e["tile_log"], e["bpt"]) in SHAPES or e["tile_log"] < 6
        assert 0 <= e["min_run"] < e["tile_log"]
        assert 1 <= e["stages"] <= e["log"]


@pytest.mark.gpu
@pytest.mark.parametrize("L,skip,M", [(7, 0, 1), (12, 0, 1), (12, 2, 1), (17, 1, 1), (10, 0, 3), (20, 2, 1)])
def test_every_shape_matches_oracle(gpu, oracle, tmp_path, L, skip, M):
    v = random_fr(M << L, seed=L * 10 + skip, stream="shapes")
    expect = v.copy()
    oracle.ntt_nr(expect, 1 << L, 1 << skip)
    path = str(tmp_path / "force.json")
    try:
        for tl, bpt in SHAPES:
            for mr in (0, 2, 3):
                if mr >= tl:
                    continue
                write_table(path, [{"log": L, "stages": L - skip, "tile_log": tl, "bpt": bpt,
                                    "min_run": mr}])
                gpu.ntt_tuning_load(path)
                print(f"[case] L={L} skip={skip} M={M} tile_log={tl} bpt={bpt} min_run={mr}", flush=True)
                a = v.copy()
                gpu.ntt_nr(a, 1 << L, 1 << skip)
                assert np.array_equal(a, expect), (tl, bpt, mr)
    finally:
        restore_default(gpu, tmp_path)


@pytest.mark.gpu
def test_autotune_keeps_results_exact(gpu, oracle, tmp_path):
    path = str(tmp_path / "tuned.json")
    try:
        secs = gpu.ntt_autotune(6, 13, 2, path)
        print(f"[case] autotune 2^6..2^13 took {secs:.2f} s", flush=True)
        with open(path) as f:
            t = json.load(f)
        keys = {(e["log"], e["stages"]) for e in t["entries"]}
        for L in range(6, 14):
            for skip in range(3):
                assert (L, L - skip) in keys
        for L, skip in [(9, 1), (13, 2), (12, 0)]:
            v = random_fr(1 << L, seed=L, stream="tuned")
            a, b = v.copy(), v.copy()
            gpu.ntt_nr(a, 1 << L, 1 << skip)
            oracle.ntt_nr(b, 1 << L, 1 << skip)
            assert np.array_equal(a, b)
    finally:
        restore_default(gpu, tmp_path)
