#!/usr/bin/env python3
"""Benchmark: WHIR-style commit of one polynomial on one MI355X.

Workload (one step): 2^20 uniform Fr coefficients (seeded), zero-padded to
blowup 4 and evaluated by a coset NTT (coset 5*H) into a 2^22 codeword in
RSFr order, leaves of 16 consecutive codeword elements hashed by Skyscraper
left fold, binary Skyscraper tree to the root.  Inputs are resident in HBM
before the timed region; each step writes the codeword and the whole tree.

Prints one JSON line (rank 0).  `value` is committed input elements per
second over the timed steps.  Per-stage times come from HIP events recorded
on the stream the kernels run on.  The CPU baseline times the oracle (the C
restatement of the reference's Rayon path, OpenMP over all host threads) on
the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fr elements/s committed (RS-NTT+Merkle), batched vs plain; p50/p99 commit ms"

# MI355X peaks used for the roofline fractions
HBM_PEAK_GBS = 8000.0
# VALU: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz; 32-bit integer ops have no
# packed form, so this is the integer issue peak (lane-instructions/s).
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
# A Montgomery product of 8x32-bit digits cannot issue fewer than 64 + 64
# digit multiply-adds (a*b and m*p) plus as many carry adds: 256 VALU
# instructions, so the multiplier's speed of light is VALU peak / 256.
MODMUL_PEAK_G = VALU_PEAK_TOPS * 1e3 / 256


class Hip:
    """Minimal ctypes binding of the HIP runtime for streams and events."""

    def __init__(self):
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.L = ctypes.CDLL(name)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("libamdhip64.so not found")
        vp = ctypes.c_void_p
        self.L.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
        self.L.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        self.L.hipEventRecord.argtypes = [vp, vp]
        self.L.hipEventSynchronize.argtypes = [vp]
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        self.L.hipStreamSynchronize.argtypes = [vp]
        self.L.hipEventDestroy.argtypes = [vp]

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with hipError {rc}")

    def stream(self):
        s = ctypes.c_void_p()
        self._ok(self.L.hipStreamCreateWithFlags(ctypes.byref(s), 1), "hipStreamCreateWithFlags")
        return s.value

    def event(self):
        e = ctypes.c_void_p()
        self._ok(self.L.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e.value

    def record(self, e, s):
        self._ok(self.L.hipEventRecord(e, s), "hipEventRecord")

    def elapsed_ms(self, a, b):
        f = ctypes.c_float()
        self._ok(self.L.hipEventElapsedTime(ctypes.byref(f), a, b), "hipEventElapsedTime")
        return f.value

    def sync(self, s):
        self._ok(self.L.hipStreamSynchronize(s), "hipStreamSynchronize")


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    lo = int(k)
    hi = min(lo + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def load_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def cpu_baseline(coeffs, blowup, leaf_elems, budget_s):
    from oracle import pyoracle

    pyoracle.lib()
    n = coeffs.shape[0]
    t0 = time.perf_counter()
    root, _, _ = pyoracle.commit(coeffs, blowup, True, leaf_elems)  # warm-up
    first = time.perf_counter() - t0
    runs = max(2, min(5, int(budget_s / max(first, 1e-3))))
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        pyoracle.commit(coeffs, blowup, True, leaf_elems)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {
        "value": n / med,
        "unit": "Fr elements/s",
        "cores": pyoracle.threads(),
        "kind": "port",
        "sample": f"full workload (2^{n.bit_length() - 1} coefficients) x {runs} runs after 1 warm-up, median "
                  f"{med * 1e3:.1f} ms; C restatement of the reference Rayon path, OpenMP",
        "ms_per_commit": med * 1e3,
    }, root


def run_batched(pk, hip, s, seed, steps, warmup, leaf_elems=16):
    """Config 3: 64 polynomials, k ~ U{12..22}, blowup 2 or 4 by coin flip, coset
    NTT, one tree each; batched (one persistent hashing launch) vs 64 plain
    per-call commits on the same GPU, everything resident in HBM."""
    from provekit_amd.inputs import random_fr

    rng = np.random.default_rng(seed)
    ks = rng.integers(12, 23, size=64)
    blowups = rng.choice([2, 4], size=64)
    # coefficients: seeded rotations of one 2^22-element ChaCha20 pool
    pool = random_fr(1 << 22, seed=seed, stream="batch-pool")
    offs = rng.integers(0, 1 << 22, size=64)
    polys = []
    for k, b, off in zip(ks, blowups, offs):
        c = np.roll(pool, -int(off), axis=0)[: 1 << int(k)]
        polys.append((pk.DeviceBuffer.from_array(c), 1 << int(k), int(b), True))
    del pool
    batch = pk.BatchCommit(polys, leaf_elems=leaf_elems)
    total = int(sum(1 << int(k) for k in ks))

    def plain_all():
        for i, (c, n, b, _) in enumerate(polys):
            pk._call("pkgpu_commit_device", c.ptr, n, b, 1, leaf_elems, batch.codewords[i].ptr,
                     batch.nodes[i].ptr, s)

    def timed(fn, k_steps, per_call_sync=False):
        e0, e1 = hip.event(), hip.event()
        hip.sync(s)
        t0 = time.perf_counter()
        hip.record(e0, s)
        for _ in range(k_steps):
            fn()
        hip.record(e1, s)
        hip.sync(s)
        return hip.elapsed_ms(e0, e1) / k_steps, (time.perf_counter() - t0) * 1e3 / k_steps

    for _ in range(warmup):
        plain_all()
        batch.run(s)
        batch.graph_launch(s)
    hip.sync(s)
    plain_ms, _ = timed(plain_all, steps)
    # host-synchronous per-call use: wait for each commit before the next
    def plain_synced():
        for i, (c, n, b, _) in enumerate(polys):
            pk._call("pkgpu_commit_device", c.ptr, n, b, 1, leaf_elems, batch.codewords[i].ptr,
                     batch.nodes[i].ptr, s)
            hip.sync(s)
    _, plain_sync_wall = timed(plain_synced, max(1, steps // 2))
    batch_ms, _ = timed(lambda: batch.run(s), steps)
    graph_ms, _ = timed(lambda: batch.graph_launch(s), steps)
    plain_roots = []
    plain_all()
    hip.sync(s)
    for i in range(64):
        N = polys[i][1] * polys[i][2]
        nodes = 2 * (N // leaf_elems) - 1
        plain_roots.append(batch.nodes[i].download(np.uint8, nbytes=32 * nodes)[-32:].tobytes())
    batch.run(s)
    roots = batch.roots(s)
    best = min(batch_ms, graph_ms)
    return {
        "workload": "64 polys, k~U{12..22}, blowup 2/4 coin flip, coset RS-NTT, 16-element leaves, "
                    "Skyscraper tree each; coefficients = seeded rotations of a 2^22 ChaCha20 pool",
        "total_coefficients": total,
        "ks": [int(k) for k in ks], "blowups": [int(b) for b in blowups],
        "plain_ms": plain_ms, "plain_elements_per_s": total / (plain_ms / 1e3),
        "plain_synced_ms": plain_sync_wall, "plain_synced_elements_per_s": total / (plain_sync_wall / 1e3),
        "batched_ms": batch_ms, "batched_graph_ms": graph_ms,
        "batched_elements_per_s": total / (best / 1e3),
        "speedup_vs_plain": plain_ms / best,
        "speedup_vs_plain_synced": plain_sync_wall / best,
        "roots_match_plain": roots == plain_roots,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--blowup", type=int, default=4)
    ap.add_argument("--leaf", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--batched-steps", type=int, default=5)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run).", file=sys.stderr)
        sys.exit(2)

    import provekit_amd as pk
    from provekit_amd.inputs import random_fr

    pk.init(0)
    hip = Hip()
    s = hip.stream()

    n = 1 << args.log_n
    N = n * args.blowup
    leaves = N // args.leaf
    nodes = 2 * leaves - 1
    coeffs = random_fr(n, seed=args.seed, stream="commit-bench")
    d_coeffs = pk.DeviceBuffer.from_array(coeffs)
    d_cw = pk.DeviceBuffer(32 * N)
    d_nodes = pk.DeviceBuffer(32 * nodes)

    def step(evs=None):
        if evs:
            hip.record(evs[0], s)
        pk._call("pkgpu_rs_interleaved_encode_device", d_coeffs.ptr, 1, n, None, 0, N, 1, d_cw.ptr, s)
        if evs:
            hip.record(evs[1], s)
        pk._call("pkgpu_skyscraper_hash_many_device", 32 * args.leaf, d_cw.ptr, leaves, d_nodes.ptr, s)
        if evs:
            hip.record(evs[2], s)
        pk._call("pkgpu_skyscraper_merkle_device", d_nodes.ptr, leaves, d_nodes.ptr, s)
        if evs:
            hip.record(evs[3], s)

    for _ in range(args.warmup):
        step()
    hip.sync(s)
    pk.synchronize()

    K = args.steps
    evs = [[hip.event() for _ in range(4)] for _ in range(K)]
    t0 = time.perf_counter()
    for k in range(K):
        step(evs[k])
    hip.sync(s)
    pk.synchronize()
    wall = time.perf_counter() - t0

    lat = [hip.elapsed_ms(e[0], e[3]) for e in evs]
    ntt = [hip.elapsed_ms(e[0], e[1]) for e in evs]
    leaf = [hip.elapsed_ms(e[1], e[2]) for e in evs]
    tree = [hip.elapsed_ms(e[2], e[3]) for e in evs]
    gpu_total_ms = hip.elapsed_ms(evs[0][0], evs[-1][3])
    ms_per_step = gpu_total_ms / K
    value = n * K / (gpu_total_ms / 1e3)

    root = d_nodes.download(np.uint8, nbytes=32 * nodes)[-32:].tobytes()

    compressions_leaf = leaves * (args.leaf - 1)
    leaf_ms = statistics.mean(leaf)
    ntt_ms = statistics.mean(ntt)
    tree_ms = statistics.mean(tree)
    # dominant kernel: Skyscraper leaf hashing (one launch, k_hash_fold)
    modmul_leaf = 14 * compressions_leaf
    achieved = modmul_leaf / (leaf_ms / 1e3) / 1e9
    roofline = {
        "bound": "valu",
        "kernel": "k_hash_fold (Skyscraper leaf hashing)",
        "achieved": achieved,
        "peak": MODMUL_PEAK_G,
        "unit": "Gmodmul/s",
        "frac": achieved / MODMUL_PEAK_G,
        "traffic": load_traffic("k_hash_fold"),
        "work_per_launch": f"{compressions_leaf} compressions x 14 Montgomery squarings",
        "peak_basis": f"VALU int32 issue {VALU_PEAK_TOPS:.1f} T lane-instr/s / 256 instr per 8x32 Montgomery product",
    }
    ntt_bytes = 32 * n + 32 * N  # read the coefficient chunk once, write the codeword once
    ntt_modmul = (N // 2) * (N.bit_length() - 1 - (args.blowup.bit_length() - 1))
    roofline_ntt = {
        "bound": "hbm",
        "kernel": "k_ntt_pass (RS encode, all passes)",
        "achieved": ntt_bytes / (ntt_ms / 1e3) / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": ntt_bytes / (ntt_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "traffic": load_traffic("k_ntt_pass"),
        "modmul_rate_G": ntt_modmul / (ntt_ms / 1e3) / 1e9,
        "modmul_frac": ntt_modmul / (ntt_ms / 1e3) / 1e9 / MODMUL_PEAK_G,
    }

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Fr elements/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "plain commit: 2^20 Fr coefficients -> coset RS-NTT blowup 4 (2^22) -> "
                        "16-element leaves -> Skyscraper binary Merkle tree",
            "coefficients": n, "blowup": args.blowup, "codeword": N, "leaf_elems": args.leaf,
            "hash": "skyscraper", "seed": args.seed, "parallelism": "single GPU",
        },
        "p50_commit_ms": pct(lat, 0.5),
        "p99_commit_ms": pct(lat, 0.99),
        "stages_ms": {"rs_ntt": ntt_ms, "leaf_hash": leaf_ms, "merkle_tree": tree_ms},
        "codeword_elements_per_s": N * K / (gpu_total_ms / 1e3),
        "skyscraper_hashes_per_s": (compressions_leaf + leaves - 1) * K / (gpu_total_ms / 1e3),
        "wall_s": wall,
        "root": root.hex(),
        "roofline": roofline,
        "roofline_ntt": roofline_ntt,
    }
    if not args.no_batched:
        out["batched"] = run_batched(pk, hip, s, args.seed, args.batched_steps, 1, args.leaf)
    if not args.no_cpu:
        cb, croot = cpu_baseline(coeffs, args.blowup, args.leaf, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["root_matches_cpu"] = croot == root
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
