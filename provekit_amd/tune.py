"""Autotune the NTT pass shapes on the current GPU and write the table the
library reads at start-up (ntt_tuning.json beside the library, or the file
named by PKGPU_NTT_TUNING).

    python -m provekit_amd.tune --out provekit_amd/ntt_tuning.json
"""
import argparse
import json
import os

import provekit_amd as pk


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "ntt_tuning.json"))
    ap.add_argument("--log-min", type=int, default=6)
    ap.add_argument("--log-max", type=int, default=26)
    ap.add_argument("--max-skip", type=int, default=3, help="largest log2 blowup (stages skipped)")
    args = ap.parse_args()
    pk.init(0)
    secs = pk.ntt_autotune(args.log_min, args.log_max, args.max_skip, args.out)
    with open(args.out) as f:
        n = len(json.load(f)["entries"])
    print(json.dumps({"tuning_file": args.out, "entries": n, "seconds": secs}), flush=True)


if __name__ == "__main__":
    main()
