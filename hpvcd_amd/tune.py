"""Autotune the kernels' launch shapes on this GPU and store them in a tuning
cache file that hpvcd_gpu_open reads (environment variable HPVCD_TUNING_FILE).

    python -m hpvcd_amd.tune --workload grid4032 --file hpvcd_tuning.txt

The search runs on a resident job of the named workload (the generator's
records); every shape gives the same samples, so the file only changes speed.
"""
import argparse
import json
import sys

import hpvcd_amd as H


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="grid4032", choices=sorted(H.WORKLOADS))
    ap.add_argument("--file", required=True, help="tuning cache file to write (the GPU's key line is replaced)")
    ap.add_argument("--iters", type=int, default=10, help="timed runs per candidate")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args()
    rs = H.RecordSet.generate(args.workload)
    with H.Gpu(args.device) as g:
        g.tuning = H.default_tuning()
        with g.job(rs) as job:
            best, best_ms, default_ms = job.tune(args.iters)
        key = g.tuning_key
    H.save_tuning(args.file, key, best)
    print(json.dumps({"key": key, "workload": args.workload, "best": best, "best_ms": round(best_ms, 4),
                      "default_ms": round(default_ms, 4)}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
