"""Summaries of rocprofv3 runs of ``bench.py`` for the ``profiles/`` folder.

Three runs of the same bench command feed one summary:

* ``--kernel-trace --stats``: per-kernel dispatch counts and durations;
* ``--pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE`` (separate runs, the counter
  blocks cannot hold both with the trace): HBM bytes read and written per
  dispatch, in KiB as the counters report them.

The PMC runs pass ``--copy-calibration MIB`` to ``bench.py``, which adds
device-to-device copies of a known size through the HIP runtime's copy kernel
(MIB MiB read, MIB MiB written).  Each counter is scaled by the factor that maps its reading on that
copy onto the known byte count.

    python -m hpvcd_amd.profiling --workload grid4032 --copy-mib 256 \\
        --trace gpurun_out/prof_trace --fetch gpurun_out/pmc_fetch \\
        --write gpurun_out/pmc_write --command "<the bench command>" --out profiles

writes ``profiles/kernel_stats_<workload>.md`` and
``profiles/traffic_<workload>.{json,md}``; ``bench.py`` reads the JSON for
``roofline.traffic``.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import statistics

# kernel kinds as bench.py names them -> substring of the HIP kernel symbol
KERNEL_SYMBOLS = {
    "intra": "k_intra<",
    "prep": "k_prep(",
    "residual": "k_residual<",
    "deblock": "k_deblock<",
    "sao": "k_sao<",
    "rgb": "k_rgb<",
    "mc": "k_mc<",
    "act": "k_act<",
}


def _one(pattern: str) -> str:
    """The newest file matching pattern (a folder reused across runs keeps
    older outputs beside the new one)."""
    hits = glob.glob(pattern, recursive=True)
    if not hits:
        raise FileNotFoundError(pattern)
    return max(hits, key=os.path.getmtime)


def kernel_stats(trace_dir: str) -> list[dict]:
    """Rows of the rocprofv3 kernel statistics CSV, sorted by total time."""
    path = _one(os.path.join(trace_dir, "**", "*kernel_stats.csv"))
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append({
                "name": r["Name"],
                "calls": int(r["Calls"]),
                "total_ns": float(r["TotalDurationNs"]),
                "avg_ns": float(r["AverageNs"]),
                "min_ns": float(r["MinNs"]),
                "max_ns": float(r["MaxNs"]),
                "percent": float(r["Percentage"]),
            })
    rows.sort(key=lambda r: -r["total_ns"])
    return rows


def counter_per_dispatch(pmc_dir: str, counter: str) -> list[tuple[str, float]]:
    """(kernel name, counter value) per dispatch, summed over counter instances."""
    path = _one(os.path.join(pmc_dir, "**", "*counter_collection.csv"))
    acc: dict[str, list] = collections.OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            key = r["Dispatch_Id"]
            if key not in acc:
                acc[key] = [r["Kernel_Name"], 0.0]
            acc[key][1] += float(r["Counter_Value"])
    return [(k, v) for k, v in acc.values()]


def calibration(values: list[tuple[str, float]], copy_mib: int) -> dict:
    """Factor mapping the counter's KiB reading to bytes, from the copy kernel."""
    copies = [v for name, v in values if "copyBuffer" in name and v > 0.5 * (copy_mib << 10)]
    if not copies:
        raise ValueError("no copy dispatch of the calibration size in the counter run")
    reading_kib = statistics.median(copies)
    return {"copy_bytes": copy_mib << 20, "reading_kib": reading_kib,
            "bytes_per_kib": (copy_mib << 20) / reading_kib}


def write_kernel_stats(workload: str, stats: list[dict], command: str, out_dir: str) -> None:
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, f"kernel_stats_{workload}.md"), "w") as f:
        f.write(f"# Kernel durations, {workload}\n\nCommand:\n\n    {command}\n\n")
        f.write("| kernel | calls | avg us | min us | max us | % of GPU time |\n|---|---|---|---|---|---|\n")
        for r in stats[:16]:
            name = r["name"].split("(")[0].replace("void ", "")
            f.write(f"| `{name}` | {r['calls']} | {r['avg_ns'] / 1e3:.2f} | {r['min_ns'] / 1e3:.2f} | "
                    f"{r['max_ns'] / 1e3:.2f} | {r['percent']:.1f} |\n")


def summarize(workload: str, copy_mib: int, trace_dir: str, fetch_dir: str, write_dir: str, command: str,
              out_dir: str) -> dict:
    stats = kernel_stats(trace_dir)
    fetch = counter_per_dispatch(fetch_dir, "FETCH_SIZE")
    write = counter_per_dispatch(write_dir, "WRITE_SIZE")
    cf, cw = calibration(fetch, copy_mib), calibration(write, copy_mib)
    kernels = {}
    for kind, sym in KERNEL_SYMBOLS.items():
        fv = [v for name, v in fetch if sym in name]
        wv = [v for name, v in write if sym in name]
        tr = [r for r in stats if sym in r["name"]]
        if not fv or not wv:
            continue
        calls = sum(r["calls"] for r in tr)
        # launches per step: kernels of one kind may have several instantiations
        # or launches (deblock: 4 per step); average over dispatches
        f_b = statistics.mean(fv) * cf["bytes_per_kib"]
        w_b = statistics.mean(wv) * cw["bytes_per_kib"]
        kernels[kind] = {
            "dispatches_fetch_run": len(fv),
            "fetch_bytes_per_launch": int(f_b),
            "write_bytes_per_launch": int(w_b),
            "hbm_bytes_per_launch": int(f_b + w_b),
            "trace_calls": calls,
            "trace_avg_us": (sum(r["total_ns"] for r in tr) / calls / 1e3) if calls else None,
        }
    summary = {
        "workload": workload,
        "command_trace": command.format(run="--kernel-trace --stats"),
        "command_fetch": command.format(run="--pmc FETCH_SIZE"),
        "command_write": command.format(run="--pmc WRITE_SIZE"),
        "calibration": {"fetch": cf, "write": cw},
        "kernels": kernels,
    }
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, f"traffic_{workload}.json"), "w") as f:
        json.dump(summary, f, indent=1)
    with open(os.path.join(out_dir, f"traffic_{workload}.md"), "w") as f:
        f.write(f"# HBM traffic per launch, {workload}\n\n")
        f.write("Commands (rocprofv3 on one MI355X; PMC runs separate from the trace):\n\n")
        for key in ("command_fetch", "command_write"):
            f.write(f"    {summary[key]}\n")
        f.write("\nCalibration: the HIP runtime's device-to-device copy kernel, "
                f"{copy_mib} MiB read + {copy_mib} MiB written.\n\n")
        f.write("| counter | copy reading (KiB) | bytes per KiB |\n|---|---|---|\n")
        f.write(f"| FETCH_SIZE | {cf['reading_kib']:.0f} | {cf['bytes_per_kib']:.4f} |\n")
        f.write(f"| WRITE_SIZE | {cw['reading_kib']:.0f} | {cw['bytes_per_kib']:.4f} |\n\n")
        f.write("| kernel | fetch MB/launch | write MB/launch | HBM MB/launch |\n|---|---|---|---|\n")
        for kind, k in kernels.items():
            f.write(f"| {kind} | {k['fetch_bytes_per_launch'] / 1e6:.2f} | {k['write_bytes_per_launch'] / 1e6:.2f} | "
                    f"{k['hbm_bytes_per_launch'] / 1e6:.2f} |\n")
    write_kernel_stats(workload, stats, summary["command_trace"], out_dir)
    return summary


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", required=True)
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", help="FETCH_SIZE run (omit with --write for a durations-only summary)")
    ap.add_argument("--write", help="WRITE_SIZE run")
    ap.add_argument("--command", required=True,
                    help="the profiled command with {run} in place of the rocprofv3 mode options")
    ap.add_argument("--out", default="profiles")
    a = ap.parse_args()
    if not a.fetch or not a.write:
        write_kernel_stats(a.workload, kernel_stats(a.trace), a.command.format(run="--kernel-trace --stats"), a.out)
        return
    s = summarize(a.workload, a.copy_mib, a.trace, a.fetch, a.write, a.command, a.out)
    print(json.dumps(s["kernels"], indent=1))


if __name__ == "__main__":
    main()
