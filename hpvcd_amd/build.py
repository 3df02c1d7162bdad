"""Build recipe for the native parts of hpvcd_amd.

``build_library()`` compiles the gfx950 HIP kernels, the host runtime and the
syntax generator into ``hpvcd_amd/libhpvcd.so`` with hipcc.  The library is
built in-tree so that it travels with the source tree; nothing is installed.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhpvcd.so")

# host C++ units: record generator and its HEVC writer, coding-tree walker,
# motion derivations, CABAC, parameter sets, syntax parser, HEIF container, decoder object,
# serving loop
HOST_UNITS = ["gen.cpp", "gen_hevc.cpp", "walker.cpp", "motion.cpp", "cabac.cpp", "hevc_ps.cpp", "parse.cpp", "heif.cpp", "decoder.cpp", "server.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HPVCD_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are switched off for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = proc.stdout
    # the host pass does not know the device target feature; that note is noise
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l]
    if lines:
        print("\n".join(lines), flush=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed ({proc.returncode}): {' '.join(cmd)}")


def build_library(verbose: bool = True) -> str:
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    host = ["-O3", "-std=c++17", "-fPIC", "-Wall"]
    units = [("runtime.hip", HIP_FLAGS + ["-x", "hip"])] + [(u, host) for u in HOST_UNITS]
    jobs = []
    for src, flags in units:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        jobs.append([HIPCC] + flags + ["-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as pool:
        for f in [pool.submit(_run, j) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build_library()
    sys.exit(0)
