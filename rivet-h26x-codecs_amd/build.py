"""Build the HIP tier in-tree: rivet_hip/librivet_hip.so for gfx950, and
rivet_hip/librivet_gen.so, the host-only generator of synthetic work lists.

Each HIP source is compiled with hipcc for gfx950 only; the generator with
g++. Both libraries are written next to the Python package so they travel
with the source tree and are what the package loads.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "rivet_hip" / "librivet_hip.so"
GEN_OUT = HERE / "rivet_hip" / "librivet_gen.so"
OBJ = HERE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # No packed-FP32 instruction selection (none is wanted in integer code).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]

HIP_SOURCES = ["hevc.hip", "h264.hip", "runtime.hip", "entry.hip", "hash.hip"]
CXX_SOURCES = ["gen.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    # hipcc's host pass reports the device-only target feature; not an error.
    noise = [l for l in (r.stdout + r.stderr).splitlines()
             if l.strip() and "packed-fp32-ops" not in l and "not a recognized feature" not in l]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = OBJ / (s + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(o)])
    for s in CXX_SOURCES:
        o = OBJ / (s + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(CSRC / s), "-o", str(o)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    objs = [str(OBJ / (s + ".o")) for s in HIP_SOURCES]
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", str(tmp), *objs])
    os.replace(tmp, OUT)
    # The synthetic work-list generator is a host library of its own: the
    # device tier carries no test-workload code.
    gtmp = GEN_OUT.with_suffix(".so.tmp")
    _run(["g++", "-shared", "-o", str(gtmp), *[str(OBJ / (s + ".o")) for s in CXX_SOURCES]])
    os.replace(gtmp, GEN_OUT)
    if verbose:
        print("built", OUT, "and", GEN_OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
