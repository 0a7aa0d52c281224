"""Builds libpureadsorb_amd.so (host library + gfx950 kernels) in-tree with hipcc.

Usage: python pureadsorb.jl_amd/build.py   (no GPU needed: hipcc cross-compiles for gfx950)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "pureadsorb_amd", "libpureadsorb_amd.so")
SOURCES = ["pureadsorb_amd.hip", "widom_kernels.hip", "widom_binned.hip", "widom_grad.hip", "mc_kernels.hip",
           "mc_host.hip", "audit_kernels.hip", "nvt_kernels.hip", "batch_kernels.hip", "io_readers.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off (see the kernels' header comment)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
]


SYNTH_OUT = os.path.join(HERE, "pureadsorb_amd", "libpa_synth.so")


def build_synth(verbose: bool = False) -> str:
    """Host-only helper of the synthetic-input generator (gcc, no GPU code)."""
    tmp = SYNTH_OUT + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", tmp, os.path.join(CSRC, "synth.c"), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, SYNTH_OUT)
    return SYNTH_OUT


def build(verbose: bool = False) -> str:
    build_synth(verbose)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
