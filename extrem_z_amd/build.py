"""Builds the HIP engine for gfx950 into extrem_z_amd/_build/libextremz.so."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libextremz.so")
SOURCES = ["engine.hip", "kernels.hip", "pool.hip", "cost.hip", "encode_noalloc.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # IEEE float64 without contraction on host and device, as the reference
    "-ffp-contract=off",
    # packed fp32 instructions stay off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-pthread", "-ldl",
]


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, *[os.path.join(SRC, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    cli = os.path.join(OUT_DIR, "extremz_cli")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(SRC, "cli.cpp"), "-o", cli + ".tmp",
                    "-L" + OUT_DIR, "-lextremz", "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(cli + ".tmp", cli)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
