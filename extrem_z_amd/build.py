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


def build_compat(hipcc, verbose=False):
    """libExtreMz: the reference library's API (include/api/ExtreM.h) over
    libextremz.so, as a shared library and as the static archive the
    reference's CMake target ExtreMz::ExtreMz names."""
    src = os.path.join(SRC, "compat.hip")
    obj = os.path.join(OUT_DIR, "compat.o")
    flags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-ldl", "-pthread")]
    cmd = [hipcc, *flags, "-c", src, "-o", obj + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(obj + ".tmp", obj)
    so = os.path.join(OUT_DIR, "libExtreMz.so")
    subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", obj, "-o", so + ".tmp",
                    "-L" + OUT_DIR, "-lextremz", "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(so + ".tmp", so)
    ar = os.path.join(OUT_DIR, "libExtreMz.a")
    if os.path.exists(ar + ".tmp"):
        os.remove(ar + ".tmp")
    subprocess.run(["ar", "rcs", ar + ".tmp", obj], check=True)
    os.replace(ar + ".tmp", ar)
    # the test program that drives the API in main.cpp's order
    root = os.path.dirname(HERE)
    drv = os.path.join(root, "tests", "compat", "_build", "dropin_main")
    os.makedirs(os.path.dirname(drv), exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I" + os.path.join(root, "include", "api"),
                    os.path.join(root, "tests", "compat", "dropin_main.cpp"), "-o", drv + ".tmp",
                    "-L" + OUT_DIR, "-lExtreMz", "-lextremz",
                    "-Wl,-rpath,$ORIGIN/../../../extrem_z_amd/_build"], check=True)
    os.replace(drv + ".tmp", drv)
    return so


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, *[os.path.join(SRC, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_compat(hipcc, verbose)
    cli = os.path.join(OUT_DIR, "extremz_cli")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(SRC, "cli.cpp"), "-o", cli + ".tmp",
                    "-L" + OUT_DIR, "-lextremz", "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(cli + ".tmp", cli)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
