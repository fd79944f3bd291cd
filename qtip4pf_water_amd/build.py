"""Build the gfx950 shared library libqtip4pf.so in-tree with hipcc.

Usage: python -m qtip4pf_water_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc", "qtip4pf.hip")
SETUP_SRC = os.path.join(PKG, "csrc", "qtp_setup.cpp")  # host-only
NOALLOC_SRC = os.path.join(PKG, "csrc", "qtp_noalloc.hip")
OUT = os.path.join(PKG, "libqtip4pf.so")
ARCH = os.environ.get("QTP_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no FMA contraction: the reference's separate multiply/add order is kept
    # where it matters (neighbour-list distances); hot loops call fma() explicitly
    "-ffp-contract=off",
    # packed FP32 instructions disabled
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, SRC, NOALLOC_SRC, SETUP_SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, OUT)
    return OUT


DRIVER_SRC = os.path.join(ROOT, "tests", "adapter", "adapter_driver.cpp")
DRIVER_BIN = os.path.join(ROOT, "tests", "_bin", "qtp_adapter_driver")


def build_adapter_driver(verbose: bool = False) -> str:
    """Test driver for include/qtip4pf_ref_adapter.hpp (links libqtip4pf.so)."""
    os.makedirs(os.path.dirname(DRIVER_BIN), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), DRIVER_SRC,
           "-L" + PKG, "-lqtip4pf", "-Wl,-rpath,$ORIGIN/../../qtip4pf_water_amd", "-o", DRIVER_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return DRIVER_BIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_adapter_driver(verbose=True))
