"""Builds libmxllg.so (the HIP kernels and their C ABI) in place for gfx950.

    python -m magex_amd.csrc.build

hipcc cross-compiles without a GPU.  Flags: -O3, C++17, gfx950 only; packed FP32 instructions are
disabled (-target-feature -packed-fp32-ops) because some of their forms have given wrong results on
this hardware generation.  -ffp-contract=on fuses a * b + c into one FMA only inside a single source
expression, so every kernel that spells a formula the same way rounds it the same way: the layouts,
the batched and the plain kernels and the captured and direct step loops agree bit for bit (the
default cross-statement contraction decides per kernel, from the surrounding code).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
SOURCES = ["mxllg_api.hip", "mxllg_kernels.hip", "mxllg_batch.hip", "mxllg_adjoint.hip", "mxllg_mc.hip", "mxllg_sampler.hip"]
TARGET = os.path.join(PKG, "libmxllg.so")
ARCH = "gfx950"


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def command(out=TARGET):
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=on",
            "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", HERE,
            *[os.path.join(HERE, s) for s in SOURCES], "-o", out]


def build(verbose=False):
    """Compile every source into libmxllg.so (always recompiles); returns the library path."""
    tmp = TARGET + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr.strip():
        print(r.stderr)
    os.replace(tmp, TARGET)
    return TARGET


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
