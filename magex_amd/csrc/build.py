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
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
SOURCES = ["mxllg_api.hip", "mxllg_kernels.hip", "mxllg_steps.hip", "mxllg_batch.hip", "mxllg_adjoint.hip", "mxllg_mc.hip", "mxllg_sampler.hip"]
TARGET = os.path.join(PKG, "libmxllg.so")
ARCH = "gfx950"


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=on",
         "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", HERE]


def compile_command(src, obj):
    """One translation unit to an object file (device code is whole-program per unit: no -fgpu-rdc)."""
    return [hipcc(), *FLAGS, "-c", os.path.join(HERE, src), "-o", obj]


def link_command(objs, out=TARGET):
    return [hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", *objs, "-o", out]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def build(verbose=False, jobs=None):
    """Compile every source into libmxllg.so (always recompiles); returns the library path.

    The units compile in parallel (one hipcc per source, at most `jobs` at once, default the CPU
    count capped at the number of sources) and are then linked into one shared library.
    """
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with tempfile.TemporaryDirectory(prefix="mxllg_build_") as tmpdir:
        objs = [os.path.join(tmpdir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
        cmds = [compile_command(s, o) for s, o in zip(SOURCES, objs)]
        if verbose:
            for c in cmds:
                print(" ".join(c), flush=True)
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            logs = list(pool.map(_run, cmds))
        tmp = TARGET + ".tmp"
        logs.append(_run(link_command(objs, tmp)))
    if verbose:
        for log in logs:
            if log.strip():
                print(log)
    os.replace(tmp, TARGET)
    return TARGET


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
