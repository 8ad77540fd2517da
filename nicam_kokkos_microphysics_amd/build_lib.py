"""Build the native libraries of the package in-tree.

* ``libnsw6.so``       -- HIP kernels for gfx950 (the step and its reverse mode) +
                          the C ABI of ``include/nsw6.h`` (hipcc), plus the host
                          set-up code (gcc).
* ``libnsw6_synth.so`` -- the seeded synthetic-input generator (gcc + OpenMP).

The FP64 kernels are compiled without FMA contraction (``-ffp-contract=off``) so
that every add/multiply rounds like the reference's CPU build; packed FP32 VALU
forms are disabled (``-packed-fp32-ops``).  The reverse-mode kernels additionally
allow reciprocal products for divisions (``-freciprocal-math``).
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("NSW6_OFFLOAD_ARCH", "gfx950")

LIB_NSW6 = os.path.join(HERE, "libnsw6.so")
LIB_SYNTH = os.path.join(HERE, "libnsw6_synth.so")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# The reverse mode is held to 1e-6 relative, not to the forward's round-off: its
# divisions may become products with one shared reciprocal (about a quarter fewer
# FP64 instructions in the cell and wall passes), and the remaining FP64 divisions take
# the reciprocal sequence instead of the scaled IEEE one (another 5-12 %).
ADJ_FLAGS = ["-freciprocal-math", "-fapprox-func"]
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-fno-fast-math", "-ffp-contract=off", "-Wall"]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")
    if verbose and res.stdout.strip():
        print(res.stdout)


def build(verbose: bool = True) -> None:
    """Compile both libraries (always recompiles)."""
    os.makedirs(BUILD, exist_ok=True)
    obj_k = os.path.join(BUILD, "nsw6_kernels.o")
    obj_r = os.path.join(BUILD, "nsw6_adjoint.o")
    obj_a = os.path.join(BUILD, "nsw6_api.o")
    obj_h = os.path.join(BUILD, "nsw6_host.o")
    obj_s = os.path.join(BUILD, "nsw6_snapshot.o")
    steps = [
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "nsw6_kernels.hip"), "-o", obj_k],
        [HIPCC, *HIP_FLAGS, *ADJ_FLAGS, "-c", os.path.join(CSRC, "nsw6_adjoint.hip"), "-o", obj_r],
        [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, "nsw6_api.cpp"), "-o", obj_a],
        ["gcc", *C_FLAGS, "-c", os.path.join(CSRC, "nsw6_host.c"), "-o", obj_h],
        ["gcc", *C_FLAGS, "-c", os.path.join(CSRC, "nsw6_snapshot.c"), "-o", obj_s],
    ]
    with ThreadPoolExecutor(max_workers=len(steps)) as pool:
        for fut in [pool.submit(_run, cmd, verbose) for cmd in steps]:
            fut.result()
    tmp = LIB_NSW6 + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, obj_k, obj_r, obj_a, obj_h, obj_s],
         verbose)
    os.replace(tmp, LIB_NSW6)
    tmp = LIB_SYNTH + ".tmp"
    _run(["gcc", *C_FLAGS, "-fopenmp", "-shared", "-o", tmp, os.path.join(CSRC, "nsw6_synth.c"),
          os.path.join(CSRC, "nsw6_host.c"), "-lm"], verbose)
    os.replace(tmp, LIB_SYNTH)


if __name__ == "__main__":
    build(verbose=True)
