"""Build the native libraries of the package in-tree.

* ``libnsw6.so``       -- HIP kernels for gfx950 (the step and its reverse mode) +
                          the C ABI of ``include/nsw6.h`` (hipcc), plus the host
                          set-up code (gcc).
* ``libnsw6_synth.so`` -- the seeded synthetic-input generator (gcc + OpenMP).

Floating-point options per unit: the FP64 step kernels with FMA contraction and
divisions on the reciprocal sequence (``KERNEL_FLAGS``); the reverse mode with
contraction inside expressions and shared reciprocals (``ADJ_FLAGS``); the
mixed-precision rate kernel with contraction only (``MIXED_FLAGS``); the host-side
ABI code without contraction.  Packed
FP32 VALU forms are disabled everywhere (``-packed-fp32-ops``).
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
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# The FP64 step is held to 1e-10 of each field's max-abs against the oracle, not to the
# oracle's rounding sequence.  FMA contraction (a*b+c rounded once) and divisions on
# the reciprocal sequence (-fapprox-func: v_rcp_f64 and Newton steps instead of the
# scaled IEEE division) take about 6 % of the rate kernel's and 17 % of the finish
# kernel's instructions.  Over three steps of two 2 x 4,000-column states the worst
# difference from the oracle is 6.9e-12 of max-abs with them and 1.1e-11 without (the
# device's exp/log and the oracle's libm differ by a few ulp either way).
KERNEL_FLAGS = ["-ffp-contract=fast", "-fapprox-func"]
# The mixed form's FP32 rate kernel keeps the accurate exp, log, sqrt and division: the
# native FP32 approximations, amplified by the limiters over a 6-step rollout, turn its
# gradient away from the FP64 one (cosine 0.92 against the 0.99 required).  Contraction
# alone takes the kernel from 0.94 to 0.89 ms at 169,000 x 96 and its one-step error
# from 1.9e-6 to 2.4e-6 of max-abs (bar 1e-5).
MIXED_FLAGS = ["-ffp-contract=fast"]
# The reverse mode is held to 1e-6 relative: its divisions may also become products
# with one shared reciprocal (about a quarter fewer FP64 instructions in the cell and
# wall passes).  Contraction is kept within single expressions there: across
# statements it lengthens live ranges in the register-bound wall and point passes
# (reverse passes at 16,900 x 96: 1.76 ms contracting within expressions, 1.86 ms
# across statements, 1.89 ms not contracting).
ADJ_FLAGS = ["-ffp-contract=on", "-fapprox-func", "-freciprocal-math"]
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
    obj_m = os.path.join(BUILD, "nsw6_rates_mixed.o")
    obj_f = os.path.join(BUILD, "nsw6_finish.o")
    steps = [
        [HIPCC, *HIP_FLAGS, *KERNEL_FLAGS, "-c", os.path.join(CSRC, "nsw6_kernels.hip"), "-o", obj_k],
        [HIPCC, *HIP_FLAGS, *ADJ_FLAGS, "-c", os.path.join(CSRC, "nsw6_adjoint.hip"), "-o", obj_r],
        [HIPCC, *HIP_FLAGS, *MIXED_FLAGS, "-c", os.path.join(CSRC, "nsw6_rates_mixed.hip"), "-o", obj_m],
        [HIPCC, *HIP_FLAGS, *KERNEL_FLAGS, "-c", os.path.join(CSRC, "nsw6_finish.hip"), "-o", obj_f],
        [HIPCC, *HIP_FLAGS, "-ffp-contract=off", "-x", "hip", "-c", os.path.join(CSRC, "nsw6_api.cpp"), "-o", obj_a],
        ["gcc", *C_FLAGS, "-c", os.path.join(CSRC, "nsw6_host.c"), "-o", obj_h],
        ["gcc", *C_FLAGS, "-c", os.path.join(CSRC, "nsw6_snapshot.c"), "-o", obj_s],
    ]
    with ThreadPoolExecutor(max_workers=len(steps)) as pool:
        for fut in [pool.submit(_run, cmd, verbose) for cmd in steps]:
            fut.result()
    tmp = LIB_NSW6 + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, obj_k, obj_m, obj_f, obj_r, obj_a, obj_h, obj_s],
         verbose)
    os.replace(tmp, LIB_NSW6)
    tmp = LIB_SYNTH + ".tmp"
    _run(["gcc", *C_FLAGS, "-fopenmp", "-shared", "-o", tmp, os.path.join(CSRC, "nsw6_synth.c"),
          os.path.join(CSRC, "nsw6_host.c"), "-lm"], verbose)
    os.replace(tmp, LIB_SYNTH)


if __name__ == "__main__":
    build(verbose=True)
