"""NICAM snapshot files of the reference's stand-alone kernel run, and the run itself.

The files (read and written by the C ABI, ``nsw6_snapshot_read/write`` and
``nsw6_verify_read/write``) are those src/main.cpp:128-174 reads from ``data/`` and
:326-337 reads from ``ref_verify/``: raw native doubles in the shapes of
src/data_io.cpp:207-402, ``data/vgrid/GRD_gz.dat`` / ``GRD_gzh.dat`` for the grid.

``run()`` is the reference's ``main`` for this path: read a run directory, call
mp_driver ``iterations`` times on the GPU with the harness's dt = TIME_DTL = 6 s and
l_region = SET_l = 1 (include/problemsize.h:79-82, src/problemsize.cpp:95), and, if
``ref_verify/`` is present, return the PROF_val_check table of the state against it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _native as N
from .check import prof_val_check_line
from .fields import TRACERS, alloc_host, field_ptrs

# the reference harness (include/problemsize.h:63,68,79-82)
ADM_GALL_IN = 16641
ADM_LALL = 1
TIME_DTL = 6.0
TIME_CTIME = 0.0
SET_L = 1
# keys of the verification set, in the order of src/main.cpp:339-350
VERIFY_KEYS = ("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge") + tuple(f"rhogq:{t}" for t in TRACERS)


class SnapshotError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {N.status_string(code)}")
        self.code = code


def _check(rc: int, what: str):
    if rc != N.NSW6_OK:
        raise SnapshotError(rc, what)


def read(run_dir: str, lall: int = ADM_LALL, ijdim: int = ADM_GALL_IN, *, grid: bool = True):
    """Host state dict (every mp_driver array; the outputs not in ``data/`` are zero, as
    the reference's freshly allocated Views) and the grid (gz, gzh) of a run directory."""
    st = alloc_host(lall, ijdim)
    gz = np.zeros(N.KALL) if grid else None
    gzh = np.zeros(N.KALL) if grid else None
    rc = N.lib().nsw6_snapshot_read(os.fsencode(run_dir), lall, ijdim, ctypes.byref(field_ptrs(st)),
                                    None if gz is None else gz.ctypes.data,
                                    None if gzh is None else gzh.ctypes.data)
    _check(rc, f"reading {run_dir}/data")
    return gz, gzh, st


def write(run_dir: str, state: dict, gz=None, gzh=None):
    """Write the ``data/`` files of a host state (and the grid when given)."""
    lall, ijdim = state["rhog"].shape[0], state["rhog"].shape[-1]
    os.makedirs(run_dir, exist_ok=True)
    g = None if gz is None else np.ascontiguousarray(gz, dtype=np.float64)
    gh = None if gzh is None else np.ascontiguousarray(gzh, dtype=np.float64)
    rc = N.lib().nsw6_snapshot_write(os.fsencode(run_dir), lall, ijdim, ctypes.byref(field_ptrs(state)),
                                     None if g is None else g.ctypes.data, None if gh is None else gh.ctypes.data)
    _check(rc, f"writing {run_dir}/data")


def read_verify(run_dir: str, lall: int = ADM_LALL, ijdim: int = ADM_GALL_IN) -> dict:
    """The verification set as {VERIFY_KEYS[i]: [l][96][ij] array}."""
    buf = np.zeros((N.NSW6_VERIFY_FIELDS, lall, N.KALL, ijdim))
    _check(N.lib().nsw6_verify_read(os.fsencode(run_dir), lall, ijdim, buf.ctypes.data),
           f"reading {run_dir}/ref_verify")
    return {k: buf[i] for i, k in enumerate(VERIFY_KEYS)}


def write_verify(run_dir: str, state: dict):
    lall, ijdim = state["rhog"].shape[0], state["rhog"].shape[-1]
    os.makedirs(run_dir, exist_ok=True)
    _check(N.lib().nsw6_verify_write(os.fsencode(run_dir), lall, ijdim, ctypes.byref(field_ptrs(state))),
           f"writing {run_dir}/ref_verify")


def verify_table(state: dict, verify: dict) -> list:
    """PROF_val_check lines (src/mod_debug.cpp:363-410) of a host state against a
    verification set, labelled as in src/main.cpp:339-350."""
    out = []
    for key in VERIFY_KEYS:
        if ":" in key:
            name, tr = key.split(":")
            computed = state[name][:, TRACERS.index(tr)]
            label = tr
        else:
            computed, label = state[key], key
        out.append(prof_val_check_line(label, computed, verify[key]))
    return out


def run(run_dir: str, *, lall: int = ADM_LALL, ijdim: int = ADM_GALL_IN, iterations: int = 1,
        dt: float = TIME_DTL, ct: float = TIME_CTIME, l_region: int = SET_L, precision: int = N.NSW6_FP64):
    """The reference's stand-alone run of mp_driver on a snapshot, on cuda:0.  Returns
    (host state after the run, check-table lines or None when ref_verify/ is absent)."""
    from .driver import NSW6
    from .fields import to_device, to_host
    gz, gzh, st = read(run_dir, lall, ijdim)
    dev = to_device(st)
    drv = NSW6(gz, gzh, max_columns=lall * ijdim, precision=precision)
    try:
        for _ in range(iterations):
            drv.step(dev, dt, ct, l_region=l_region)
    finally:
        drv.close()
    out = to_host(dev)
    lines = None
    if os.path.isdir(os.path.join(run_dir, "ref_verify")):
        lines = verify_table(out, read_verify(run_dir, lall, ijdim))
    return out, lines


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Run mp_driver on a NICAM snapshot directory (data/, ref_verify/)")
    ap.add_argument("run_dir")
    ap.add_argument("--lall", type=int, default=ADM_LALL)
    ap.add_argument("--ijdim", type=int, default=ADM_GALL_IN)
    ap.add_argument("--iterations", type=int, default=1)
    ap.add_argument("--dt", type=float, default=TIME_DTL)
    args = ap.parse_args(argv)
    _, lines = run(args.run_dir, lall=args.lall, ijdim=args.ijdim, iterations=args.iterations, dt=args.dt)
    if lines is None:
        print("no ref_verify/ in", args.run_dir)
    else:
        print("Checking Results")
        for line in lines:
            print(line)


if __name__ == "__main__":
    main()
