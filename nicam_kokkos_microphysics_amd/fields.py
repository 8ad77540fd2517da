"""The 50 arrays of ``mp_driver`` (include/mod_mp_driver.h:71-124): shapes, roles,
host/device containers and the seeded synthetic states.

Layout is NICAM's (ij,k,l) Fortran order seen from C: a row-major array
[l][k][ij] for 3-D fields, [l][nq][k][ij] for tracers, [l][ij] for column
vectors, [l][2][ij] for ``precip`` and [l][1][ij] for the cloud-top rows.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native as N

KALL, NQ = N.KALL, N.NQMAX
TRACERS = ("QV", "QC", "QR", "QI", "QS", "QG")

_KIND = {name: "3d" for name in N.FIELD_NAMES}
_KIND.update({"rhogq": "trc", "q": "trc"})
_KIND.update({n: "row2" for n in ("precip", "ISO1_precip", "ISO2_precip")})
_KIND.update({n: "col" for n in ("precip_rhoe", "precip_lh_heat", "precip_rhophi", "precip_rhokin",
                                  "ix", "iy", "iz", "jx", "jy", "jz")})
_KIND.update({n: "row1" for n in ("rctop", "rwtop", "tctop")})

# Arrays mp_driver reads (the rest are outputs, or unused by the NSW6 path).
INPUTS = ("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rhogq", "vx", "vy", "vz", "tem", "q",
          "gsgam2", "gsgam2h", "gam2", "gam2h", "ix", "iy", "iz", "jx", "jy", "jz", "z")
# Arrays mp_driver writes.
OUTPUTS = ("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rhogq", "rho", "tem", "pre", "q", "qd",
           "precip", "ISO1_precip", "ISO2_precip", "precip_rhoe", "precip_lh_heat", "precip_rhophi",
           "precip_rhokin", "re_liquid", "re_solid", "re_cld", "rctop", "rwtop", "tctop", "GDCLW", "GDCFRC",
           "GPREC")
# Reverse mode: outputs that take a cotangent, inputs that receive one.
ADJ_OUTPUTS = ("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rhogq", "rho", "tem", "pre", "q", "qd",
               "precip", "precip_rhoe", "precip_lh_heat", "precip_rhophi", "precip_rhokin", "GPREC")
ADJ_INPUTS = ("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rhogq", "vx", "vy", "vz", "tem", "q")
# The reference's check table (src/main.cpp:339-350): rhogq per tracer, not q.
CHECK_FIELDS = ("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge") + tuple(f"rhogq:{t}" for t in TRACERS)


def shape_of(name: str, lall: int, ijdim: int):
    k = _KIND[name]
    if k == "3d":
        return (lall, KALL, ijdim)
    if k == "trc":
        return (lall, NQ, KALL, ijdim)
    if k == "col":
        return (lall, ijdim)
    if k == "row2":
        return (lall, 2, ijdim)
    return (lall, 1, ijdim)


def alloc_host(lall: int, ijdim: int) -> dict:
    return {n: np.zeros(shape_of(n, lall, ijdim), dtype=np.float64) for n in N.FIELD_NAMES}


def field_view(state: dict, key: str):
    """`rhogq:QR` style keys select one tracer of a tracer array."""
    if ":" in key:
        name, tr = key.split(":")
        return state[name][:, TRACERS.index(tr)]
    return state[key]


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"] or a.dtype != np.float64:
            raise ValueError("arrays must be C-contiguous float64")
        return a.ctypes.data
    # torch tensor
    if not a.is_contiguous() or str(a.dtype) != "torch.float64":
        raise ValueError("tensors must be contiguous float64")
    return a.data_ptr()


def field_ptrs(state: dict) -> N.FieldPtrs:
    fp = N.FieldPtrs()
    for name in N.FIELD_NAMES:
        setattr(fp, name, _ptr(state.get(name)))
    return fp


# ---------------------------------------------------------------------------------
# synthetic states (workloads of the benchmark and the parity tests)
# ---------------------------------------------------------------------------------
@dataclass(frozen=True)
class Workload:
    name: str
    lall: int
    ijdim: int
    dt: float
    steps: int
    frac_supersat: float = 0.0
    frac_towers: float = 0.0
    seed: int = 20240101
    frac_rainshafts: float = 0.0


WORKLOADS = {
    # 130x130 columns, one region: the reference's CPU-runnable shape
    "nsw6_16900x96_fp64": Workload("nsw6_16900x96_fp64", 1, 16900, 15.0, 1),
    # one GPU's share of a global grid: 10 regions x 16,900 columns, 10 steps
    "nsw6_169000x96_fp64": Workload("nsw6_169000x96_fp64", 10, 16900, 15.0, 10),
    # mixed-precision workload: plus 5 % supersaturated columns
    "nsw6_169000x96_mixed": Workload("nsw6_169000x96_mixed", 10, 16900, 15.0, 10, frac_supersat=0.05),
    # adjoint workload
    "nsw6_16900x96_adjoint": Workload("nsw6_16900x96_adjoint", 1, 16900, 15.0, 1),
    # the FP64 step on the reference's vgrid (72 m bottom layers) at dt = 30 s with deep
    # convection (10 % towers, 1 % rain shafts): rain falls through two layers, so every
    # region takes the general wall fluxes (6 steps: the reference's own saturation
    # adjustment stops converging at a few points from the 9th step on)
    "nsw6_169000x96_refgrid_dt30": Workload("nsw6_169000x96_refgrid_dt30", 10, 16900, 30.0, 6, frac_towers=0.10,
                                            frac_rainshafts=0.01),
    # long rollout: 40 regions, 20 steps, 10 % deep convective towers
    "nsw6_676000x96_rollout": Workload("nsw6_676000x96_rollout", 40, 16900, 15.0, 20, frac_towers=0.10),
}


def synth_grid():
    """Power-law 94-layer grid over 0-40 km (dz 80 m to 1 km); returns (gz, gzh)."""
    gz = np.zeros(KALL)
    gzh = np.zeros(KALL)
    N.synth_lib().nsw6_synth_grid(gz.ctypes.data, gzh.ctypes.data)
    return gz, gzh


def synth_state(lall: int, ijdim: int, *, seed: int = 20240101, frac_supersat: float = 0.0,
                frac_towers: float = 0.0, rh_lo: float = 0.5, rh_hi: float = 1.05, l_offset: int = 0,
                ij_offset: int = 0, gz=None, gzh=None, tsfc: float | None = None,
                frac_rainshafts: float = 0.0) -> tuple:
    """Seeded, self-consistent host state: (gz, gzh, state).  Columns are a pure
    function of (seed, l + l_offset, ij + ij_offset), so sub-blocks of a large run can
    be regenerated exactly."""
    if gz is None or gzh is None:
        gz, gzh = synth_grid()
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    st = alloc_host(lall, ijdim)
    cfg = N.SynthConfig()
    S = N.synth_lib()
    S.nsw6_synth_default_config(ctypes.byref(cfg))
    cfg.seed = seed
    cfg.frac_supersat = frac_supersat
    cfg.frac_towers = frac_towers
    cfg.frac_rainshafts = frac_rainshafts
    cfg.rh_lo = rh_lo
    cfg.rh_hi = rh_hi
    if tsfc is not None:
        cfg.tsfc = tsfc
    fp = field_ptrs(st)
    rc = S.nsw6_synth_fill(ctypes.byref(cfg), gz.ctypes.data, gzh.ctypes.data, lall, ijdim, l_offset,
                           ij_offset, ctypes.byref(fp))
    if rc != 0:
        raise RuntimeError(f"nsw6_synth_fill failed: {rc}")
    return gz, gzh, st


def workload_state(w: Workload, lall: int | None = None, ijdim: int | None = None, **kw):
    return synth_state(lall or w.lall, ijdim or w.ijdim, seed=w.seed, frac_supersat=w.frac_supersat,
                       frac_towers=w.frac_towers, frac_rainshafts=w.frac_rainshafts, **kw)


def to_device(state: dict, device="cuda:0") -> dict:
    import torch
    return {k: torch.from_numpy(v).to(device) for k, v in state.items()}


def to_host(state: dict) -> dict:
    return {k: v.detach().cpu().numpy().copy() for k, v in state.items()}


def copy_state(state: dict) -> dict:
    """Deep copy of a host (numpy) or device (torch) state."""
    return {k: (v.clone() if hasattr(v, "clone") else v.copy()) for k, v in state.items()}
