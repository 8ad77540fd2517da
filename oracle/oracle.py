"""TEST INFRASTRUCTURE ONLY: ctypes loader of the CPU oracle (liboracle_nsw6.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_nsw6.so")

_lib = None


def build(verbose: bool = False) -> None:
    res = subprocess.run(["make", "-B", "-C", HERE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(res.stdout)
    if verbose:
        print(res.stdout)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        from nicam_kokkos_microphysics_amd import _native as N
        L = ctypes.CDLL(LIB)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.ora_gammafunc.restype = D
        L.ora_gammafunc.argtypes = [D]
        L.ora_params_init.argtypes = [ctypes.POINTER(N.Params)]
        L.ora_params_init.restype = None
        L.ora_grid_setup.argtypes = [P, P, ctypes.POINTER(N.Grid)]
        L.ora_grid_setup.restype = None
        L.ora_psat_liq.restype = D
        L.ora_psat_liq.argtypes = [D]
        L.ora_psat_ice.restype = D
        L.ora_psat_ice.argtypes = [D]
        L.ora_bergeron_point.argtypes = [D, P, P, P]
        L.ora_bergeron_point.restype = None
        L.ora_bigloop_point.argtypes = [ctypes.POINTER(N.Params), D, D, D, P, D, D, D, D, D, D, D, D, P]
        L.ora_bigloop_point.restype = None
        L.ora_vadv1d_prep.argtypes = [I, I, I, P, P, P, P, P, P, P, D]
        L.ora_vadv1d_prep.restype = None
        L.ora_vadv1d_getflux.argtypes = [I, I, I, P, P, P, P, P, P, P]
        L.ora_vadv1d_getflux.restype = None
        L.ora_satadjust_liq_point.argtypes = [D, D, D, D, P, P, P, P]
        L.ora_satadjust_liq_point.restype = I
        L.ora_mp_driver.argtypes = [ctypes.POINTER(N.Params), ctypes.POINTER(N.Grid), I, I,
                                    ctypes.POINTER(N.FieldPtrs), D, D, I, ctypes.POINTER(N.Report)]
        L.ora_mp_driver.restype = I
        L.ora_last_elapsed.restype = D
        L.ora_last_elapsed.argtypes = []
        L.ora_num_threads.restype = I
        L.ora_num_threads.argtypes = []
        _lib = L
    return _lib


class BigLoopOut(ctypes.Structure):
    _fields_ = [("drhogq", ctypes.c_double * 6), ("Vt", ctypes.c_double * 6), ("rceff", ctypes.c_double),
                ("rceff_cld", ctypes.c_double), ("rctop", ctypes.c_double), ("tctop", ctypes.c_double),
                ("rctop_cld", ctypes.c_double), ("tctop_cld", ctypes.c_double), ("wk", ctypes.c_double * 49),
                ("q_t", ctypes.c_double * 6)]


def params():
    from nicam_kokkos_microphysics_amd import _native as N
    p = N.Params()
    lib().ora_params_init(ctypes.byref(p))
    return p


def grid(gz, gzh):
    from nicam_kokkos_microphysics_amd import _native as N
    g = N.Grid()
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    lib().ora_grid_setup(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(g))
    return g


def mp_driver(state: dict, gz, gzh, dt: float, ct: float = 0.0, P=None, ice_adjust: int = -1):
    """Run the oracle step in place on a host state dict; returns the report."""
    from nicam_kokkos_microphysics_amd import _native as N
    from nicam_kokkos_microphysics_amd.fields import field_ptrs
    if P is None:
        P = params()
    G = grid(gz, gzh)
    lall, ijdim = state["rhog"].shape[0], state["rhog"].shape[-1]
    fp = field_ptrs(state)
    rep = N.Report()
    lib().ora_mp_driver(ctypes.byref(P), ctypes.byref(G), lall, ijdim, ctypes.byref(fp), float(dt), float(ct),
                        int(ice_adjust), ctypes.byref(rep))
    return rep


def bigloop_point(P, dt, dens, temp, q6, qsatl, qsati, pre, rhog, Nc, a1, a2, ma2):
    o = BigLoopOut()
    q = np.ascontiguousarray(q6, dtype=np.float64)
    lib().ora_bigloop_point(ctypes.byref(P), dt, dens, temp, q.ctypes.data, qsatl, qsati, pre, rhog, Nc, a1, a2,
                            ma2, ctypes.byref(o))
    return o


def last_elapsed() -> float:
    return lib().ora_last_elapsed()


def num_threads() -> int:
    return lib().ora_num_threads()
