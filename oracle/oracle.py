"""TEST INFRASTRUCTURE ONLY: ctypes loaders of the CPU oracle (liboracle_nsw6.so) and
of the host dual-number gradient check of the adjoint (libnsw6_grad.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NSW6_ORACLE_LIB / NSW6_GRAD_LIB select other builds of the two libraries (the
# sanitizer builds of `make -C oracle sanitize`, under oracle/_san/)
LIB = os.environ.get("NSW6_ORACLE_LIB") or os.path.join(HERE, "liboracle_nsw6.so")
GRAD_LIB = os.environ.get("NSW6_GRAD_LIB") or os.path.join(HERE, "libnsw6_grad.so")
GRAD_NDIR = 8

_lib = None
_glib = None


def build(verbose: bool = False) -> None:
    res = subprocess.run(["make", "-B", "-C", HERE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(res.stdout)
    if verbose:
        print(res.stdout)


def declare(L):
    """ctypes signatures of liboracle_nsw6.so's entry points on a loaded library `L`
    (the oracle, or a build of it with a planted mistake)."""
    from nicam_kokkos_microphysics_amd import _native as N
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
    L.ora_satadjust_all_point.argtypes = [D, D, D, D, P, D, D, P, P, P, P]
    L.ora_satadjust_all_point.restype = I
    L.ora_mp_driver.argtypes = [ctypes.POINTER(N.Params), ctypes.POINTER(N.Grid), I, I,
                                ctypes.POINTER(N.FieldPtrs), D, D, I, ctypes.POINTER(N.Report)]
    L.ora_mp_driver.restype = I
    L.ora_mp_driver_div.argtypes = [ctypes.POINTER(N.Params), ctypes.POINTER(N.Grid), I, I,
                                    ctypes.POINTER(N.FieldPtrs), D, D, I, I, ctypes.POINTER(N.Report)]
    L.ora_mp_driver_div.restype = I
    L.ora_record_kcell_bounds.argtypes = [P]
    L.ora_record_kcell_bounds.restype = None
    L.ora_record_newton_margin.argtypes = [P]
    L.ora_record_newton_margin.restype = None
    L.ora_thrmdyn.argtypes = [I, P, P, P, P, P, P, P]
    L.ora_thrmdyn.restype = None
    L.ora_negative_filter.argtypes = [I, P, P, P, P, P, P, P, P]
    L.ora_negative_filter.restype = None
    L.ora_rhogkin.argtypes = [ctypes.POINTER(N.Grid), I] + [P] * 10
    L.ora_rhogkin.restype = None
    L.ora_precip_transport.argtypes = [ctypes.POINTER(N.Params), ctypes.POINTER(N.Grid), I,
                                       ctypes.POINTER(N.FieldPtrs), P, D]
    L.ora_precip_transport.restype = None
    L.ora_saturation_adjustment.argtypes = [ctypes.POINTER(N.Params), I] + [P] * 7
    L.ora_saturation_adjustment.restype = None
    L.ora_last_elapsed.restype = D
    L.ora_last_elapsed.argtypes = []
    L.ora_num_threads.restype = I
    L.ora_num_threads.argtypes = []
    return L


def load(path: str):
    """A separately loaded, declared copy of an oracle build at `path`."""
    return declare(ctypes.CDLL(path))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = declare(ctypes.CDLL(LIB))
    return _lib


class BigLoopOut(ctypes.Structure):
    _fields_ = [("drhogq", ctypes.c_double * 6), ("Vt", ctypes.c_double * 6), ("rceff", ctypes.c_double),
                ("rceff_cld", ctypes.c_double), ("rctop", ctypes.c_double), ("tctop", ctypes.c_double),
                ("rctop_cld", ctypes.c_double), ("tctop_cld", ctypes.c_double), ("cloud_sw", ctypes.c_double),
                ("wk", ctypes.c_double * 49),
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


def mp_driver(state: dict, gz, gzh, dt: float, ct: float = 0.0, P=None, ice_adjust: int = -1,
              div_num: int = 1):
    """Run the oracle step in place on a host state dict; returns the report.
    div_num > 1 is the MP_DIV_NUM sub-cycling (src/mod_mp_driver.cpp:420-612)."""
    from nicam_kokkos_microphysics_amd import _native as N
    from nicam_kokkos_microphysics_amd.fields import field_ptrs
    if P is None:
        P = params()
    G = grid(gz, gzh)
    lall, ijdim = state["rhog"].shape[0], state["rhog"].shape[-1]
    fp = field_ptrs(state)
    rep = N.Report()
    if div_num == 1:
        lib().ora_mp_driver(ctypes.byref(P), ctypes.byref(G), lall, ijdim, ctypes.byref(fp), float(dt), float(ct),
                            int(ice_adjust), ctypes.byref(rep))
    else:
        lib().ora_mp_driver_div(ctypes.byref(P), ctypes.byref(G), lall, ijdim, ctypes.byref(fp), float(dt),
                                float(ct), int(ice_adjust), int(div_num), ctypes.byref(rep))
    return rep


def bigloop_point(P, dt, dens, temp, q6, qsatl, qsati, pre, rhog, Nc, a1, a2, ma2, L=None):
    o = BigLoopOut()
    q = np.ascontiguousarray(q6, dtype=np.float64)
    (L or lib()).ora_bigloop_point(ctypes.byref(P), dt, dens, temp, q.ctypes.data, qsatl, qsati, pre, rhog, Nc, a1, a2,
                            ma2, ctypes.byref(o))
    return o


def mp_driver_with_bounds(state: dict, gz, gzh, dt: float, P=None, nt: int = 3):
    """ora_mp_driver in place, also returning the kcell bounds of its vadv1d_prep calls
    as an int array [l][nt][2 (full, half)][96][2 (kcell_min, kcell_max)]."""
    rep, b, _ = mp_driver_traced(state, gz, gzh, dt, P=P, nt=nt, newton=False)
    return rep, b


def mp_driver_traced(state: dict, gz, gzh, dt: float, P=None, nt: int = 3, bounds: bool = True,
                     newton: bool = True):
    """ora_mp_driver in place with its test hooks on: returns (report, kcell bounds
    [l][nt][2][96][2] or None, Newton stopping margins [l][96][ij] or None; see
    ora_record_newton_margin)."""
    lall, ijdim = state["rhog"].shape[0], state["rhog"].shape[-1]
    b = np.full((lall, nt, 2, 96, 2), -99, dtype=np.int32) if bounds else None
    m = np.full((lall, 96, ijdim), 1e300) if newton else None
    L = lib()
    if bounds:
        L.ora_record_kcell_bounds(b.ctypes.data)
    if newton:
        L.ora_record_newton_margin(m.ctypes.data)
    try:
        rep = mp_driver(state, gz, gzh, dt, P=P)
    finally:
        L.ora_record_kcell_bounds(None)
        L.ora_record_newton_margin(None)
    return rep, b, m


# ---------------------------------------------------------------------------------
# single phases of the step on one region ([k][ij], tracers [6][k][ij]); L = a build
# of the oracle (default: the oracle)
# ---------------------------------------------------------------------------------
def _c(a):
    assert a.flags["C_CONTIGUOUS"] and a.dtype == np.float64
    return a.ctypes.data


def thrmdyn(q, ein, rho, L=None):
    """THRMDYN_qd/cv/tempre: returns (qd, cv, tem, pre)."""
    ijdim = q.shape[-1]
    qd, cv, tem, pre = (np.zeros_like(ein) for _ in range(4))
    (L or lib()).ora_thrmdyn(ijdim, _c(q), _c(ein), _c(rho), _c(qd), _c(cv), _c(tem), _c(pre))
    return qd, cv, tem, pre


def negative_filter(rhog, rhoge, rhogq, rho, tem, pre, q, gsgam2, L=None):
    """negative_filter in place on one region."""
    (L or lib()).ora_negative_filter(rhog.shape[-1], _c(rhog), _c(rhoge), _c(rhogq), _c(rho), _c(tem), _c(pre),
                                     _c(q), _c(gsgam2))


def rhogkin(gz, gzh, rhog, rhogvx, rhogvy, rhogvz, rhogw, gsgam2, gsgam2h, L=None):
    """cnvvar_rhogkin_in: returns (rhogkin, rhogkin_h, rhogkin_v)."""
    G = grid(gz, gzh)
    out = [np.zeros_like(rhog) for _ in range(3)]
    (L or lib()).ora_rhogkin(ctypes.byref(G), rhog.shape[-1], _c(rhog), _c(rhogvx), _c(rhogvy), _c(rhogvz),
                             _c(rhogw), _c(gsgam2), _c(gsgam2h), *[_c(o) for o in out])
    return tuple(out)


def precip_transport(state: dict, gz, gzh, Vterm, dt: float, P=None, L=None):
    """precip_transport_new in place on region 0 of a host state; Vterm [6][96][ij]."""
    from nicam_kokkos_microphysics_amd.fields import field_ptrs
    P = params() if P is None else P
    G = grid(gz, gzh)
    fp = field_ptrs(state)
    V = np.ascontiguousarray(Vterm, dtype=np.float64)
    (L or lib()).ora_precip_transport(ctypes.byref(P), ctypes.byref(G), state["rhog"].shape[-1], ctypes.byref(fp),
                                      V.ctypes.data, float(dt))


def saturation_adjustment(rhog, rhoge, rhogq, tem, q, qd, gsgam2, P=None, L=None):
    """SATURATION_adjustment in place on one region."""
    P = params() if P is None else P
    (L or lib()).ora_saturation_adjustment(ctypes.byref(P), rhog.shape[-1], _c(rhog), _c(rhoge), _c(rhogq),
                                           _c(tem), _c(q), _c(qd), _c(gsgam2))


def last_elapsed() -> float:
    return lib().ora_last_elapsed()


def num_threads() -> int:
    return lib().ora_num_threads()


# ---------------------------------------------------------------------------------
# host gradient check of the adjoint (libnsw6_grad.so)
# ---------------------------------------------------------------------------------
def grad_lib():
    global _glib
    if _glib is None:
        if not os.path.exists(GRAD_LIB):
            build()
        from nicam_kokkos_microphysics_amd import _native as N
        L = ctypes.CDLL(GRAD_LIB)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.grad_step_fp64.argtypes = [P, P, ctypes.POINTER(N.Params), I, I, ctypes.POINTER(N.FieldPtrs), D]
        L.grad_step_fp64.restype = I
        L.grad_loss_fp64.argtypes = [P, P, ctypes.POINTER(N.Params), P, I, I, ctypes.POINTER(N.FieldPtrs), D, P]
        L.grad_loss_fp64.restype = I
        L.grad_dual_loss.argtypes = [P, P, ctypes.POINTER(N.Params), I, I, ctypes.POINTER(N.FieldPtrs), D, I,
                                     ctypes.POINTER(N.FieldPtrs), P, P, P]
        L.grad_dual_loss.restype = I
        L.grad_check_pieces.argtypes = [P, P, ctypes.POINTER(N.Params), I, I, ctypes.POINTER(N.FieldPtrs), D, P, P]
        L.grad_check_pieces.restype = I
        FP = ctypes.POINTER(N.FieldPtrs)
        L.grad_sflux_fp64.argtypes = [P, P, ctypes.POINTER(N.Params), I, I, FP, D, P]
        L.grad_sflux_fp64.restype = I
        L.grad_adjoint_host.argtypes = [P, P, ctypes.POINTER(N.Params), I, I, FP, FP, P, FP, P, D]
        L.grad_adjoint_host.restype = I
        L.grad_dual_vjp.argtypes = [P, P, ctypes.POINTER(N.Params), I, I, FP, D, FP, P, I, FP, P, P, P]
        L.grad_dual_vjp.restype = I
        L.grad_math_tang.argtypes = [I, P, P, ctypes.c_int64]
        L.grad_math_tang.restype = I
        L.grad_math_tables.argtypes = [P, P]
        L.grad_math_tables.restype = None
        _glib = L
    return _glib


def _shape(state):
    return state["rhog"].shape[0], state["rhog"].shape[-1]


def step_fp64(state: dict, gz, gzh, dt: float, P=None) -> int:
    """The templated step in FP64 on the host, in place (0 = ok, 1 = Newton failure)."""
    from nicam_kokkos_microphysics_amd.fields import field_ptrs
    P = params() if P is None else P
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    lall, ijdim = _shape(state)
    fp = field_ptrs(state)
    return grad_lib().grad_step_fp64(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(P), lall, ijdim,
                                     ctypes.byref(fp), float(dt))


def loss_fp64(state: dict, gz, gzh, dt: float, coef=None, P=None) -> float:
    """The adjoint-check loss of one FP64 host step (state not modified); `coef`
    overrides the 8 tunable coefficients."""
    from nicam_kokkos_microphysics_amd.fields import field_ptrs
    P = params() if P is None else P
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    lall, ijdim = _shape(state)
    fp = field_ptrs(state)
    out = ctypes.c_double()
    c = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
    st = grad_lib().grad_loss_fp64(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(P),
                                   None if c is None else c.ctypes.data, lall, ijdim, ctypes.byref(fp), float(dt),
                                   ctypes.addressof(out))
    if st < 0:
        raise RuntimeError(f"grad_loss_fp64 status {st}")
    return out.value


def dual_loss(state: dict, gz, gzh, dt: float, dirs, dcoef=None, P=None):
    """Loss and its directional derivatives along the directions `dirs` (a list of
    dicts of direction arrays keyed like the state; missing keys are zero) and the
    coefficient directions `dcoef` [len(dirs)][8], by forward-mode dual numbers."""
    from nicam_kokkos_microphysics_amd import _native as N
    from nicam_kokkos_microphysics_amd.fields import field_ptrs
    P = params() if P is None else P
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    lall, ijdim = _shape(state)
    fp = field_ptrs(state)
    ndir = len(dirs)
    out = np.zeros(ndir)
    loss = 0.0
    for c0 in range(0, ndir, GRAD_NDIR):
        chunk = dirs[c0:c0 + GRAD_NDIR]
        arr = (N.FieldPtrs * len(chunk))()
        keep = []
        for i, d in enumerate(chunk):
            dd = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in d.items()}
            keep.append(dd)
            arr[i] = field_ptrs(dd)
        dc = np.zeros((len(chunk), 8)) if dcoef is None else np.ascontiguousarray(dcoef[c0:c0 + GRAD_NDIR],
                                                                                dtype=np.float64)
        lv = ctypes.c_double()
        dl = np.zeros(len(chunk))
        st = grad_lib().grad_dual_loss(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(P), lall, ijdim,
                                       ctypes.byref(fp), float(dt), len(chunk), arr, dc.ctypes.data,
                                       ctypes.addressof(lv), dl.ctypes.data)
        if st < 0:
            raise RuntimeError(f"grad_dual_loss status {st}")
        loss = lv.value
        out[c0:c0 + len(chunk)] = dl
    return loss, out


def check_pieces(state: dict, gz, gzh, dt: float, P=None):
    """Largest relative mismatch between each hand-written reverse piece of the step
    (point rates, wall fluxes, cell divergence/KIN2EIN, point-wise tail) and forward-
    mode dual numbers over every point, wall and cell of `state`; returns (err[4],
    count[4]).  Raises if a region needs the general transport."""
    from nicam_kokkos_microphysics_amd.fields import field_ptrs
    P = params() if P is None else P
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    lall, ijdim = _shape(state)
    fp = field_ptrs(state)
    err = np.zeros(4)
    cnt = np.zeros(4, dtype=np.int64)
    st = grad_lib().grad_check_pieces(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(P), lall, ijdim,
                                      ctypes.byref(fp), float(dt), err.ctypes.data, cnt.ctypes.data)
    if st != 0:
        raise RuntimeError(f"grad_check_pieces status {st}")
    return err, cnt


def sflux_fp64(state: dict, gz, gzh, dt: float, P=None):
    """Surface mass fluxes -f_t(kmin)/dt of the falling tracers, [l][NT][ij], of one
    FP64 host step (state not modified): NT = 3 (QR, QS, QG), or 4 (QR, QI, QS, QG)
    under four_water."""
    from nicam_kokkos_microphysics_amd.fields import field_ptrs
    P = params() if P is None else P
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    lall, ijdim = _shape(state)
    out = np.zeros((lall, 4 if P.four_water != 0.0 else 3, ijdim))
    fp = field_ptrs(state)
    st = grad_lib().grad_sflux_fp64(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(P), lall, ijdim,
                                    ctypes.byref(fp), float(dt), out.ctypes.data)
    if st < 0:
        raise RuntimeError(f"grad_sflux_fp64 status {st}")
    return out


def adjoint_host(state: dict, gz, gzh, dt: float, cot: dict, cot_sflux=None, P=None):
    """Host run of the reverse-mode passes (the HIP kernels' bodies) with the
    arguments of nsw6_mp_driver_adj: returns (grad dict keyed by fields.ADJ_INPUTS,
    grad_coef[8], status)."""
    from nicam_kokkos_microphysics_amd.fields import ADJ_INPUTS, field_ptrs
    P = params() if P is None else P
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    lall, ijdim = _shape(state)
    cot = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in cot.items()}
    grad = {k: np.zeros_like(state[k], dtype=np.float64) for k in ADJ_INPUTS}
    gc = np.zeros(8)
    cs = None if cot_sflux is None else np.ascontiguousarray(cot_sflux, dtype=np.float64)
    fs, fc, fg = field_ptrs(state), field_ptrs(cot), field_ptrs(grad)
    st = grad_lib().grad_adjoint_host(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(P), lall, ijdim,
                                      ctypes.byref(fs), ctypes.byref(fc), None if cs is None else cs.ctypes.data,
                                      ctypes.byref(fg), gc.ctypes.data, float(dt))
    return grad, gc, st


def dual_vjp(state: dict, gz, gzh, dt: float, cot: dict, cot_sflux, dirs, dcoef=None, P=None):
    """J = sum <cot, outputs> + <cot_sflux, sflux> of one FP64 host step and its
    directional derivatives along `dirs` (list of dicts of input directions) and the
    coefficient directions `dcoef` [len(dirs)][8], by forward-mode dual numbers."""
    from nicam_kokkos_microphysics_amd import _native as N
    from nicam_kokkos_microphysics_amd.fields import field_ptrs
    P = params() if P is None else P
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    lall, ijdim = _shape(state)
    cot = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in cot.items()}
    cs = None if cot_sflux is None else np.ascontiguousarray(cot_sflux, dtype=np.float64)
    fs, fc = field_ptrs(state), field_ptrs(cot)
    out = np.zeros(len(dirs))
    J = 0.0
    for c0 in range(0, len(dirs), GRAD_NDIR):
        chunk = dirs[c0:c0 + GRAD_NDIR]
        arr = (N.FieldPtrs * len(chunk))()
        keep = []
        for i, d in enumerate(chunk):
            dd = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in d.items()}
            keep.append(dd)
            arr[i] = field_ptrs(dd)
        dc = np.zeros((len(chunk), 8)) if dcoef is None else np.ascontiguousarray(dcoef[c0:c0 + GRAD_NDIR],
                                                                                dtype=np.float64)
        jv = ctypes.c_double()
        dj = np.zeros(len(chunk))
        st = grad_lib().grad_dual_vjp(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(P), lall, ijdim,
                                      ctypes.byref(fs), float(dt), ctypes.byref(fc),
                                      None if cs is None else cs.ctypes.data, len(chunk), arr, dc.ctypes.data,
                                      ctypes.addressof(jv), dj.ctypes.data)
        if st < 0:
            raise RuntimeError(f"grad_dual_vjp status {st}")
        J = jv.value
        out[c0:c0 + len(chunk)] = dj
    return J, out
