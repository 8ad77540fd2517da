"""Host-side mirror of the reference's microphysics driver interface.

``mp_init("NSW6", gz, gzh)`` plays the role of MP_DRIVER::mp_init together with
GRD_Setup / CVW_CPW_Setup / SATURATION_Setup (src/main.cpp:179-188,
src/mod_mp_driver.cpp:11-24) and returns an :class:`NSW6` context; its
:meth:`NSW6.mp_driver` takes the arguments of MP_DRIVER::mp_driver
(include/mod_mp_driver.h:71-124) in the same order, as float64 device tensors
with a leading region axis, and runs the HIP step.

Errors: an unknown MP_TYPE raises (the reference prints and exits,
src/mod_debug.cpp:17-21); Newton non-convergence or a NaN temperature raise
:class:`NSW6Error` (the reference calls Kokkos::abort,
src/mod_satadjust.cpp:883-891).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .fields import ADJ_INPUTS, field_ptrs, shape_of


class NSW6Error(RuntimeError):
    def __init__(self, code: int, report=None):
        msg = N.status_string(code)
        if report is not None:
            msg += (f" (not converged: {report.n_not_converged}, NaN: {report.n_nan}, first point "
                    f"l={report.first_l} k={report.first_k} ij={report.first_ij})")
        super().__init__(msg)
        self.code = code
        self.report = report


def default_params() -> N.Params:
    p = N.Params()
    N.lib().nsw6_params_default(ctypes.byref(p))
    return p


def grid_setup(gz, gzh) -> N.Grid:
    g = N.Grid()
    gz = np.ascontiguousarray(gz, dtype=np.float64)
    gzh = np.ascontiguousarray(gzh, dtype=np.float64)
    N.lib().nsw6_grid_setup(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(g))
    return g


def check_device_tensors(arrays: dict, lall: int, ijdim: int, device: int, what: str = "state"):
    """Every non-None entry must be a contiguous float64 tensor on HIP device `device`
    with its mp_driver shape: the C ABI hands these pointers to kernels, which would
    fault on a host pointer and read past a short or narrower array."""
    import torch
    for k, v in arrays.items():
        if v is None:
            continue
        if not isinstance(v, torch.Tensor) or not v.is_cuda:
            raise TypeError(f"{what}[{k}]: must be a HIP device tensor, not {type(v).__name__}"
                            + ("" if not isinstance(v, torch.Tensor) else " on the CPU"))
        if v.device.index != device:
            raise ValueError(f"{what}[{k}]: on cuda:{v.device.index}, the context is on cuda:{device}")
        if v.dtype != torch.float64 or not v.is_contiguous():
            raise ValueError(f"{what}[{k}]: must be contiguous float64")
        if tuple(v.shape) != shape_of(k, lall, ijdim):
            raise ValueError(f"{what}[{k}]: shape {tuple(v.shape)} != {shape_of(k, lall, ijdim)}")


def _check_vector(v, shape, device, what):
    import torch
    if not isinstance(v, torch.Tensor) or not v.is_cuda or v.device.index != device:
        raise TypeError(f"{what}: must be a HIP device tensor on cuda:{device}")
    if v.dtype != torch.float64 or not v.is_contiguous() or tuple(v.shape) != tuple(shape):
        raise ValueError(f"{what}: must be contiguous float64 of shape {tuple(shape)}")


def _stream_handle(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


class NSW6:
    """One HIP context of the NSW6 step (workspace sized for ``max_columns``)."""

    def __init__(self, gz, gzh, params: N.Params | None = None, max_columns: int = 169000,
                 precision: int = N.NSW6_FP64, div_num: int = 1):
        self._L = N.lib()
        self.grid = grid_setup(gz, gzh)
        self.params = params if params is not None else default_params()
        self.max_columns = int(max_columns)
        ctx = ctypes.c_void_p()
        rc = self._L.nsw6_init(ctypes.byref(self.grid), ctypes.byref(self.params), self.max_columns,
                               ctypes.byref(ctx))
        if rc != 0:
            raise NSW6Error(rc)
        self._ctx = ctx
        import torch
        self.device = torch.cuda.current_device()   # nsw6_init took the current device
        self.set_precision(precision)
        self.set_div_num(div_num)

    # -- configuration ------------------------------------------------------------
    def set_params(self, params: N.Params):
        rc = self._L.nsw6_set_params(self._ctx, ctypes.byref(params))
        if rc != 0:
            raise NSW6Error(rc)
        self.params = params

    def set_precision(self, precision: int):
        rc = self._L.nsw6_set_precision(self._ctx, int(precision))
        if rc != 0:
            raise NSW6Error(rc)
        self.precision = precision

    def set_div_num(self, div_num: int):
        """MP_DIV_NUM sub-steps per call (include/problemsize.h:175; the reference fixes 1)."""
        rc = self._L.nsw6_set_div_num(self._ctx, int(div_num))
        if rc != 0:
            raise NSW6Error(rc)
        self.div_num = int(div_num)

    # -- the step -----------------------------------------------------------------
    def step(self, state: dict, dt: float, ct: float = 0.0, *, l_region: int = 0, stream=None,
             check: bool = True):
        """Run mp_driver on a dict of device tensors (keys = the 50 argument names)."""
        rhog = state["rhog"]
        lall, ijdim = int(rhog.shape[0]), int(rhog.shape[-1])
        check_device_tensors(state, lall, ijdim, self.device)
        fp = field_ptrs(state)
        rc = self._L.nsw6_mp_driver_fields(self._ctx, l_region, lall, ijdim, ctypes.byref(fp), float(dt),
                                           float(ct), _stream_handle(stream))
        if rc != 0:
            raise NSW6Error(rc)
        if check:
            self.check()

    def mp_driver(self, l_region, rhog, rhogvx, rhogvy, rhogvz, rhogw, rhoge, rhogq, vx, vy, vz, w, unccn, rho,
                  tem, pre, q, qd, precip, ISO1_precip, ISO2_precip, precip_rhoe, precip_lh_heat,
                  precip_rhophi, precip_rhokin, re_liquid, re_solid, re_cld, rctop, rwtop, tctop, frhoge_af,
                  frhogqv_af, frhoge_rad, rhogqke, gsgam2, gsgam2h, gam2, gam2h, ix, iy, iz, jx, jy, jz, z, zh,
                  dt, ct, GDCLW, GDCFRC, GPREC, CBMFX, *, stream=None, check=True):
        """Argument list of MP_DRIVER::mp_driver (include/mod_mp_driver.h:71-124)."""
        args = dict(locals())
        state = {n: args[n] for n in N.FIELD_NAMES}
        self.step(state, dt, ct, l_region=l_region, stream=stream, check=check)

    def adjoint(self, state: dict, cot: dict, dt: float, *, cot_sflux=None, grad: dict | None = None,
                grad_coef=None, ct: float = 0.0, l_region: int = 0, stream=None, check: bool = True):
        """Reverse mode of one FP64 step (nsw6_mp_driver_adj): `state` holds the step's
        input (device tensors, not modified), `cot` the cotangents of its outputs (keys
        of fields.ADJ_OUTPUTS; missing = 0), `cot_sflux` [l][NT][ij] those of the
        surface mass fluxes of the falling tracers (NT = 3: QR, QS, QG; under
        four_water NT = 4: QR, QI, QS, QG).  Returns (grad, grad_coef): the cotangents of
        the inputs (keys of fields.ADJ_INPUTS, and unccn when the state holds it: nonzero
        under OPT_INDIR only) and of the 8 tunable coefficients."""
        import torch
        rhog = state["rhog"]
        lall, ijdim = int(rhog.shape[0]), int(rhog.shape[-1])
        check_device_tensors(state, lall, ijdim, self.device, "state")
        check_device_tensors(cot, lall, ijdim, self.device, "cot")
        if cot_sflux is not None:
            nt = 4 if self.params.four_water != 0.0 else 3
            _check_vector(cot_sflux, (lall, nt, ijdim), self.device, "cot_sflux")
        if grad is None:
            grad = {k: torch.empty_like(state[k]) for k in ADJ_INPUTS}
            if state.get("unccn") is not None:
                grad["unccn"] = torch.empty_like(state["unccn"])
        check_device_tensors(grad, lall, ijdim, self.device, "grad")
        if grad_coef is None:
            grad_coef = torch.empty(N.NCOEF, dtype=torch.float64, device=rhog.device)
        _check_vector(grad_coef, (N.NCOEF,), self.device, "grad_coef")
        fs, fc, fg = field_ptrs(state), field_ptrs(cot), field_ptrs(grad)
        cs = None if cot_sflux is None else ctypes.c_void_p(cot_sflux.data_ptr())
        rc = self._L.nsw6_mp_driver_adj(self._ctx, l_region, lall, ijdim, ctypes.byref(fs), ctypes.byref(fc), cs,
                                        ctypes.byref(fg), ctypes.c_void_p(grad_coef.data_ptr()), float(dt),
                                        float(ct), _stream_handle(stream))
        if rc != 0:
            raise NSW6Error(rc)
        if check:
            self.check()
        return grad, grad_coef

    def check(self) -> N.Report:
        rep = N.Report()
        rc = self._L.nsw6_last_status(self._ctx, ctypes.byref(rep))
        if rc != 0:
            raise NSW6Error(rc, rep)
        return rep

    def kcell_bounds(self, lall: int, nt: int = 3) -> np.ndarray:
        """kcell_min/kcell_max of the last step (nsw6_kcell_bounds) as int32
        [l][nt][2 (full, half)][96][2 (kcell_min, kcell_max)]."""
        out = np.empty((lall, nt, 2, N.KALL, 2), dtype=np.int32)
        rc = self._L.nsw6_kcell_bounds(self._ctx, int(lall), out.ctypes.data, out.size)
        if rc != 0:
            raise NSW6Error(rc)
        return out

    def report(self) -> N.Report:
        rep = N.Report()
        self._L.nsw6_last_status(self._ctx, ctypes.byref(rep))
        return rep

    # -- timing (HIP events on the launch stream) ----------------------------------
    def timing(self, on: bool = True):
        self._L.nsw6_timing_enable(self._ctx, int(on))

    def timing_reset(self):
        self._L.nsw6_timing_reset(self._ctx)

    def timing_read(self, which: int = 0):
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        rc = self._L.nsw6_timing_read(self._ctx, which, ctypes.byref(ms), ctypes.byref(cnt))
        if rc != 0:
            raise NSW6Error(rc)
        return ms.value, cnt.value

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.nsw6_finalize(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mp_init(MP_TYPE: str, gz, gzh, params: N.Params | None = None, max_columns: int = 169000) -> NSW6:
    """MP_DRIVER::mp_init (src/mod_mp_driver.cpp:11-24): only "NSW6" is supported."""
    if MP_TYPE != "NSW6":
        raise ValueError(f"mp_init: NOT appropriate type. Type: {MP_TYPE}")
    return NSW6(gz, gzh, params=params, max_columns=max_columns)
