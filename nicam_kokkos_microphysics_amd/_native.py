"""ctypes bindings of the C ABI in ``include/nsw6.h``.

The struct layouts are read from the header itself, so the Python side can never
drift from the C side.  Loading fails loudly when the HIP library has not been
built: there is no CPU fallback for the step.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HEADER = os.path.join(REPO, "include", "nsw6.h")
LIB_NSW6 = os.path.join(HERE, "libnsw6.so")
LIB_SYNTH = os.path.join(HERE, "libnsw6_synth.so")

KALL, KMIN, KMAX, NQMAX = 96, 1, 94, 6
NSW6_OK, NSW6_NOT_CONVERGED, NSW6_NAN_DETECTED = 0, 1, 2
NSW6_BAD_ARG, NSW6_HIP_ERROR, NSW6_NO_DEVICE, NSW6_UNSUPPORTED = 3, 4, 5, 6
NSW6_IO_ERROR = 7
NSW6_MAX_DIV_NUM = 1000
NSW6_NOALLOC_MAX_COLUMNS, NSW6_NOALLOC_MAX_REGIONS = 16900, 64
NSW6_VERIFY_FIELDS = 12
NSW6_FP64, NSW6_MIXED = 0, 1

# The 8 tunable coefficients (order of NSW6_NCOEF in nsw6.h).
TUNABLE_COEFFS = ("kk_aut_rate", "qicrt_saut", "Esi", "Esr", "Cr", "Cs", "Cg", "icenuc_coef")
NCOEF = len(TUNABLE_COEFFS)


def _struct_body(text: str, name: str) -> str:
    m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S)
    if not m:
        raise RuntimeError(f"struct {name} not found in {HEADER}")
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    return body


def _fields_of(body: str, ctype, is_pointer=False):
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        decl = re.sub(r"^(double|int32_t)\s*", "", decl)
        for item in decl.split(","):
            item = item.strip().lstrip("*").strip()
            m = re.match(r"(\w+)(?:\[(\w+)\])?$", item)
            if not m:
                raise RuntimeError(f"cannot parse declaration '{item}'")
            nm, dim = m.group(1), m.group(2)
            if dim is None:
                out.append((nm, ctype))
            else:
                n = KALL if dim == "NSW6_KALL" else int(dim)
                out.append((nm, ctype * n))
    return out


with open(HEADER) as _fh:
    _HTEXT = _fh.read()


class Params(ctypes.Structure):
    _fields_ = _fields_of(_struct_body(_HTEXT, "nsw6_params"), ctypes.c_double)


class Grid(ctypes.Structure):
    _fields_ = _fields_of(_struct_body(_HTEXT, "nsw6_grid"), ctypes.c_double)


class FieldPtrs(ctypes.Structure):
    _fields_ = _fields_of(_struct_body(_HTEXT, "nsw6_fields"), ctypes.c_void_p)


class Report(ctypes.Structure):
    _fields_ = _fields_of(_struct_body(_HTEXT, "nsw6_report"), ctypes.c_int32)


FIELD_NAMES = tuple(n for n, _ in FieldPtrs._fields_)

# ctypes declarations of every extern "C" entry point of nsw6.h
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
SIGNATURES = {
    "nsw6_params_default": (None, [ctypes.POINTER(Params)]),
    "nsw6_grid_setup": (None, [_P, _P, ctypes.POINTER(Grid)]),
    "nsw6_init": (_I, [ctypes.POINTER(Grid), ctypes.POINTER(Params), ctypes.c_int64, ctypes.POINTER(_P)]),
    "nsw6_set_params": (_I, [_P, ctypes.POINTER(Params)]),
    "nsw6_set_precision": (_I, [_P, _I]),
    "nsw6_set_div_num": (_I, [_P, _I]),
    "nsw6_mp_driver_fields": (_I, [_P, _I, _I, _I, ctypes.POINTER(FieldPtrs), _D, _D, _P]),
    "nsw6_mp_driver": (_I, [_P, _I, _I, _I] + [_P] * 46 + [_D, _D] + [_P] * 4 + [_P]),
    "nsw6_mp_driver_noalloc": (_I, [_P, _I, _I, _I, ctypes.POINTER(FieldPtrs), _D, _D, _P]),
    "nsw6_mp_driver_adj": (_I, [_P, _I, _I, _I, ctypes.POINTER(FieldPtrs), ctypes.POINTER(FieldPtrs), _P,
                                ctypes.POINTER(FieldPtrs), _P, _D, _D, _P]),
    "nsw6_last_status": (_I, [_P, ctypes.POINTER(Report)]),
    "nsw6_kcell_bounds": (_I, [_P, _I, _P, ctypes.c_int64]),
    "nsw6_timing_enable": (_I, [_P, _I]),
    "nsw6_timing_read": (_I, [_P, _I, ctypes.POINTER(_D), ctypes.POINTER(ctypes.c_int64)]),
    "nsw6_timing_reset": (_I, [_P]),
    "nsw6_finalize": (None, [_P]),
    "nsw6_status_string": (ctypes.c_char_p, [_I]),
    "nsw6_snapshot_read": (_I, [ctypes.c_char_p, _I, _I, ctypes.POINTER(FieldPtrs), _P, _P]),
    "nsw6_snapshot_write": (_I, [ctypes.c_char_p, _I, _I, ctypes.POINTER(FieldPtrs), _P, _P]),
    "nsw6_verify_read": (_I, [ctypes.c_char_p, _I, _I, _P]),
    "nsw6_verify_write": (_I, [ctypes.c_char_p, _I, _I, ctypes.POINTER(FieldPtrs)]),
}

_lib = None
_synth = None


def lib() -> ctypes.CDLL:
    """The HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_NSW6):
            raise RuntimeError(
                f"{LIB_NSW6} is missing: build it with `python -m nicam_kokkos_microphysics_amd.build_lib` "
                "(there is no CPU fallback for the NSW6 step)")
        # torch brings its own copy of the HIP runtime (same soname): load it first so
        # that libnsw6.so binds to the runtime instance that owns torch's streams and
        # allocations
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_NSW6)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class SynthConfig(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("rh_lo", _D), ("rh_hi", _D), ("frac_supersat", _D),
                ("frac_towers", _D), ("tsfc", _D), ("frac_rainshafts", _D)]


def synth_lib() -> ctypes.CDLL:
    global _synth
    if _synth is None:
        if not os.path.exists(LIB_SYNTH):
            raise RuntimeError(f"{LIB_SYNTH} is missing: build it with `python -m nicam_kokkos_microphysics_amd.build_lib`")
        S = ctypes.CDLL(LIB_SYNTH)
        S.nsw6_synth_default_config.argtypes = [ctypes.POINTER(SynthConfig)]
        S.nsw6_synth_default_config.restype = None
        S.nsw6_synth_grid.argtypes = [_P, _P]
        S.nsw6_synth_grid.restype = None
        S.nsw6_synth_fill.argtypes = [ctypes.POINTER(SynthConfig), _P, _P, _I, _I, _I, _I, ctypes.POINTER(FieldPtrs)]
        S.nsw6_synth_fill.restype = _I
        _synth = S
    return _synth


def status_string(code: int) -> str:
    return lib().nsw6_status_string(int(code)).decode()
