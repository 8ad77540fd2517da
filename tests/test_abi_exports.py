"""The C ABI library loads (no GPU needed) and exports every function that
include/nsw6.h declares, with the ctypes signatures of the Python binding covering
exactly that set."""
import ctypes
import os
import re

from nicam_kokkos_microphysics_amd import _native as N

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "nsw6.h")


def declared_functions():
    with open(HEADER) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^(?:int|void|const char \*)\s*(nsw6_\w+)\s*\(", text, flags=re.M))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("nsw6_init", "nsw6_mp_driver", "nsw6_mp_driver_fields", "nsw6_mp_driver_adj",
                 "nsw6_last_status", "nsw6_finalize"):
        assert must in names


def test_library_exports_every_declared_function():
    L = ctypes.CDLL(N.LIB_NSW6)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header():
    assert set(N.SIGNATURES) == declared_functions()
    L = N.lib()
    for name in N.SIGNATURES:
        assert getattr(L, name) is not None


def test_status_strings():
    assert N.status_string(N.NSW6_OK) == "ok"
    assert "reverse mode" in N.status_string(N.NSW6_UNSUPPORTED)
    assert N.status_string(99) == "unknown status"
