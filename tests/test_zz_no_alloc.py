"""The allocation-free entry point nsw6_mp_driver_noalloc, through the C ABI, against
nsw6_mp_driver_fields on the same inputs.

Both run the same kernels with the same arithmetic, so the outputs must be equal bit
for bit (stronger than the FP64 bar of 1e-10 of each field's max-abs).  Checked:
two calls in a row, the second on a smaller input (the scratch keeps the first
call's contents); MP_DIV_NUM = 2 (the sub-step sums live in the scratch too); the
cloud-top variant; the device report; calls beyond the stated limits are refused.
"""
import ctypes

import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from nicam_kokkos_microphysics_amd import fields as F

pytestmark = pytest.mark.gpu

DT = 15.0


def _noalloc(drv, dev, dt=DT):
    import torch
    lall, ijdim = int(dev["rhog"].shape[0]), int(dev["rhog"].shape[-1])
    fp = F.field_ptrs(dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return N.lib().nsw6_mp_driver_noalloc(drv._ctx, 0, lall, ijdim, ctypes.byref(fp), float(dt), 0.0, stream)


def _both(drv, st):
    """(outputs of nsw6_mp_driver_fields, its report, outputs of
    nsw6_mp_driver_noalloc, its report) from the same host state."""
    import torch
    a, b = F.to_device(st), F.to_device(st)
    lall, ijdim = st["rhog"].shape[0], st["rhog"].shape[-1]
    print(f"fields vs noalloc: lall={lall} ijdim={ijdim}", flush=True)
    drv.step(a, DT)
    rep_a = drv.report()
    assert _noalloc(drv, b) == N.NSW6_OK
    torch.cuda.synchronize()
    rep_b = drv.report()
    return F.to_host(a), rep_a, F.to_host(b), rep_b


def _assert_equal(x, y):
    for k in F.OUTPUTS:
        assert np.array_equal(x[k], y[k]), (k, float(np.abs(x[k] - y[k]).max()))


def _report_tuple(r):
    return (r.status, r.n_not_converged, r.n_nan, r.first_l, r.first_k, r.first_ij, r.n_general_regions)


def test_noalloc_equals_fields_entry_two_calls_shrinking():
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st1 = F.synth_state(2, 5000, seed=71, frac_supersat=0.1)
    _, _, st2 = F.synth_state(1, 3001, seed=72, frac_towers=0.2)
    drv = NSW6(gz, gzh, max_columns=10000)
    try:
        for st in (st1, st2):
            x, ra, y, rb = _both(drv, st)
            _assert_equal(x, y)
            assert _report_tuple(ra) == _report_tuple(rb)
    finally:
        drv.close()


def test_noalloc_subcycled_and_cloud_top():
    from nicam_kokkos_microphysics_amd.driver import NSW6, default_params
    gz, gzh, st = F.synth_state(3, 1500, seed=73, frac_towers=0.3)
    P = default_params()
    P.cloudtop_intended = 1.0
    drv = NSW6(gz, gzh, params=P, max_columns=4500, div_num=2)
    try:
        x, ra, y, rb = _both(drv, st)
        _assert_equal(x, y)
        assert _report_tuple(ra) == _report_tuple(rb)
        # a second, smaller call after the sub-cycled one
        drv.set_div_num(1)
        _, _, st_small = F.synth_state(1, 700, seed=74, frac_towers=0.3)
        x, _, y, _ = _both(drv, st_small)
        _assert_equal(x, y)
    finally:
        drv.close()


def test_noalloc_refuses_calls_beyond_its_limits():
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = F.synth_state(1, N.NSW6_NOALLOC_MAX_COLUMNS + 1, seed=75)
    drv = NSW6(gz, gzh, max_columns=64)
    try:
        dev = F.to_device(st)
        assert _noalloc(drv, dev) == N.NSW6_BAD_ARG
        _, _, st65 = F.synth_state(N.NSW6_NOALLOC_MAX_REGIONS + 1, 2, seed=76)
        assert _noalloc(drv, F.to_device(st65)) == N.NSW6_BAD_ARG
        torch.cuda.synchronize()
    finally:
        drv.close()
