"""Argument and error handling of the C ABI (include/nsw6.h): bad arguments are reported
as NSW6_BAD_ARG before anything is launched, no device is NSW6_NO_DEVICE, and every
status has a message.  The reference stops the program instead (src/mod_debug.cpp:17-21).

* CPU: the checks that run before any device call (null context or grid, column
  count limits, precision and sub-step ranges, snapshot arguments), and
  NSW6_NO_DEVICE where the machine has no GPU.
* GPU: a step with too many columns, dt <= 0, a missing required array, or OPT_INDIR
  without unccn (step or reverse mode) is refused and leaves the state untouched.
"""
import ctypes

import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from nicam_kokkos_microphysics_amd import fields as F

KALL, NQ = 96, 6


def _grid():
    from nicam_kokkos_microphysics_amd.driver import grid_setup
    gz, gzh = F.synth_grid()
    return grid_setup(gz, gzh)


def test_init_argument_checks():
    L = N.lib()
    g = _grid()
    ctx = ctypes.c_void_p()
    assert L.nsw6_init(None, None, 100, ctypes.byref(ctx)) == N.NSW6_BAD_ARG
    assert L.nsw6_init(ctypes.byref(g), None, 0, ctypes.byref(ctx)) == N.NSW6_BAD_ARG
    assert L.nsw6_init(ctypes.byref(g), None, -5, ctypes.byref(ctx)) == N.NSW6_BAD_ARG
    too_many = (1 << 31) // (KALL * NQ) + 1
    assert L.nsw6_init(ctypes.byref(g), None, too_many, ctypes.byref(ctx)) == N.NSW6_BAD_ARG
    assert L.nsw6_init(ctypes.byref(g), None, 100, None) == N.NSW6_BAD_ARG


def test_no_device_is_reported():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = N.lib()
    g = _grid()
    ctx = ctypes.c_void_p()
    assert L.nsw6_init(ctypes.byref(g), None, 100, ctypes.byref(ctx)) == N.NSW6_NO_DEVICE
    assert not ctx.value


def test_null_context_calls_are_argument_errors():
    L = N.lib()
    fp = N.FieldPtrs()
    rep = N.Report()
    P = N.Params()
    L.nsw6_params_default(ctypes.byref(P))
    assert L.nsw6_set_precision(None, N.NSW6_FP64) == N.NSW6_BAD_ARG
    assert L.nsw6_set_div_num(None, 2) == N.NSW6_BAD_ARG
    assert L.nsw6_set_params(None, ctypes.byref(P)) == N.NSW6_BAD_ARG
    assert L.nsw6_mp_driver_fields(None, 0, 1, 16, ctypes.byref(fp), 15.0, 0.0, None) == N.NSW6_BAD_ARG
    assert L.nsw6_mp_driver_adj(None, 0, 1, 16, ctypes.byref(fp), ctypes.byref(fp), None, ctypes.byref(fp),
                                None, 15.0, 0.0, None) == N.NSW6_BAD_ARG
    assert L.nsw6_last_status(None, ctypes.byref(rep)) == N.NSW6_BAD_ARG
    L.nsw6_finalize(None)     # no-op


def test_snapshot_argument_checks(tmp_path):
    L = N.lib()
    fp = N.FieldPtrs()
    assert L.nsw6_snapshot_read(None, 1, 16, ctypes.byref(fp), None, None) == N.NSW6_BAD_ARG
    assert L.nsw6_snapshot_read(str(tmp_path).encode(), 0, 16, ctypes.byref(fp), None, None) == N.NSW6_BAD_ARG
    assert L.nsw6_snapshot_write(str(tmp_path).encode(), 1, 0, ctypes.byref(fp), None, None) == N.NSW6_BAD_ARG
    assert L.nsw6_verify_read(str(tmp_path).encode(), 1, 16, None) == N.NSW6_BAD_ARG
    assert L.nsw6_verify_write(str(tmp_path).encode(), 1, 16, ctypes.byref(fp)) == N.NSW6_BAD_ARG  # no arrays


def test_every_status_has_a_message():
    msgs = [N.status_string(c) for c in range(8)]
    assert len(set(msgs)) == 8 and all(m and m != "unknown status" for m in msgs)
    assert N.status_string(99) == "unknown status"


@pytest.mark.gpu
def test_gpu_refused_steps_leave_the_state_untouched():
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6, NSW6Error, default_params
    gz, gzh, st = F.synth_state(2, 64, seed=3)
    drv = NSW6(gz, gzh, max_columns=64)          # one region's worth only
    dev = F.to_device(st)
    before = {k: v.clone() for k, v in dev.items() if v is not None}

    def refused(state, dt=15.0):
        with pytest.raises(NSW6Error) as e:
            drv.step(state, dt)
        assert e.value.code == N.NSW6_BAD_ARG

    refused(dev)                                  # 128 columns > max_columns
    one = {k: (None if v is None else v[:1]) for k, v in dev.items()}
    refused(one, dt=0.0)
    refused(one, dt=-15.0)
    refused(dict(one, tem=None))                  # a required array missing
    P = default_params()
    P.opt_indir = 1.0
    drv.set_params(P)
    refused(dict(one, unccn=None))                # OPT_INDIR reads unccn
    with pytest.raises(NSW6Error) as e:           # so does its reverse mode
        drv.adjoint(dict(one, unccn=None), {"precip": one["precip"]}, 15.0)
    assert e.value.code == N.NSW6_BAD_ARG
    torch.cuda.synchronize()
    for k, v in before.items():
        assert torch.equal(dev[k], v), k
    drv.close()
