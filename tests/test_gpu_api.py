"""Boundary, error-report and binding checks of the C ABI on the GPU.

* The flat nsw6_mp_driver, called with the reference's argument order
  (include/mod_mp_driver.h:71-124) through ctypes, is bit for bit the
  nsw6_mp_driver_fields step on the same inputs (a swapped pointer among its 52 array
  arguments would show).
* A NaN temperature at one point (l, k, ij): as in the reference, whose NaN guard
  (src/mod_satadjust.cpp:753,880) sits inside the Newton loop, which a NaN input never
  enters (`qsum - qsat > EPS` is false), the step reports OK; the non-finite values stay
  in that column, in the same places as the oracle's, and every other column agrees
  with the oracle within 1e-10 of each field's max-abs.
* The device report, forced from the host: with the Newton's stopping tolerance set to
  0 every saturation-adjustment point runs to itelim = 100; the step reports
  NSW6_NOT_CONVERGED with the oracle's count and the oracle's first failing point
  (lexicographic (l, k, ij)), and its state still agrees with the oracle's.
* MPStep refuses a coefficient tensor that differs from the context's coefficients.
* The Python binding refuses host arrays before any pointer reaches the library.
"""
import ctypes

import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from nicam_kokkos_microphysics_amd import fields as F
from nicam_kokkos_microphysics_amd.check import compare_states

DT = 15.0
COMPARE = [k for k in F.OUTPUTS if k not in ("rhogq", "q")] + \
          [f"rhogq:{t}" for t in F.TRACERS] + [f"q:{t}" for t in F.TRACERS]


@pytest.mark.gpu
def test_flat_entry_point_equals_fields_entry_point():
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = F.synth_state(2, 700, seed=81, frac_towers=0.2)
    drv = NSW6(gz, gzh, max_columns=1400)
    try:
        a, b = F.to_device(st), F.to_device(st)
        drv.step(a, DT)
        ptr = {k: (ctypes.c_void_p(v.data_ptr()) if v is not None else None) for k, v in b.items()}
        order = ["rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rhogq", "vx", "vy", "vz", "w", "unccn",
                 "rho", "tem", "pre", "q", "qd", "precip", "ISO1_precip", "ISO2_precip", "precip_rhoe",
                 "precip_lh_heat", "precip_rhophi", "precip_rhokin", "re_liquid", "re_solid", "re_cld", "rctop",
                 "rwtop", "tctop", "frhoge_af", "frhogqv_af", "frhoge_rad", "rhogqke", "gsgam2", "gsgam2h",
                 "gam2", "gam2h", "ix", "iy", "iz", "jx", "jy", "jz", "z", "zh"]
        tail = ["GDCLW", "GDCFRC", "GPREC", "CBMFX"]
        assert order + tail == [n for n in N.FIELD_NAMES]   # the header's order is the reference's
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = N.lib().nsw6_mp_driver(drv._ctx, 2, 700, 0, *[ptr[n] for n in order], DT, 0.0,
                                    *[ptr[n] for n in tail], stream)
        assert rc == N.NSW6_OK
        torch.cuda.synchronize()
        for k in F.OUTPUTS:
            assert torch.equal(a[k], b[k]), k
    finally:
        drv.close()


@pytest.mark.gpu
def test_nan_temperature_stays_in_its_column():
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = F.synth_state(2, 300, seed=21)
    st["tem"][1, 40, 123] = np.nan
    ref = F.copy_state(st)
    rep_o = O.mp_driver(ref, gz, gzh, DT)
    drv = NSW6(gz, gzh, max_columns=600)
    try:
        dev = F.to_device(st)
        print("step with tem[1, 40, 123] = NaN", flush=True)
        drv.step(dev, DT, check=False)
        torch.cuda.synchronize()
        rep = drv.report()
        out = F.to_host(dev)
    finally:
        drv.close()
    assert rep_o.status == N.NSW6_OK and rep.status == N.NSW6_OK and rep.n_nan == 0
    bad_o, bad_g = ~np.isfinite(ref["tem"]), ~np.isfinite(out["tem"])
    assert bad_o.any() and np.array_equal(bad_o, bad_g)
    assert set(map(tuple, np.argwhere(bad_g)[:, [0, 2]].tolist())) == {(1, 123)}
    keep = np.ones(300, dtype=bool)
    keep[123] = False
    sub = {k: v[..., keep] for k, v in out.items()}
    sref = {k: v[..., keep] for k, v in ref.items()}
    bad = compare_states(sub, sref, COMPARE, 1e-10)
    assert not bad, bad


@pytest.mark.gpu
def test_forced_non_convergence_is_reported_like_the_oracle():
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6, default_params
    gz, gzh, st = F.synth_state(3, 500, seed=22, frac_supersat=0.2)
    P_o, P_g = O.params(), default_params()
    assert P_o.satadjust_dtemp_criteria == P_g.satadjust_dtemp_criteria == 1e-8
    P_o.satadjust_dtemp_criteria = P_g.satadjust_dtemp_criteria = 0.0
    ref = F.copy_state(st)
    rep_o = O.mp_driver(ref, gz, gzh, DT, P=P_o)
    drv = NSW6(gz, gzh, params=P_g, max_columns=1500)
    try:
        dev = F.to_device(st)
        print("step with the Newton tolerance 0", flush=True)
        drv.step(dev, DT, check=False)
        torch.cuda.synchronize()
        rep = drv.report()
        out = F.to_host(dev)
    finally:
        drv.close()
    assert rep_o.status == N.NSW6_NOT_CONVERGED and rep_o.n_not_converged > 100
    assert rep.status == N.NSW6_NOT_CONVERGED
    assert (rep.n_not_converged, rep.n_nan) == (rep_o.n_not_converged, rep_o.n_nan)
    assert (rep.first_l, rep.first_k, rep.first_ij) == (rep_o.first_l, rep_o.first_k, rep_o.first_ij)
    bad = compare_states(out, ref, COMPARE, 1e-10)
    assert not bad, bad


@pytest.mark.gpu
def test_mpstep_refuses_coefficients_other_than_the_contexts():
    import torch
    from nicam_kokkos_microphysics_amd.autograd import mp_step
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = F.synth_state(1, 128, seed=23)
    drv = NSW6(gz, gzh, max_columns=128)
    try:
        dev = F.to_device(st)
        p = drv.params
        coef = torch.tensor([getattr(p, n) for n in N.TUNABLE_COEFFS], dtype=torch.float64, device="cuda")
        out = mp_step(drv, dev, DT, coef)          # the context's values: accepted
        assert torch.isfinite(out["rhog"]).all()
        wrong = coef.clone()
        wrong[4] *= 1.01                           # Cr
        with pytest.raises(ValueError):
            mp_step(drv, dev, DT, wrong)
    finally:
        drv.close()


def test_binding_refuses_host_arrays():
    """No GPU needed: the check runs before any call into the library."""
    import torch
    from nicam_kokkos_microphysics_amd.driver import check_device_tensors
    gz, gzh, st = F.synth_state(1, 16, seed=24)
    with pytest.raises(TypeError):
        check_device_tensors(st, 1, 16, 0)                                   # numpy arrays
    with pytest.raises(TypeError):
        check_device_tensors({k: torch.from_numpy(v) for k, v in st.items()}, 1, 16, 0)   # CPU tensors


@pytest.mark.gpu
def test_timing_slots_partition_the_step():
    """The step timer (event before k_rates to event after k_finish) is the sum of the
    per-kernel timers, which are consecutive event intervals on the launch stream; the
    general sweep has its own slot."""
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = F.synth_state(2, 2000, seed=25)
    drv = NSW6(gz, gzh, max_columns=4000)
    try:
        dev = F.to_device(st)
        drv.step(dev, DT)
        torch.cuda.synchronize()
        drv.timing(True)
        drv.timing_reset()
        for _ in range(3):
            drv.step(dev, DT, check=False)
        torch.cuda.synchronize()
        total, n = drv.timing_read(0)
        parts = [drv.timing_read(i)[0] for i in (1, 2, 3, 8, 4)]
        assert n == 3 and drv.timing_read(8)[1] == 3
        assert all(p >= 0.0 for p in parts) and total > 0.0
        assert abs(sum(parts) - total) <= 1e-3 + 1e-3 * total, (parts, total)
        drv.timing(False)
    finally:
        drv.close()
