"""precip_transport_type = "4WATER" (src/mod_mp_nsw6.cpp:127, 593-600): QI sediments
with the other precipitating tracers, at Vt(QI) = Vti (src/mod_mp_nsw6.cpp:1073-1075),
or at the constant -CONST_Vti under sw_constVti = 1 (:399-407).

* CPU (oracle):
  - Vti at a point is -3.29 (rho qi)^0.16 for qi > 1e-8 and 0 below; the constant
    switch gives -CONST_Vti;
  - 4WATER with a constant zero QI fall speed is the 3WATER step, bit for bit (the
    flux machinery is the one the 3WATER pins cover; only the tracer set changes);
  - with QI falling, every column's total water changes by exactly the surface
    precipitation of the step (rain and solid rows, the solid row now carrying QI),
    to round-off, and the step differs from 3WATER.
* GPU: parity with the oracle within 1e-10 of max-abs for 4WATER with the QI fall
  speed of the scheme, with a constant fall speed, and with rain fast enough for the
  general sweep.
* Reverse mode (QI's flux, its fall speed Vti and its surface flux differentiated):
  the host passes against forward-mode dual numbers (1e-13 of the absolute terms,
  random directions and one per input field) for the same three cases, and the GPU
  against the host passes and the dual numbers.
"""
import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from nicam_kokkos_microphysics_amd import fields as F
from nicam_kokkos_microphysics_amd.check import compare_states

DT = 15.0
COMPARE = [k for k in F.OUTPUTS if k not in ("rhogq", "q")] + \
          [f"rhogq:{t}" for t in F.TRACERS] + [f"q:{t}" for t in F.TRACERS]


def _state(lall=2, ijdim=64, seed=41, tsfc=268.0):
    # by default a cold surface: ice and snow reach the ground
    return F.synth_state(lall, ijdim, seed=seed, frac_towers=0.3, tsfc=tsfc)


def _with(P, **kw):
    for k, v in kw.items():
        setattr(P, k, v)
    return P


def test_ice_fall_speed_closed_form():
    from oracle import oracle as O
    rho, T, dt = 0.7, 250.0, 15.0
    qsl = O.lib().ora_psat_liq(T) / (rho * 461.46 * T)
    for qi, want in ((3e-4, -3.29 * (rho * 3e-4) ** 0.16), (5e-9, 0.0)):
        q6 = np.array([0.5 * qsl, 0, 0, qi, 0, 0])
        pre = rho * ((1 - qi) * 287.04 + q6[0] * 461.46) * T
        o = O.bigloop_point(O.params(), dt, rho, T, q6, qsl, qsl, pre, rho, 50.0, 1.0, 0.5, 0.5)
        assert abs(o.Vt[3] - want) <= 1e-14 * max(abs(want), 1e-30)
        P = _with(O.params(), sw_constVti=1.0, CONST_Vti=0.8)
        o = O.bigloop_point(P, dt, rho, T, q6, qsl, qsl, pre, rho, 50.0, 1.0, 0.5, 0.5)
        assert o.Vt[3] == -0.8


def test_four_water_with_resting_ice_is_three_water():
    from oracle import oracle as O
    gz, gzh, st = _state()
    a, b = F.copy_state(st), F.copy_state(st)
    assert O.mp_driver(a, gz, gzh, DT).status == 0
    P = _with(O.params(), four_water=1.0, sw_constVti=1.0, CONST_Vti=0.0)
    assert O.mp_driver(b, gz, gzh, DT, P=P).status == 0
    for k in F.OUTPUTS:
        assert np.array_equal(a[k], b[k]), k


def test_falling_ice_closes_the_column_water_budget():
    from oracle import oracle as O
    gz, gzh, st = _state()
    G = O.grid(gz, gzh)
    dz = np.array(G.dgz[:])[1:95]
    w0 = np.einsum("lnki,k->li", st["rhogq"][:, :, 1:95], dz)
    a = F.copy_state(st)
    assert O.mp_driver(a, gz, gzh, DT, P=_with(O.params(), four_water=1.0)).status == 0
    w1 = np.einsum("lnki,k->li", a["rhogq"][:, :, 1:95], dz)
    out = (a["precip"][:, 0] + a["precip"][:, 1]) * DT
    np.testing.assert_allclose(w1 - w0, -out, rtol=0, atol=1e-12 * np.abs(w0).max())
    b = F.copy_state(st)
    assert O.mp_driver(b, gz, gzh, DT).status == 0
    assert np.abs(a["precip"][:, 1] - b["precip"][:, 1]).max() > 0      # QI reaches the ground
    assert not np.array_equal(a["rhogq"][:, 3], b["rhogq"][:, 3])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["scheme", "constant", "fast_rain"])
def test_gpu_four_water_matches_oracle(case):
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6, default_params
    # fast rain needs the warm default surface (the cold one has little rain)
    gz, gzh, st = _state(lall=2, ijdim=900, tsfc=300.0 if case == "fast_rain" else 268.0)
    edit = {"four_water": 1.0}
    if case == "constant":
        edit.update(sw_constVti=1.0, CONST_Vti=1.5)
    if case == "fast_rain":
        edit["Cr"] = 130.0 * 25
    P_ora, P_gpu = _with(O.params(), **edit), _with(default_params(), **edit)
    ref = F.copy_state(st)
    dev = F.to_device(st)
    drv = NSW6(gz, gzh, params=P_gpu, max_columns=2 * 900)
    reps = []
    for _ in range(1 if case == "fast_rain" else 2):
        assert O.mp_driver(ref, gz, gzh, DT, P=P_ora).status == 0
        drv.step(dev, DT)
        reps.append(drv.report())
    torch.cuda.synchronize()
    bad = compare_states(F.to_host(dev), ref, COMPARE, 1e-10)
    assert not bad, bad
    if case == "fast_rain":
        assert reps[0].n_general_regions > 0
    else:
        assert reps[-1].n_general_regions == 0
    drv.close()


def _reverse_edit(case):
    edit = {"four_water": 1.0}
    if case == "constant":
        edit.update(sw_constVti=1.0, CONST_Vti=1.5)
    if case == "fast_rain":
        edit["Cr"] = 130.0 * 25
    return edit


@pytest.mark.parametrize("case", ["scheme", "constant", "fast_rain"])
def test_four_water_reverse_matches_dual_numbers(case):
    from oracle import oracle as O
    from adjoint_helpers import STATE_KEYS, directional, output_cotangents, random_directions
    gz, gzh, st = _state(lall=2, ijdim=24, seed=43, tsfc=300.0 if case == "fast_rain" else 268.0)
    P = _with(O.params(), **_reverse_edit(case))
    cot, cs = output_cotangents(st, gz, gzh, DT, P, 7)
    assert cs.shape[1] == 4
    grad, gc, status = O.adjoint_host(st, gz, gzh, DT, cot, cs, P=P)
    assert status == 0
    dirs, dcoef = random_directions(st, 13, 4)
    rng = np.random.default_rng(4)
    for key in STATE_KEYS:
        dirs.append({key: st[key] * rng.uniform(-1, 1, st[key].shape)})
    dcoef = np.concatenate([dcoef, np.zeros((len(STATE_KEYS), 8))])
    _, dj = O.dual_vjp(st, gz, gzh, DT, cot, cs, dirs, dcoef, P=P)
    for i in range(len(dirs)):
        s, a = directional(grad, gc, dirs[i], dcoef[i])
        assert abs(s - dj[i]) <= 1e-13 * a, (i, s, dj[i], a)
    if case == "fast_rain":
        return          # warm surface: the ice melts before it reaches the ground
    # QI's surface flux reaches the inputs: its cotangent alone gives a nonzero gradient
    cs_i = np.zeros_like(cs)
    cs_i[:, 1] = cs[:, 1]
    g_i, _, status = O.adjoint_host(st, gz, gzh, DT, {}, cs_i, P=P)
    assert status == 0 and np.abs(g_i["rhogq"][:, 3]).max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["scheme", "constant", "fast_rain"])
def test_gpu_four_water_reverse(case):
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6, default_params
    from adjoint_helpers import compare_with_host, output_cotangents
    gz, gzh, st = _state(lall=2, ijdim=300, seed=44, tsfc=300.0 if case == "fast_rain" else 268.0)
    P_ora, P_gpu = _with(O.params(), **_reverse_edit(case)), _with(default_params(), **_reverse_edit(case))
    cot, cs = output_cotangents(st, gz, gzh, DT, P_ora, 7)
    g_host, gc_host, status = O.adjoint_host(st, gz, gzh, DT, cot, cs, P=P_ora)
    assert status == 0
    drv = NSW6(gz, gzh, params=P_gpu, max_columns=600)
    try:
        dcot = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in cot.items()}
        print(f"4WATER adjoint case={case}", flush=True)
        grad, gc = drv.adjoint(F.to_device(st), dcot, DT, cot_sflux=torch.from_numpy(np.ascontiguousarray(cs)).cuda())
        torch.cuda.synchronize()
        rep = drv.report()
    finally:
        drv.close()
    assert rep.status == 0 and (rep.n_general_regions > 0) == (case == "fast_rain")
    n = compare_with_host(st, gz, gzh, DT, P_ora, cot, cs, {k: v.cpu().numpy() for k, v in grad.items()},
                          gc.cpu().numpy(), g_host, gc_host)
    print(f"{n} kink entries", flush=True)


@pytest.mark.gpu
def test_gpu_reverse_after_a_change_of_tracer_set():
    """One context: a 3WATER reverse call, then 4WATER by set_params (the records grow
    from 92 to 102 slots), then 3WATER again; each call against the host passes."""
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6, default_params
    from adjoint_helpers import compare_with_host, output_cotangents
    gz, gzh, st = _state(lall=2, ijdim=200, seed=45)
    drv = NSW6(gz, gzh, params=default_params(), max_columns=400)
    try:
        for four in (0.0, 1.0, 0.0):
            P_ora, P_gpu = _with(O.params(), four_water=four), _with(default_params(), four_water=four)
            drv.set_params(P_gpu)
            cot, cs = output_cotangents(st, gz, gzh, DT, P_ora, 9)
            g_host, gc_host, status = O.adjoint_host(st, gz, gzh, DT, cot, cs, P=P_ora)
            assert status == 0
            print(f"adjoint four_water={four}", flush=True)
            dcot = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in cot.items()}
            grad, gc = drv.adjoint(F.to_device(st), dcot, DT,
                                   cot_sflux=torch.from_numpy(np.ascontiguousarray(cs)).cuda())
            torch.cuda.synchronize()
            compare_with_host(st, gz, gzh, DT, P_ora, cot, cs, {k: v.cpu().numpy() for k, v in grad.items()},
                              gc.cpu().numpy(), g_host, gc_host)
    finally:
        drv.close()
