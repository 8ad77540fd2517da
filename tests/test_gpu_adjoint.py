"""Reverse mode of the FP64 step on the GPU (nsw6_mp_driver_adj through the C ABI).

* Elementwise agreement with the host run of the same reverse passes
  (oracle.adjoint_host): every input cotangent within 1e-9 of its field's max-abs,
  the coefficient cotangents within 1e-9 relative; also with Berry's autoconversion
  (sw_kk2000 = 0), without explicit ice generation (sw_expice = 0: Bergeron growth and
  the mixed-phase saturation adjustment), without Roh & Satoh snow (sw_roh2014 = 0) and
  with all three switched.
* Directional agreement with forward-mode dual numbers on the host (oracle.dual_vjp):
  |<grad, v> - dJ/dv| <= 1e-6 |dJ/dv| along random directions through every input
  field and the 8 coefficients.
* Central finite differences (h = 1e-6 relative) of the GPU forward step itself on
  the adjoint-check loss (sum of squared surface precipitation plus the rhoge
  tendency) agree with the GPU gradient to 1e-6 relative.
* Through two chained steps, the torch.autograd.Function gives the gradient of the
  chained direct calls to round-off and matches central differences of the
  two-step loss.
* Determinism (two calls are bitwise equal).
* States whose every region takes the general wall fluxes (rain 25x faster; deep
  convection on the reference vgrid at dt = 60 s): the GPU reverse mode against the
  host passes and against dual numbers.
"""
import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from nicam_kokkos_microphysics_amd import fields as F

from adjoint_helpers import (DT, STATE_KEYS, all_species, coefs, compare_with_host, directional, non_roh,
                             output_cotangents, random_cotangents, random_directions)

pytestmark = pytest.mark.gpu


def _state(case, lall=2, ijdim=300):
    if case == "default":
        return F.synth_state(lall, ijdim, seed=11)
    if case == "supersaturated":
        return F.synth_state(lall, ijdim, seed=5, frac_supersat=0.2)
    if case == "towers":
        return F.synth_state(lall, ijdim, seed=9, frac_towers=0.3)
    raise ValueError(case)


def _gpu_adjoint(st, gz, gzh, cot, cs, max_columns=None):
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6
    lall, ijdim = st["rhog"].shape[0], st["rhog"].shape[-1]
    drv = NSW6(gz, gzh, max_columns=max_columns or lall * ijdim)
    dev = F.to_device(st)
    dcot = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in cot.items()}
    dcs = None if cs is None else torch.from_numpy(np.ascontiguousarray(cs)).cuda()
    print(f"adjoint lall={lall} ijdim={ijdim}", flush=True)
    grad, gc = drv.adjoint(dev, dcot, DT, cot_sflux=dcs)
    torch.cuda.synchronize()
    rep = drv.report()
    out = {k: v.cpu().numpy() for k, v in grad.items()}
    gcoef = gc.cpu().numpy()
    drv.close()
    return out, gcoef, rep


@pytest.mark.parametrize("case", ["default", "supersaturated", "towers"])
def test_gpu_reverse_matches_host_passes(case):
    from oracle import oracle as O
    gz, gzh, st = _state(case)
    cot, cs = random_cotangents(st, gz, gzh, 7)
    g_host, gc_host, status = O.adjoint_host(st, gz, gzh, DT, cot, cs)
    assert status == 0
    g_gpu, gc_gpu, rep = _gpu_adjoint(st, gz, gzh, cot, cs)
    assert rep.status == 0 and rep.n_general_regions == 0
    for k in F.ADJ_INPUTS:
        scale = np.abs(g_host[k]).max()
        err = np.abs(g_gpu[k] - g_host[k]).max()
        assert err <= 1e-9 * scale + 1e-300, (k, err, scale)
    np.testing.assert_allclose(gc_gpu, gc_host, rtol=1e-9, atol=0)


def test_gpu_berry_reverse_matches_host_passes():
    """sw_kk2000 = 0 (Berry's autoconversion, the original rain accretion of cloud
    water): the GPU reverse mode against the host passes (themselves checked against
    dual numbers), elementwise within 1e-9 of max-abs."""
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6, default_params
    gz, gzh, st = _state("supersaturated")
    Ph = O.params()
    Ph.sw_kk2000 = 0.0
    cot, cs = random_cotangents(st, gz, gzh, 7)
    g_host, gc_host, status = O.adjoint_host(st, gz, gzh, DT, cot, cs, P=Ph)
    assert status == 0
    Pg = default_params()
    Pg.sw_kk2000 = 0.0
    lall, ijdim = st["rhog"].shape[0], st["rhog"].shape[-1]
    drv = NSW6(gz, gzh, params=Pg, max_columns=lall * ijdim)
    dcot = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in cot.items()}
    grad, gc = drv.adjoint(F.to_device(st), dcot, DT, cot_sflux=torch.from_numpy(np.ascontiguousarray(cs)).cuda())
    torch.cuda.synchronize()
    for k in F.ADJ_INPUTS:
        g = grad[k].cpu().numpy()
        scale = np.abs(g_host[k]).max()
        err = np.abs(g - g_host[k]).max()
        assert err <= 1e-9 * scale + 1e-300, (k, err, scale)
    np.testing.assert_allclose(gc.cpu().numpy(), gc_host, rtol=1e-9, atol=0)
    drv.close()


def _scheme_case(case):
    """States and parameters of the non-default schemes: without explicit ice generation
    (sw_expice = 0), without Roh & Satoh snow (sw_roh2014 = 0, with the reference's
    parameters for it), with Berry's autoconversion (sw_kk2000 = 0) and with every
    switch off."""
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import default_params
    Po, Pg = O.params(), default_params()
    scheme, state = case.split(":")
    edits = {"no_expice": {"sw_expice": 0.0}, "no_expice_berry": {"sw_expice": 0.0, "sw_kk2000": 0.0},
             "non_roh": {"sw_roh2014": 0.0},
             "all_off": {"sw_roh2014": 0.0, "sw_expice": 0.0, "sw_kk2000": 0.0}}[scheme]
    for k, v in edits.items():
        setattr(Po, k, v)
        setattr(Pg, k, v)
    if Po.sw_roh2014 == 0.0:
        non_roh(Po)
        non_roh(Pg)
    if state == "all_species":
        gz, gzh, st = F.synth_state(2, 300, seed=21)
        st = all_species(st, 1)
    elif state == "supersaturated":
        gz, gzh, st = F.synth_state(2, 300, seed=5, frac_supersat=0.3)
    else:
        gz, gzh, st = F.synth_state(2, 300, seed=22, frac_towers=0.5)
        st = all_species(st, 2)
        for _ in range(3):
            assert O.mp_driver(st, gz, gzh, DT, P=Po).status == 0
    return gz, gzh, st, Po, Pg


@pytest.mark.parametrize("case", ["no_expice:all_species", "no_expice:supersaturated", "no_expice_berry:evolved",
                                  "non_roh:all_species", "all_off:evolved"])
def test_gpu_scheme_variants_reverse(case):
    """sw_expice = 0 (Bergeron growth, mixed-phase saturation adjustment), sw_roh2014 = 0
    (exponential snow moments), with KK2000 or Berry's autoconversion: the GPU reverse
    mode against the host run of the same passes and against dual numbers
    (compare_with_host: elementwise 1e-9 of max-abs, directional 1e-9, points without a
    derivative handled apart)."""
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st, Po, Pg = _scheme_case(case)
    cot, cs = output_cotangents(st, gz, gzh, DT, Po, 7)
    g_host, gc_host, status = O.adjoint_host(st, gz, gzh, DT, cot, cs, P=Po)
    assert status == 0
    lall, ijdim = st["rhog"].shape[0], st["rhog"].shape[-1]
    drv = NSW6(gz, gzh, params=Pg, max_columns=lall * ijdim)
    try:
        dcot = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in cot.items()}
        print(f"adjoint {case}", flush=True)
        grad, gc = drv.adjoint(F.to_device(st), dcot, DT,
                               cot_sflux=torch.from_numpy(np.ascontiguousarray(cs)).cuda())
        torch.cuda.synchronize()
        rep = drv.report()
    finally:
        drv.close()
    assert rep.status == 0
    g_gpu = {k: v.cpu().numpy() for k, v in grad.items()}
    n = compare_with_host(st, gz, gzh, DT, Po, cot, cs, g_gpu, gc.cpu().numpy(), g_host, gc_host)
    print(f"{n} kink entries", flush=True)


@pytest.mark.parametrize("case", ["default", "supersaturated"])
def test_gpu_reverse_matches_dual_numbers(case):
    from oracle import oracle as O
    gz, gzh, st = _state(case, lall=1, ijdim=200)
    cot, cs = random_cotangents(st, gz, gzh, 3)
    g_gpu, gc_gpu, rep = _gpu_adjoint(st, gz, gzh, cot, cs)
    assert rep.status == 0
    dirs, dcoef = random_directions(st, 17, 8)
    _, dj = O.dual_vjp(st, gz, gzh, DT, cot, cs, dirs, dcoef)
    for i in range(len(dirs)):
        s, _ = directional(g_gpu, gc_gpu, dirs[i], dcoef[i])
        assert abs(s - dj[i]) <= 1e-6 * abs(dj[i]), (i, s, dj[i])


def _check_loss_gpu(drv, st_dev, metrics_keys=None):
    """Loss of the adjoint-check form on GPU outputs: sum of squared surface
    precipitation (rain row and solid row) plus the rhoge tendency."""
    import torch
    s = F.copy_state(st_dev)
    rhoge_in = s["rhoge"].clone()
    drv.step(s, DT)
    torch.cuda.synchronize()
    p = s["precip"]
    return float((p * p).sum() + ((s["rhoge"] - rhoge_in) / DT).sum()), s


def test_gpu_reverse_matches_central_differences():
    """Gradient of L = sum precip^2 + sum (rhoge_out - rhoge_in)/dt with respect to
    every input field and coefficient, against central differences of the GPU step
    (h = 1e-6 relative) along combined random directions."""
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6, default_params
    gz, gzh, st = _state("supersaturated", lall=1, ijdim=256)
    drv = NSW6(gz, gzh, max_columns=256)
    dev = F.to_device(st)
    L0, out = _check_loss_gpu(drv, dev)
    cot = {"precip": 2.0 * out["precip"], "rhoge": torch.full_like(out["rhoge"], 1.0 / DT)}
    grad, gc = drv.adjoint(dev, cot, DT)
    grad["rhoge"] = grad["rhoge"] - 1.0 / DT   # explicit -rhoge_in/dt term of the loss
    g = {k: v.cpu().numpy() for k, v in grad.items()}
    gcoef = gc.cpu().numpy()
    dirs, dcoef = random_directions(st, 29, 3)
    c0 = coefs()
    h = 1e-6
    for i in range(3):
        vals = []
        for sgn in (1.0, -1.0):
            P = default_params()
            for c, name in enumerate(N.TUNABLE_COEFFS):
                setattr(P, name, c0[c] + sgn * h * dcoef[i][c])
            drv.set_params(P)
            sp = F.copy_state(st)
            for k, d in dirs[i].items():
                sp[k] = sp[k] + sgn * h * d
            vals.append(_check_loss_gpu(drv, F.to_device(sp))[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        s, _ = directional(g, gcoef, dirs[i], dcoef[i])
        assert abs(s - fd) <= 1e-6 * abs(fd), (i, s, fd)
    drv.close()


def test_gpu_reverse_is_deterministic():
    gz, gzh, st = _state("default", lall=2, ijdim=300)
    cot, cs = random_cotangents(st, gz, gzh, 5)
    a, ga, _ = _gpu_adjoint(st, gz, gzh, cot, cs)
    b, gb, _ = _gpu_adjoint(st, gz, gzh, cot, cs)
    for k in F.ADJ_INPUTS:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(ga, gb)


def _general_case(case):
    """States whose regions all take the general wall fluxes: rain 25 times faster
    (the fast-rain parity case), and deep convection on the reference vgrid at
    dt = 60 s (kcell ranges of several cells, upward departures)."""
    import os
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import default_params
    Po, Pg = O.params(), default_params()
    if case == "fast_rain":
        gz, gzh, st = F.synth_state(2, 400, seed=3)
        Po.Cr = Pg.Cr = 130.0 * 25
        return gz, gzh, st, 15.0, Po, Pg
    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    gz = np.fromfile(os.path.join(g, "GRD_gz.dat"), dtype="<f8")
    gzh = np.fromfile(os.path.join(g, "GRD_gzh.dat"), dtype="<f8")
    gz, gzh, st = F.synth_state(2, 400, seed=41, gz=gz, gzh=gzh, frac_towers=0.1, frac_rainshafts=0.05)
    return gz, gzh, st, 60.0, Po, Pg


@pytest.mark.parametrize("case", ["fast_rain", "reference_grid_dt60"])
def test_gpu_reverse_general_regions(case):
    """Every region general: the GPU reverse mode (k_adj_wall_general) against the host
    run of the same passes and against dual numbers (compare_with_host: elementwise
    1e-9 of max-abs and directional 1e-9, points without a derivative handled
    apart)."""
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st, dt, Po, Pg = _general_case(case)
    cot, cs = output_cotangents(st, gz, gzh, dt, Po, 7)
    g_host, gc_host, status = O.adjoint_host(st, gz, gzh, dt, cot, cs, P=Po)
    assert status == 0
    lall, ijdim = st["rhog"].shape[0], st["rhog"].shape[-1]
    drv = NSW6(gz, gzh, params=Pg, max_columns=lall * ijdim)
    try:
        dcot = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in cot.items()}
        print(f"adjoint {case}", flush=True)
        grad, gc = drv.adjoint(F.to_device(st), dcot, dt,
                               cot_sflux=torch.from_numpy(np.ascontiguousarray(cs)).cuda())
        torch.cuda.synchronize()
        rep = drv.report()
    finally:
        drv.close()
    assert rep.status == 0 and rep.n_general_regions == lall
    g_gpu = {k: v.cpu().numpy() for k, v in grad.items()}
    n = compare_with_host(st, gz, gzh, dt, Po, cot, cs, g_gpu, gc.cpu().numpy(), g_host, gc_host)
    print(f"{n} kink entries", flush=True)


def test_autograd_function_and_two_step_chain():
    """Two chained differentiable steps through MPStep: the autograd gradient equals
    the chain of direct reverse-mode calls (step 2's input cotangents as step 1's
    output cotangents) to round-off, and central differences of the two-step loss.
    The differences use h = 1e-7: at h = 1e-6 the stencil of this state crosses
    limiter switches at some of its 24,576 points after the first step (the
    difference quotient then moves by ~1e-3 between h = 1e-5 and 1e-6 and settles
    only below), which is a kink of the step, not a gradient error."""
    import torch
    from nicam_kokkos_microphysics_amd.autograd import mp_step
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = _state("supersaturated", lall=1, ijdim=256)
    drv = NSW6(gz, gzh, max_columns=256)

    def two_step_loss(host_state, track):
        dev = F.to_device(host_state)
        leaves = {}
        for k in F.ADJ_INPUTS:
            if track:
                dev[k].requires_grad_(True)
                leaves[k] = dev[k]
        o1 = mp_step(drv, dev, DT)
        s2 = dict(dev)
        s2.update(o1)
        o2 = mp_step(drv, s2, DT)
        L = (o1["precip"] ** 2).sum() + (o2["precip"] ** 2).sum() + ((o2["rhoge"] - dev["rhoge"]) / DT).sum()
        return L, leaves, o1, o2

    L, leaves, o1, o2 = two_step_loss(st, True)
    L.backward()
    g = {k: v.grad.cpu().numpy() for k, v in leaves.items()}

    # the same chain by hand through the C ABI
    with torch.no_grad():
        x = F.to_device(st)
        s1 = dict(x)
        s1.update({k: v.detach() for k, v in o1.items()})
        g2, _ = drv.adjoint(s1, {"precip": 2.0 * o2["precip"].detach(),
                                 "rhoge": torch.full_like(x["rhoge"], 1.0 / DT)}, DT)
        cot1 = {k: g2[k] for k in F.ADJ_INPUTS if k in F.ADJ_OUTPUTS}
        cot1["precip"] = 2.0 * o1["precip"].detach()
        g1, _ = drv.adjoint(x, cot1, DT)
        ref = {k: v.cpu().numpy() for k, v in g1.items()}
        for k in ("vx", "vy", "vz"):
            ref[k] = ref[k] + g2[k].cpu().numpy()
        ref["rhoge"] = ref["rhoge"] - 1.0 / DT
    for k in F.ADJ_INPUTS:
        scale = np.abs(ref[k]).max()
        assert np.abs(g[k] - ref[k]).max() <= 1e-12 * scale + 1e-300, k

    dirs, _ = random_directions(st, 31, 2)
    h = 1e-7
    with torch.no_grad():
        for i in range(2):
            sp = {k: (st[k] + h * dirs[i][k] if k in dirs[i] else st[k]) for k in st}
            sm = {k: (st[k] - h * dirs[i][k] if k in dirs[i] else st[k]) for k in st}
            fd = (float(two_step_loss(sp, False)[0]) - float(two_step_loss(sm, False)[0])) / (2 * h)
            s = sum(float(np.sum(g[k] * dirs[i][k])) for k in dirs[i] if k in g)
            assert abs(s - fd) <= 1e-5 * abs(fd), (i, s, fd)
    drv.close()
