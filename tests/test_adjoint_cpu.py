"""Host checks of the adjoint machinery (no GPU needed).

* The host restatement of the templated step (the statements the HIP kernels run)
  reproduces the oracle to round-off, so the dual-number gradient differentiates the
  step that the parity tests pin.
* Forward-mode dual-number directional derivatives of the adjoint-check loss agree
  with central finite differences (h = 1e-6 relative) along random directions that
  perturb every input field and the 8 coefficients at once.
* The general regions (departures beyond the next cell): the host step's general
  wall fluxes against the oracle, and the whole reverse mode against dual numbers.
* Without explicit ice generation (sw_expice = 0: Bergeron growth and the mixed-phase
  saturation adjustment): every reverse piece and the whole reverse mode against dual
  numbers, and those dual numbers against central differences.
* Without the Roh & Satoh snow moments (sw_roh2014 = 0), alone and with every other
  switch off: the same, the dual numbers checked against differences at single points.
* Every hand-written reverse piece (point rates, wall fluxes, cell divergence and
  KIN2EIN, point-wise tail with the implicit-function Newton adjoint) agrees with
  forward-mode dual numbers at every point, wall and cell of several states, including
  ones that activate every process (hydrometeors of all species at all levels).
"""
import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import fields as F
from nicam_kokkos_microphysics_amd.check import field_errors
from oracle import oracle as O

from adjoint_helpers import all_species, directional, non_roh, random_cotangents, random_directions

DT = 15.0
STATE_KEYS = ["rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rhogq", "tem", "q", "vx", "vy", "vz"]


_everywhere = all_species


def _coefs():
    P = O.params()
    return np.array([P.kk_aut_rate, P.qicrt_saut, P.Esi, P.Esr, P.Cr, P.Cs, P.Cg, P.icenuc_coef])


@pytest.mark.parametrize("seed,fs,ft", [(5, 0.2, 0.0), (9, 0.0, 0.3)])
def test_host_step_matches_oracle(seed, fs, ft):
    gz, gzh, st = F.synth_state(1, 64, seed=seed, frac_supersat=fs, frac_towers=ft)
    a, b = F.copy_state(st), F.copy_state(st)
    assert O.mp_driver(a, gz, gzh, DT).status == 0
    assert O.step_fp64(b, gz, gzh, DT) == 0
    keys = ["rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rho", "tem", "pre", "qd", "precip",
            "precip_rhoe", "precip_lh_heat", "precip_rhophi", "precip_rhokin", "GPREC"] + \
        [f"rhogq:{t}" for t in F.TRACERS] + [f"q:{t}" for t in F.TRACERS]
    worst = max(v[2] for v in field_errors(b, a, keys).values())
    assert worst <= 1e-12, worst


def test_dual_matches_central_differences():
    gz, gzh, st = F.synth_state(1, 32, seed=5, frac_supersat=0.2)
    rng = np.random.default_rng(3)
    c0 = _coefs()
    dirs, dcoef = [], []
    for _ in range(3):
        dirs.append({k: st[k] * rng.uniform(-1, 1, st[k].shape) for k in STATE_KEYS})
        dcoef.append(c0 * rng.uniform(-1, 1, 8))
    dcoef = np.array(dcoef)
    _, dl = O.dual_loss(st, gz, gzh, DT, dirs, dcoef)
    h = 1e-6
    for i in range(3):
        sp = {k: (st[k] + h * dirs[i][k] if k in dirs[i] else st[k]) for k in st}
        sm = {k: (st[k] - h * dirs[i][k] if k in dirs[i] else st[k]) for k in st}
        fd = (O.loss_fp64(sp, gz, gzh, DT, coef=c0 + h * dcoef[i]) -
              O.loss_fp64(sm, gz, gzh, DT, coef=c0 - h * dcoef[i])) / (2 * h)
        assert abs(dl[i] - fd) <= 1e-6 * abs(fd), (dl[i], fd)


@pytest.mark.parametrize("case", ["default", "supersaturated", "towers", "all_species", "evolved"])
def test_adjoint_pieces_match_dual_numbers(case):
    if case == "default":
        gz, gzh, st = F.synth_state(1, 48, seed=11)
    elif case == "supersaturated":
        gz, gzh, st = F.synth_state(1, 48, seed=5, frac_supersat=0.2)
    elif case == "towers":
        gz, gzh, st = F.synth_state(1, 48, seed=9, frac_towers=0.3)
    elif case == "all_species":
        gz, gzh, st = F.synth_state(1, 48, seed=21)
        st = _everywhere(st, 1)
    else:
        gz, gzh, st = F.synth_state(1, 48, seed=22, frac_towers=0.5)
        st = _everywhere(st, 2)
        for _ in range(3):
            assert O.mp_driver(st, gz, gzh, DT).status == 0
    err, cnt = O.check_pieces(st, gz, gzh, DT)
    assert np.all(cnt > 0)
    assert np.all(err <= 1e-6), dict(zip(["point", "wall", "sediment", "tail"], err))


@pytest.mark.parametrize("case", ["all_species", "evolved"])
def test_berry_reverse_matches_dual_numbers(case):
    """sw_kk2000 = 0 (Berry's autoconversion and the original collection of cloud water
    by rain in place of KK2000, src/mod_mp_nsw6.cpp:1090-1100, 1140-1160): every reverse
    piece against dual numbers at every point, and the whole step's reverse against
    the dual-number directional derivative."""
    P = O.params()
    P.sw_kk2000 = 0.0
    if case == "all_species":
        gz, gzh, st = F.synth_state(2, 40, seed=21)
        st = _everywhere(st, 1)
    else:
        gz, gzh, st = F.synth_state(2, 40, seed=22, frac_towers=0.5)
        st = _everywhere(st, 2)
        for _ in range(3):
            assert O.mp_driver(st, gz, gzh, DT, P=P).status == 0
    err, cnt = O.check_pieces(st, gz, gzh, DT, P=P)
    assert np.all(cnt > 0)
    assert np.all(err <= 1e-6), dict(zip(["point", "wall", "sediment", "tail"], err))
    cot, cs = random_cotangents(st, gz, gzh, 7)
    grad, gc, status = O.adjoint_host(st, gz, gzh, DT, cot, cs, P=P)
    assert status == 0
    dirs, dcoef = random_directions(st, 13, 4)
    _, dj = O.dual_vjp(st, gz, gzh, DT, cot, cs, dirs, dcoef, P=P)
    for i in range(len(dirs)):
        s_, a = directional(grad, gc, dirs[i], dcoef[i])
        assert abs(s_ - dj[i]) <= 1e-13 * a, (i, s_, dj[i], a)


@pytest.mark.parametrize("case", ["all_species", "supersaturated", "evolved", "berry"])
def test_no_explicit_ice_reverse_matches_dual_numbers(case):
    """sw_expice = 0: the Bergeron growth Psfw, Psfi between -30 and 0 degC (with a1,
    a2 of the table interpolation differentiated in temperature) in place of explicit
    ice generation, and the mixed-phase saturation adjustment (satadjust_all, its
    converged solution differentiated through the implicit-function theorem).  Every
    reverse piece against dual numbers at every point, and the whole step's reverse
    against the dual-number directional derivative (the dual numbers of this scheme are
    themselves checked against central differences below)."""
    P = O.params()
    P.sw_expice = 0.0
    if case == "all_species":
        gz, gzh, st = F.synth_state(2, 40, seed=21)
        st = _everywhere(st, 1)
    elif case == "supersaturated":
        gz, gzh, st = F.synth_state(2, 40, seed=5, frac_supersat=0.3)
    else:
        if case == "berry":
            P.sw_kk2000 = 0.0
        gz, gzh, st = F.synth_state(2, 40, seed=22, frac_towers=0.5)
        st = _everywhere(st, 2)
        for _ in range(3):
            assert O.mp_driver(st, gz, gzh, DT, P=P).status == 0
    # points where the Bergeron growth acts: -30 < T - 273.15 < 0 with cloud water and ice
    # (the supersaturated state exercises the mixed-phase adjustment instead)
    berg = (st["tem"] > 243.15) & (st["tem"] < 273.15) & (st["q"][:, 1] > 0) & (st["q"][:, 3] > 0)
    assert case == "supersaturated" or berg[:, 1:-1].sum() >= 100, berg.sum()
    err, cnt = O.check_pieces(st, gz, gzh, DT, P=P)
    assert np.all(cnt > 0)
    assert np.all(err <= 1e-6), dict(zip(["point", "wall", "sediment", "tail"], err))
    cot, cs = random_cotangents(st, gz, gzh, 7)
    grad, gc, status = O.adjoint_host(st, gz, gzh, DT, cot, cs, P=P)
    assert status == 0
    dirs, dcoef = random_directions(st, 13, 4)
    _, dj = O.dual_vjp(st, gz, gzh, DT, cot, cs, dirs, dcoef, P=P)
    for i in range(len(dirs)):
        s_, a = directional(grad, gc, dirs[i], dcoef[i])
        assert abs(s_ - dj[i]) <= 1e-13 * a, (i, s_, dj[i], a)


@pytest.mark.parametrize("case", ["supersaturated", "all_species"])
def test_no_explicit_ice_dual_matches_central_differences(case):
    """The dual numbers of the sw_expice = 0 step (Bergeron a1, a2 tangents, the
    satadjust_all implicit-function tangent) against central differences of the host
    step (h = 1e-6 relative) along random directions through every input and the 8
    coefficients."""
    P = O.params()
    P.sw_expice = 0.0
    if case == "supersaturated":
        gz, gzh, st = F.synth_state(1, 32, seed=5, frac_supersat=0.3)
    else:
        gz, gzh, st = F.synth_state(1, 32, seed=21)
        st = _everywhere(st, 1)
    a, b = F.copy_state(st), F.copy_state(st)
    assert O.mp_driver(a, gz, gzh, DT, P=P).status == 0
    assert O.step_fp64(b, gz, gzh, DT, P=P) == 0
    assert np.abs(a["tem"] - b["tem"]).max() <= 1e-9
    rng = np.random.default_rng(3)
    c0 = _coefs()
    dirs, dcoef = [], []
    for _ in range(3):
        dirs.append({k: st[k] * rng.uniform(-1, 1, st[k].shape) for k in STATE_KEYS})
        dcoef.append(c0 * rng.uniform(-1, 1, 8))
    dcoef = np.array(dcoef)
    _, dl = O.dual_loss(st, gz, gzh, DT, dirs, dcoef, P=P)
    h = 1e-6
    for i in range(3):
        sp = {k: (st[k] + h * dirs[i][k] if k in dirs[i] else st[k]) for k in st}
        sm = {k: (st[k] - h * dirs[i][k] if k in dirs[i] else st[k]) for k in st}
        fd = (O.loss_fp64(sp, gz, gzh, DT, coef=c0 + h * dcoef[i], P=P) -
              O.loss_fp64(sm, gz, gzh, DT, coef=c0 - h * dcoef[i], P=P)) / (2 * h)
        assert abs(dl[i] - fd) <= 1e-7 * abs(fd), (i, dl[i], fd)


@pytest.mark.parametrize("case", ["all_species", "evolved", "every_switch_off"])
def test_non_roh_reverse_matches_dual_numbers(case):
    """sw_roh2014 = 0 with the reference's parameters for it (non_roh): the snow moments
    of the exponential size distribution from the slope (dens qs / (As N0s
    GAM_1bs))^(1/4) in place of Roh & Satoh's fit; alone, and with explicit ice
    generation and KK2000 off too (sw_expice = sw_kk2000 = 0).  Every reverse piece
    against dual numbers at every point, and the whole step's reverse against the
    dual-number directional derivative."""
    P = non_roh(O.params())
    if case == "all_species":
        gz, gzh, st = F.synth_state(2, 40, seed=21)
        st = _everywhere(st, 1)
    else:
        if case == "every_switch_off":
            P.sw_expice = 0.0
            P.sw_kk2000 = 0.0
        gz, gzh, st = F.synth_state(2, 40, seed=22, frac_towers=0.5)
        st = _everywhere(st, 2)
        for _ in range(3):
            assert O.mp_driver(st, gz, gzh, DT, P=P).status == 0
    assert (st["q"][:, 4, 1:-1] > 1e-8).sum() >= 1000
    err, cnt = O.check_pieces(st, gz, gzh, DT, P=P)
    assert np.all(cnt > 0)
    assert np.all(err <= 1e-6), dict(zip(["point", "wall", "sediment", "tail"], err))
    cot, cs = random_cotangents(st, gz, gzh, 7)
    grad, gc, status = O.adjoint_host(st, gz, gzh, DT, cot, cs, P=P)
    assert status == 0
    dirs, dcoef = random_directions(st, 13, 4)
    _, dj = O.dual_vjp(st, gz, gzh, DT, cot, cs, dirs, dcoef, P=P)
    for i in range(len(dirs)):
        s_, a = directional(grad, gc, dirs[i], dcoef[i])
        assert abs(s_ - dj[i]) <= 1e-13 * a, (i, s_, dj[i], a)


@pytest.mark.parametrize("case", ["supersaturated", "all_species"])
def test_non_roh_dual_matches_differences_at_points(case):
    """The dual numbers of the sw_roh2014 = 0 step against differences of the host step
    at single points with snow (rhog and tem, 40 points each, h = 1e-6 relative).  The
    limiters put some points on a kink (the one-sided differences differ by more than
    the curvature): there the dual number must equal one of the one-sided differences,
    elsewhere the central one; within 1e-6 of max(|derivative|, 1)."""
    P = non_roh(O.params())
    if case == "supersaturated":
        gz, gzh, st = F.synth_state(1, 32, seed=5, frac_supersat=0.3)
    else:
        gz, gzh, st = F.synth_state(1, 32, seed=21)
        st = _everywhere(st, 1)
    c0 = _coefs()
    rng = np.random.default_rng(1)
    cand = np.argwhere(st["q"][0, 4, 1:-1] > 1e-8) + [1, 0]
    dirs, keys = [], []
    for k, j in cand[rng.choice(len(cand), 40, replace=False)]:
        for key in ("rhog", "tem"):
            d = np.zeros_like(st[key])
            d[0, k, j] = st[key][0, k, j]
            dirs.append({key: d})
            keys.append(key)
    _, dl = O.dual_loss(st, gz, gzh, DT, dirs, np.zeros((len(dirs), 8)), P=P)
    L0 = O.loss_fp64(st, gz, gzh, DT, coef=c0, P=P)
    h = 1e-6
    for i, key in enumerate(keys):
        sp, sm = dict(st), dict(st)
        sp[key] = st[key] + h * dirs[i][key]
        sm[key] = st[key] - h * dirs[i][key]
        r = (O.loss_fp64(sp, gz, gzh, DT, coef=c0, P=P) - L0) / h
        l = (L0 - O.loss_fp64(sm, gz, gzh, DT, coef=c0, P=P)) / h
        err = min(abs(dl[i] - r), abs(dl[i] - l), abs(dl[i] - 0.5 * (r + l)))
        assert err <= 1e-6 * max(abs(dl[i]), 1.0), (i, key, dl[i], r, l)


def _state(case):
    if case == "default":
        return F.synth_state(2, 40, seed=11)
    if case == "supersaturated":
        return F.synth_state(2, 40, seed=5, frac_supersat=0.2)
    if case == "towers":
        return F.synth_state(2, 40, seed=9, frac_towers=0.3)
    gz, gzh, st = F.synth_state(2, 40, seed=22, frac_towers=0.5)
    st = _everywhere(st, 2)
    for _ in range(3):
        assert O.mp_driver(st, gz, gzh, DT).status == 0
    return gz, gzh, st


@pytest.mark.parametrize("case", ["default", "supersaturated", "towers", "evolved"])
def test_whole_step_reverse_matches_dual_numbers(case):
    """The three reverse passes (record gathers included), run on the host, give
    <grad, v> = dJ/dv of J = <cot, outputs> along random directions through every
    input field and the 8 coefficients."""
    gz, gzh, st = _state(case)
    cot, cs = random_cotangents(st, gz, gzh, 7)
    grad, gc, status = O.adjoint_host(st, gz, gzh, DT, cot, cs)
    assert status == 0
    dirs, dcoef = random_directions(st, 13, 4)
    _, dj = O.dual_vjp(st, gz, gzh, DT, cot, cs, dirs, dcoef)
    for i in range(len(dirs)):
        s, a = directional(grad, gc, dirs[i], dcoef[i])
        assert abs(s - dj[i]) <= 1e-13 * a, (i, s, dj[i], a)


def test_whole_step_reverse_per_field():
    """Each input field's cotangent on its own (one direction per field and one per
    coefficient), so a gradient routed to the wrong field cannot cancel out."""
    gz, gzh, st = _state("supersaturated")
    cot, cs = random_cotangents(st, gz, gzh, 3)
    grad, gc, status = O.adjoint_host(st, gz, gzh, DT, cot, cs)
    assert status == 0
    rng = np.random.default_rng(4)
    dirs, dcoef, names = [], [], []
    for k in STATE_KEYS:
        dirs.append({k: st[k] * rng.uniform(-1, 1, st[k].shape)})
        dcoef.append(np.zeros(8))
        names.append(k)
    c0 = _coefs()
    for c in range(8):
        dirs.append({})
        e = np.zeros(8)
        e[c] = c0[c]
        dcoef.append(e)
        names.append(f"coef{c}")
    dcoef = np.array(dcoef)
    _, dj = O.dual_vjp(st, gz, gzh, DT, cot, cs, dirs, dcoef)
    for i, name in enumerate(names):
        s, a = directional(grad, gc, dirs[i], dcoef[i])
        assert abs(s - dj[i]) <= 1e-13 * a, (name, s, dj[i], a)


def _general_state(case):
    """States whose every region takes the general wall fluxes: rain 25 times faster,
    and deep convection on the reference vgrid at dt = 60 s (kcell ranges of several
    cells, upward departures).  Returns (gz, gzh, state, dt, params)."""
    import os
    P = O.params()
    if case == "fast_rain":
        gz, gzh, st = F.synth_state(2, 24, seed=3)
        P.Cr = 130.0 * 25
        return gz, gzh, st, 15.0, P
    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    gz = np.fromfile(os.path.join(g, "GRD_gz.dat"), dtype="<f8")
    gzh = np.fromfile(os.path.join(g, "GRD_gzh.dat"), dtype="<f8")
    gz, gzh, st = F.synth_state(2, 24, seed=41, gz=gz, gzh=gzh, frac_towers=0.1, frac_rainshafts=0.1)
    return gz, gzh, st, 60.0, P


def _general_cotangents(st, gz, gzh, dt, P, seed):
    rng = np.random.default_rng(seed)
    out = F.copy_state(st)
    assert O.step_fp64(out, gz, gzh, dt, P=P) == 0
    cot = {k: rng.uniform(-1, 1, out[k].shape) / (float(np.sqrt(np.mean(out[k] ** 2))) or 1.0)
           for k in F.ADJ_OUTPUTS}
    return cot, rng.uniform(-1, 1, (st["rhog"].shape[0], 3, st["rhog"].shape[-1])) / 1e-3


@pytest.mark.parametrize("case", ["fast_rain", "reference_grid_dt60"])
def test_general_regions_host_step_matches_oracle(case):
    """The templated step's general wall fluxes (wall_flux_general, the statements the
    GPU's k_flux_general runs) against the oracle, and every region is general."""
    gz, gzh, st, dt, P = _general_state(case)
    a, b = F.copy_state(st), F.copy_state(st)
    rep, bounds, _ = O.mp_driver_traced(a, gz, gzh, dt, P=P, newton=False)
    assert rep.status == 0
    bnd = np.asarray(bounds).reshape(st["rhog"].shape[0], 3, 2, 96, 2)
    k = np.arange(96)
    assert (((bnd[:, :, 0, :, 0] != k) | (bnd[:, :, 0, :, 1] > k + 1))[:, :, 1:95]).any(axis=(1, 2)).all()
    assert O.step_fp64(b, gz, gzh, dt, P=P) == 0
    keys = ["rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rho", "tem", "pre", "qd", "precip",
            "precip_rhoe", "precip_lh_heat", "precip_rhophi", "precip_rhokin", "GPREC"] + \
        [f"rhogq:{t}" for t in F.TRACERS] + [f"q:{t}" for t in F.TRACERS]
    worst = max(v[2] for v in field_errors(b, a, keys).values())
    assert worst <= 1e-12, worst


@pytest.mark.parametrize("case", ["fast_rain", "reference_grid_dt60"])
def test_general_regions_reverse_matches_dual_numbers(case):
    """The reverse of the general wall fluxes (adj_walls_general, one thread per
    column adding into the wall records) with the other passes, on the host, against
    dual numbers: random directions through every input and coefficient, and one
    direction per input field so that a cotangent routed to the wrong field or level
    cannot cancel."""
    gz, gzh, st, dt, P = _general_state(case)
    cot, cs = _general_cotangents(st, gz, gzh, dt, P, 7)
    grad, gc, status = O.adjoint_host(st, gz, gzh, dt, cot, cs, P=P)
    assert status == 0
    dirs, dcoef = random_directions(st, 13, 4)
    rng = np.random.default_rng(4)
    for key in STATE_KEYS:
        dirs.append({key: st[key] * rng.uniform(-1, 1, st[key].shape)})
    dcoef = np.concatenate([dcoef, np.zeros((len(STATE_KEYS), 8))])
    for lo in range(0, len(dirs), 8):
        _, dj = O.dual_vjp(st, gz, gzh, dt, cot, cs, dirs[lo:lo + 8], dcoef[lo:lo + 8], P=P)
        for i in range(len(dj)):
            s, a = directional(grad, gc, dirs[lo + i], dcoef[lo + i])
            assert abs(s - dj[i]) <= 1e-13 * a, (lo + i, s, dj[i], a)
