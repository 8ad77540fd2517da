"""Pins of the CPU oracle (oracle/nsw6_oracle.c) against what fixes each function
independently of the oracle's own text: the reference's data files and constant
tables (tests/golden/, extracted by make_fixtures.py), textbook closed forms,
exact identities and brute-force integrals.
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _ref_constants():
    with open(os.path.join(GOLD, "reference_constants.json")) as fh:
        return json.load(fh)


def _ref_vgrid():
    gz = np.fromfile(os.path.join(GOLD, "GRD_gz.dat"), dtype="<f8")
    gzh = np.fromfile(os.path.join(GOLD, "GRD_gzh.dat"), dtype="<f8")
    return gz, gzh


# ---------------------------------------------------------------- MISC_gammafunc
@pytest.mark.parametrize("x", [4.0, 5.0, 6.0, 3.5, 6.5, 4.5, 2.75, 3.0, 3.25, 2.625, 11.6, 0.7])
def test_gamma_matches_math_gamma(x):
    L = O.lib()
    assert abs(L.ora_gammafunc(x) / math.gamma(x) - 1.0) < 2e-10


def test_gamma_at_integers_is_factorial():
    L = O.lib()
    for n in range(1, 8):
        assert abs(L.ora_gammafunc(float(n + 1)) / math.factorial(n) - 1.0) < 2e-10


# ---------------------------------------------------------------- parameters
def test_params_follow_reference_defaults_and_roh_overrides():
    ref = _ref_constants()["globalparams_defaults"]
    P = O.params()
    # mp_nsw6_init overrides (Roh & Satoh 2014, explicit ice, KK2000, no Pgaut)
    overrides = {"N0g": 4.0e8, "Esi": 0.25, "Egi": 0.0, "Egs": 0.0, "gamma_sacr": 6.0e-2,
                 "gamma_saut": 6.0e-2, "beta_saut": 5.0e-3, "beta_gaut": 0.0, "sw_expice": 1.0,
                 "sw_kk2000": 1.0, "sw_roh2014": 1.0}
    for name, val in ref.items():
        want = overrides.get(name, val)
        assert getattr(P, name) == want, name
    for i, v in enumerate(_ref_constants()["roh_coefficients"]["coef_a"]):
        assert P.coef_a[i] == v
    for i, v in enumerate(_ref_constants()["roh_coefficients"]["coef_b"]):
        assert P.coef_b[i] == v
    # derived: As (Roh), Ar, Ag, Cg
    assert P.As == 0.069 and P.Bs == 2.0
    assert P.Ar == math.pi * 1000.0 / 6.0
    assert abs(P.Cg - math.sqrt(4 * 400.0 * 9.79764 / (3 * 1.28 * 0.6))) < 1e-12
    # gamma constants against math.gamma
    for nm, arg in (("GAM_1br", 4.0), ("GAM_3dr", 3.5), ("GAM_6dr", 6.5), ("GAM_1brdr", 4.5), ("GAM_5dr_h", 2.75),
                    ("GAM_1bs", 3.0), ("GAM_3ds", 3.25), ("GAM_1bsds", 3.25), ("GAM_5ds_h", 2.625),
                    ("GAM_1bg", 4.0), ("GAM_3dg", 3.5), ("GAM_1bgdg", 4.5), ("GAM_5dg_h", 2.75)):
        assert abs(getattr(P, nm) / math.gamma(arg) - 1) < 2e-10, nm


def test_product_params_equal_oracle_params_bitwise():
    P_ora = O.params()
    P_prod = N.Params()
    N.lib().nsw6_params_default(ctypes.byref(P_prod))
    assert bytes(P_ora) == bytes(P_prod)


# ---------------------------------------------------------------- GRD_Setup
def test_grid_setup_on_reference_vgrid_closed_forms():
    gz, gzh = _ref_vgrid()
    g = O.grid(gz, gzh)
    dgz = np.array(g.dgz)
    dgzh = np.array(g.dgzh)
    np.testing.assert_array_equal(dgz[:95], gzh[1:96] - gzh[:95])
    assert dgz[95] == dgz[94]
    np.testing.assert_array_equal(dgzh[1:96], gz[1:96] - gz[:95])
    assert dgzh[0] == dgzh[1]
    np.testing.assert_array_equal(np.array(g.rdgz), 1.0 / dgz)
    af = np.array(g.afact)
    # cell centres are wall midpoints on this grid: afact = dz(k-1) / (dz(k-1) + dz(k))
    dzc = gzh[1:96] - gzh[:95]
    want = dzc[:-1] / (dzc[:-1] + dzc[1:])
    np.testing.assert_allclose(af[1:95], want, rtol=1e-13)
    assert af[0] == 1.0
    np.testing.assert_array_equal(np.array(g.bfact), 1.0 - af)
    cf = np.array(g.cfact)
    np.testing.assert_allclose(cf[1:95], 0.5, rtol=1e-13)   # centre halfway between walls
    assert cf[0] == 1.0 and cf[95] == 0.0


def test_product_grid_setup_equals_oracle():
    gz, gzh = _ref_vgrid()
    a = O.grid(gz, gzh)
    b = N.Grid()
    N.lib().nsw6_grid_setup(gz.ctypes.data, gzh.ctypes.data, ctypes.byref(b))
    assert bytes(a) == bytes(b)


# ---------------------------------------------------------------- saturation
def test_psat_at_triple_point_and_clausius_clapeyron_slope():
    L = O.lib()
    assert L.ora_psat_liq(273.15) == 610.7
    assert L.ora_psat_ice(273.15) == 610.7
    for T in (230.0, 260.0, 290.0, 310.0):
        h = 1e-3
        for fn, Lh in ((L.ora_psat_liq, 2501000.0), (L.ora_psat_ice, 2834000.0)):
            dln = (math.log(fn(T + h)) - math.log(fn(T - h))) / (2 * h)
            assert abs(dln / (Lh / 461.46 / T ** 2) - 1) < 1e-7
    # TEM_MIN clamp
    assert L.ora_psat_liq(5.0) == L.ora_psat_liq(10.0)


def test_satadjust_newton_converges_to_saturation():
    L = O.lib()
    q6 = np.zeros(6)
    rho, T0 = 1.0, 290.0
    Lv, Rv = 2501000.0, 461.46
    qsat0 = L.ora_psat_liq(T0) / (rho * Rv * T0)
    qsum = 1.1 * qsat0
    cv = 717.56
    Em = T0 * cv + qsum * Lv            # moist energy with all water as vapour
    t = ctypes.c_double(T0)
    qv = ctypes.c_double(qsum)
    qc = ctypes.c_double(0.0)
    q6[0] = qsum
    ite = L.ora_satadjust_liq_point(rho, Em, qsum, 1.0, q6.ctypes.data, ctypes.byref(t), ctypes.byref(qv),
                                    ctypes.byref(qc))
    assert ite > 0
    T = t.value
    qs = L.ora_psat_liq(T) / (rho * Rv * T)
    assert T > T0                       # condensation warms
    assert abs(qv.value - qs) <= 1e-8 * abs(qs) * Lv / Rv / T ** 2 * 10 + 1e-18
    assert abs(qv.value + qc.value - qsum) < 1e-18
    # energy: T*CVtot + qv*Lv = Em with CVtot = cv*(1 + qc) (qd_loc = 1), to within
    # the last Newton update (|dT| < 1e-8) times dE/dT
    E = T * cv * (1.0 + qc.value) + qv.value * Lv
    dEdT = cv * (1.0 + qc.value) + Lv * qs * Lv / Rv / T ** 2
    assert abs(E - Em) <= 2e-8 * dEdT


def test_satadjust_unsaturated_point_is_unchanged():
    L = O.lib()
    q6 = np.zeros(6)
    q6[0] = 1e-3
    t = ctypes.c_double(290.0)
    qv = ctypes.c_double(1e-3)
    qc = ctypes.c_double(0.0)
    ite = L.ora_satadjust_liq_point(1.0, 0.0, 1e-3, 1.0, q6.ctypes.data, ctypes.byref(t), ctypes.byref(qv),
                                    ctypes.byref(qc))
    assert ite == 0 and t.value == 290.0 and qv.value == 1e-3 and qc.value == 0.0


def _all_point(rho, Em, qsum, q6, T0, UL=273.16, LL=233.16):
    L = O.lib()
    q6 = np.ascontiguousarray(q6, dtype=np.float64)
    t = ctypes.c_double(T0)
    qv, qc, qi = ctypes.c_double(q6[0]), ctypes.c_double(q6[1]), ctypes.c_double(q6[3])
    ite = L.ora_satadjust_all_point(rho, Em, qsum, 1.0, q6.ctypes.data, UL, LL, ctypes.byref(t), ctypes.byref(qv),
                                    ctypes.byref(qc), ctypes.byref(qi))
    return ite, t.value, qv.value, qc.value, qi.value


def test_satadjust_all_reduces_to_liquid_adjustment_above_the_range():
    """alpha = 1 above the upper limit: the mixed-phase Newton is the liquid one,
    term for term (dalpha/dT = 0, the ice share and its fusion heat vanish)."""
    L = O.lib()
    rho, T0, Rv = 1.0, 290.0, 461.46
    qsum = 1.1 * L.ora_psat_liq(T0) / (rho * Rv * T0)
    Em = T0 * 717.56 + qsum * 2501000.0
    q6 = np.zeros(6)
    q6[0] = qsum
    t = ctypes.c_double(T0)
    qv, qc = ctypes.c_double(qsum), ctypes.c_double(0.0)
    ite_l = L.ora_satadjust_liq_point(rho, Em, qsum, 1.0, q6.ctypes.data, ctypes.byref(t), ctypes.byref(qv),
                                      ctypes.byref(qc))
    ite_a, T, qv_a, qc_a, qi_a = _all_point(rho, Em, qsum, q6, T0)
    assert ite_l > 0 and ite_a == ite_l
    assert T == t.value and qv_a == qv.value and qc_a == qc.value and qi_a == 0.0


def test_satadjust_all_cold_point_deposits_ice():
    """alpha = 0 below the lower limit: all condensate is ice, q(QV) = qsat_ice(T),
    and the moist energy with the fusion term is conserved to the last update."""
    L = O.lib()
    rho, T0, Rv, Lv, Lf, cv = 0.5, 220.0, 461.46, 2501000.0, 333000.0, 717.56
    qsat0 = L.ora_psat_ice(T0) / (rho * Rv * T0)
    qsum = 1.5 * qsat0
    Em = T0 * cv + qsum * Lv            # all water as vapour, no ice yet
    q6 = np.zeros(6)
    q6[0] = qsum
    ite, T, qv, qc, qi = _all_point(rho, Em, qsum, q6, T0)
    assert ite > 0 and T > T0
    assert qc == 0.0
    qs = L.ora_psat_ice(T) / (rho * Rv * T)
    assert abs(qv - qs) <= 1e-7 * qs
    assert abs(qv + qi - qsum) <= 1e-18
    E = T * cv * (1.0 + qi) + qv * Lv - qi * Lf
    dEdT = cv * (1.0 + qi) + (Lv + Lf) * qs * 2834000.0 / Rv / T ** 2
    assert abs(E - Em) <= 2e-8 * dEdT


def test_satadjust_all_mixed_phase_split():
    """Inside the range the condensate splits as alpha : (1 - alpha) with
    alpha = (T - LL)/(UL - LL), and q(QV) is the alpha-blend of the two saturation
    mixing ratios (checked at the final temperature, within the last update)."""
    L = O.lib()
    rho, T0, Rv, UL, LL = 0.8, 255.0, 461.46, 273.16, 233.16
    a0 = (T0 - LL) / (UL - LL)
    qsat0 = (a0 * L.ora_psat_liq(T0) + (1 - a0) * L.ora_psat_ice(T0)) / (rho * Rv * T0)
    qsum = 1.3 * qsat0
    Em = T0 * 717.56 + qsum * 2501000.0
    q6 = np.zeros(6)
    q6[0] = qsum
    ite, T, qv, qc, qi = _all_point(rho, Em, qsum, q6, T0)
    assert ite > 0
    a = (T - LL) / (UL - LL)
    assert abs(qc / (qc + qi) - a) <= 1e-6
    qs = (a * L.ora_psat_liq(T) + (1 - a) * L.ora_psat_ice(T)) / (rho * Rv * T)
    assert abs(qv - qs) <= 1e-6 * qs
    assert abs(qv + qc + qi - qsum) <= 1e-18


def test_satadjust_all_unsaturated_point_is_unchanged():
    q6 = np.zeros(6)
    q6[0] = 1e-6
    ite, T, qv, qc, qi = _all_point(1.0, 0.0, 1e-6, q6, 250.0)
    assert ite == 0 and T == 250.0 and qv == 1e-6 and qc == 0.0 and qi == 0.0


# ---------------------------------------------------------------- Bergeron
def test_bergeron_uses_reference_tables_with_offset_interpolation():
    L = O.lib()
    tabs = _ref_constants()["bergeron"]
    a1t, a2t = np.array(tabs["a1_tab"]), np.array(tabs["a2_tab"])
    out = (ctypes.c_double * 3)()
    for T in (280.0, 273.15, 263.0, 252.5, 240.2, 230.0, 200.0):
        L.ora_bergeron_point(T, ctypes.byref(out, 0), ctypes.byref(out, 8), ctypes.byref(out, 16))
        temc = min(max(T - 273.15, -30.99), 0.0)
        it = int(-temc)
        fact = -(temc + (it - 1))            # reference keeps the Fortran 1-based offset
        assert 1.0 <= fact < 2.0 + 1e-12
        a2 = (1 - fact) * a2t[it] + fact * a2t[it + 1]
        a1 = ((1 - fact) * a1t[it] + fact * a1t[it + 1]) * 1e-3 ** (1 - a2)
        assert abs(out[1] - a2) <= 1e-15 * abs(a2)
        assert abs(out[2] - (1 - a2)) <= 1e-15
        assert abs(out[0] - a1) <= 1e-14 * abs(a1)


# ---------------------------------------------------------------- process rates
def _qsat(T, rho):
    L = O.lib()
    return L.ora_psat_liq(T) / (rho * 461.46 * T), L.ora_psat_ice(T) / (rho * 461.46 * T)


def _point(P, T, rho, q6, dt=15.0):
    qsl, qsi = _qsat(T, rho)
    qd = 1 - sum(q6[1:])
    pre = rho * (qd * 287.04 + q6[0] * 461.46) * T
    return O.bigloop_point(P, dt, rho, T, np.array(q6), qsl, qsi, pre, rho, 50.0, 1.0, 0.5, 0.5)


def test_dry_warm_point_has_no_conversion():
    P = O.params()
    rho, T = 1.1, 290.0
    qsl, _ = _qsat(T, rho)
    o = _point(P, T, rho, [0.5 * qsl, 0, 0, 0, 0, 0])
    assert all(o.drhogq[i] == 0.0 for i in range(6))
    assert o.Vt[2] == 0.0 and o.Vt[4] == 0.0 and o.Vt[5] == 0.0


def test_kk2000_autoconversion_closed_form():
    # saturated warm cloud, no rain: only autoconversion (KK2000 eq. 29) acts on qc
    P = O.params()
    rho, T, dt = 1.0, 290.0, 15.0
    qsl, _ = _qsat(T, rho)
    qc = 5e-4
    o = _point(P, T, rho, [qsl, qc, 0, 0, 0, 0], dt)
    praut = min(1350.0 * qc ** 2.47 * 50.0 ** -1.79, qc / dt)
    assert abs(o.drhogq[2] / (rho * praut * dt) - 1) < 1e-12
    assert abs(o.drhogq[1] + o.drhogq[2]) <= 1e-20


def test_berry_autoconversion_closed_form():
    # sw_kk2000 = 0: Berry (1968) autoconversion in the Berry & Reinhardt (1974) form
    # P = (1/rho) 1.67e-5 (rho qc)^2 / (5 + 3.66e-2 Nc / (Dc rho qc)), rho qc in g m-3,
    # Dc = 0.146 - 5.964e-2 ln(Nc/2000); no rain, so the original Pracw is 0
    P = O.params()
    P.sw_kk2000 = 0.0
    rho, T, dt = 0.9, 290.0, 15.0
    qsl, _ = _qsat(T, rho)
    for qc in (2e-4, 5e-4, 1.5e-3):
        o = _point(P, T, rho, [qsl, qc, 0, 0, 0, 0], dt)
        rhoqc = rho * qc * 1000.0
        Dc = 0.146 - 5.964e-2 * np.log(50.0 / 2000.0)
        praut = min(1.67e-5 * rhoqc ** 2 / rho / (5.0 + 3.66e-2 * 50.0 / (Dc * rhoqc)), qc / dt)
        assert abs(o.drhogq[2] / (rho * praut * dt) - 1) < 1e-12
        assert abs(o.drhogq[1] + o.drhogq[2]) <= 1e-20
        # and it differs from KK2000 at the same point
        P.sw_kk2000 = 1.0
        assert _point(P, T, rho, [qsl, qc, 0, 0, 0, 0], dt).drhogq[2] != o.drhogq[2]
        P.sw_kk2000 = 0.0


def test_rates_conserve_water_and_keep_species_nonnegative():
    P = O.params()
    rng = np.random.default_rng(0)
    for _ in range(300):
        T = rng.uniform(215, 305)
        rho = rng.uniform(0.2, 1.2)
        qsl, _ = _qsat(T, rho)
        q6 = [rng.uniform(0.3, 1.2) * qsl] + list(rng.uniform(0, 1, 5) * [1e-3, 5e-4, 1e-3, 5e-4, 5e-4] *
                                                  (rng.uniform(size=5) < 0.7))
        o = _point(P, T, rho, q6)
        d = np.array(o.drhogq[:])
        assert abs(d.sum()) <= 1e-13 * max(np.abs(d).max(), 1e-30)
        qt = np.array(o.q_t[:])
        for i in range(1, 6):
            # the limiters allow at most complete depletion (up to the rounding of q/dt*dt)
            assert q6[i] + qt[i] * 15.0 >= -4.5e-16 * q6[i]


def test_saturated_point_has_no_rain_evaporation():
    P = O.params()
    rho, T = 1.0, 285.0
    qsl, _ = _qsat(T, rho)
    o = _point(P, T, rho, [qsl, 0, 3e-4, 0, 0, 0])
    assert o.wk[27] == 0.0      # Prevp with Sliq = 1


# ---------------------------------------------------------------- vadv1d
def _prep(ijdim, wp, dz, zh, dt, mkmin=1, mkmax=94):
    zdis = np.zeros((96, ijdim))
    kcell = np.zeros((96, ijdim), dtype=np.int32)
    kmax = np.zeros(96, dtype=np.int32)
    kmin = np.zeros(96, dtype=np.int32)
    O.lib().ora_vadv1d_prep(ijdim, mkmin, mkmax, np.ascontiguousarray(dz).ctypes.data,
                            np.ascontiguousarray(zh).ctypes.data, np.ascontiguousarray(wp).ctypes.data,
                            zdis.ctypes.data, kcell.ctypes.data, kmax.ctypes.data, kmin.ctypes.data, dt)
    return zdis, kcell, kmax, kmin


def _flux(ijdim, rhof, zdis, kcell, kmax, kmin, dz, mkmin=1, mkmax=94):
    f = np.zeros((96, ijdim))
    O.lib().ora_vadv1d_getflux(ijdim, mkmin, mkmax, np.ascontiguousarray(dz).ctypes.data,
                               np.ascontiguousarray(rhof).ctypes.data, zdis.ctypes.data, kcell.ctypes.data,
                               kmax.ctypes.data, kmin.ctypes.data, f.ctypes.data)
    return f


def _grid_arrays():
    gz, gzh = _ref_vgrid()
    g = O.grid(gz, gzh)
    return np.array(g.dgz), gzh


def test_zero_fall_speed_gives_zero_flux():
    dz, zh = _grid_arrays()
    wp = np.zeros((96, 3))
    zdis, kcell, kmax, kmin = _prep(3, wp, dz, zh, 15.0)
    assert np.all(zdis == 0.0)
    assert np.all(kcell[1:95] == np.arange(1, 95)[:, None])
    f = _flux(3, np.ones((96, 3)), zdis, kcell, kmax, kmin, dz)
    assert np.all(f == 0.0)


def test_uniform_fall_speed_analytic_flux():
    dz, zh = _grid_arrays()
    c = 3.0
    wp = np.zeros((96, 2))
    wp[1:95] = -c
    zdis, kcell, kmax, kmin = _prep(2, wp, dz, zh, 15.0)
    np.testing.assert_allclose(zdis[1:95], -c * 15.0, rtol=1e-15)
    rhof = np.random.default_rng(1).uniform(0.1, 1.0, (96, 2))
    f = _flux(2, rhof, zdis, kcell, kmax, kmin, dz)
    np.testing.assert_allclose(f[1:95], -c * 15.0 * rhof[1:95], rtol=1e-15)


def test_flux_divergence_closes_column_budget():
    dz, zh = _grid_arrays()
    rng = np.random.default_rng(2)
    wp = np.zeros((96, 50))
    wp[1:95] = -rng.uniform(0, 12, (94, 50))
    zdis, kcell, kmax, kmin = _prep(50, wp, dz, zh, 15.0)
    rhof = rng.uniform(0, 1e-3, (96, 50))
    f = _flux(50, rhof, zdis, kcell, kmax, kmin, dz)
    dr = -(f[2:96] - f[1:95]) / dz[1:95, None]
    budget = (dr * dz[1:95, None]).sum(axis=0)
    np.testing.assert_allclose(budget, f[1], rtol=1e-10, atol=1e-18)


def _swept_flux_check(wp, sign_expect):
    """Exact swept-mass integral over the departure interval of every wall.  For a
    departure beyond the first cell the reference keeps zdis0 in the partial-cell
    term (the k2 loop assigns zdis instead of accumulating), which shifts the flux by
    rhof(kc) times the distance between the wall and the departure cell's near wall."""
    dz, zh = _grid_arrays()
    n = wp.shape[1]
    zdis, kcell, kmax, kmin = _prep(n, wp, dz, zh, 15.0)
    rhof = np.random.default_rng(5).uniform(0.1, 1.0, (96, n))
    f = _flux(n, rhof, zdis, kcell, kmax, kmin, dz)
    checked = multi = 0
    for j in range(n):
        for k in range(1, 95):
            x = zh[k] - zdis[k, j]
            kc = kcell[k, j]
            if not (zh[kc] <= x <= zh[kc + 1]):
                continue            # departure outside the domain: covered by the exit tests
            if x <= zh[k]:          # upward: mass of [x, zh(k)] crosses the wall upward
                exact = sum(rhof[k2, j] * (min(zh[k2 + 1], zh[k]) - max(zh[k2], x))
                            for k2 in range(kc, k) if min(zh[k2 + 1], zh[k]) > max(zh[k2], x))
                want = exact + rhof[kc, j] * (zh[k] - zh[kc + 1]) if kc < k else exact
            else:                   # downward: mass of [zh(k), x] crosses the wall downward
                exact = -sum(rhof[k2, j] * (min(zh[k2 + 1], x) - zh[k2])
                             for k2 in range(k, kc + 1) if min(zh[k2 + 1], x) > zh[k2])
                want = exact - rhof[kc, j] * (zh[kc] - zh[k])
            if abs(want) < 2.220446e-16:
                want = 0.0
            multi += abs(kc - k) > 0
            checked += 1
            assert np.sign(f[k, j]) in (0.0, sign_expect)
            assert abs(f[k, j] - want) <= 1e-12 * max(abs(want), 1e-12), (k, j, kc)
    assert checked > 0.9 * n * 94 and multi > 0


def test_flux_matches_brute_force_swept_volume_downward():
    n = 30
    c = np.linspace(5.0, 40.0, n)
    k = np.arange(96)[:, None]
    wp = -(c[None, :] * (1.0 + 0.3 * np.sin(k / 10.0)))
    wp[0] = wp[95] = 0.0
    _swept_flux_check(wp, -1.0)


def test_flux_matches_brute_force_swept_volume_upward():
    n = 20
    c = np.linspace(2.0, 25.0, n)
    k = np.arange(96)[:, None]
    wp = c[None, :] * (1.0 + 0.2 * np.cos(k / 15.0))
    wp[0] = wp[95] = 0.0
    _swept_flux_check(wp, 1.0)


def test_departure_cells_bounded_by_region_extremes():
    dz, zh = _grid_arrays()
    rng = np.random.default_rng(4)
    n = 30
    wp = np.zeros((96, n))
    wp[1:95] = -rng.uniform(0, 20, (94, n))
    zdis, kcell, kmax, kmin = _prep(n, wp, dz, zh, 15.0)
    for k in range(1, 95):
        assert kmin[k] <= k <= kmax[k]
        x = zh[k] - zdis[k]
        for j in range(n):
            kc = kcell[k, j]
            assert kmin[k] <= kc <= kmax[k]
            if not (kmin[k] == k and kmax[k] == k):
                assert zh[kc] <= x[j] <= zh[kc + 1]


# ---------------------------------------------------------------- check table format
def test_check_table_line_format():
    from nicam_kokkos_microphysics_amd.check import prof_val_check_line
    line = prof_val_check_line("rhog", np.array([1.0, 2.0]), np.array([0.5, 2.5]))
    assert line == "Checking [rhog] Max = 5.0000000000000000e-01; Min = -5.0000000000000000e-01; " \
                   "Sum = 0.0000000000000000e+00; "
