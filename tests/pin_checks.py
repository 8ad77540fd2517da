"""Pins of the oracle's step phases, each written as a function of a loaded oracle build
`L` (oracle.load / oracle.lib) that raises AssertionError when the pin fails.

tests/test_oracle_pins_step.py runs them on the oracle (they must pass);
tests/test_oracle_plants.py runs each on a build of the oracle with a mistake planted
(ORA_PLANT = n) and requires it to fail.  What fixes each pin, independently of the
oracle's own text:

* THRMDYN (src/mod_thrmdyn.cpp:98-180): qd is the dry fraction with QV counted as part
  of the gas (qd + qc + qr + qi + qs + qg = 1, independent of qv); under EIN_TYPE
  SIMPLE every heat capacity is CVdry, so cv = CVdry; T = ein / cv; the pressure is the
  sum of the dry-air and vapour partial pressures (Dalton).
* negative_filter (src/mod_mp_nsw6.cpp:1750-1832): total water is conserved while the
  vapour stays non-negative; QR..QG end non-negative, QC is not filtered; where the
  vapour would go negative the density changes by the vapour deficit; the filter is the
  identity on non-negative states; halo levels are untouched.
* cnvvar_rhogkin_in (src/mod_debug.cpp:269-330): the kinetic energies of a stratified
  state from linear interpolation in height (cell values to walls for the wall density,
  wall values to the cell for the vertical kinetic energy), derived from gz/gzh alone.
* precip_transport_new (src/mod_precip_transport.cpp:573-1087): column budgets.  Per
  tracer and in total, the column mass changes by the surface precipitation; with
  column-uniform winds the column momentum changes by that mass flux times the wind;
  with w = 0 the column total energy (internal + horizontal kinetic + potential) changes
  by the surface energy fluxes (KIN2EIN turns the kinetic energy the transport loses
  into internal energy).
* SATURATION_adjustment (src/mod_satadjust.cpp:274-362): moist internal energy
  e + qv LHV is conserved; unsaturated interior points are unchanged; on the halo
  levels (no Newton) the condensate is evaporated; the internal energy is rebuilt as
  rhog T cv.
* Process rates (src/mod_mp_nsw6.cpp:930-1247), at single-species points with dt small
  enough that no limiter binds: Praci, Pgacr and Pgfrz, and the size integral of Prevp,
  as numerical quadratures of the physical integrals over the exponential size
  distributions (the slope from the mass content); Psacw and Psmlt from the snow
  moments of Field et al. (2005), evaluated from the published polynomial with the
  reference's coefficient arrays (tests/golden/reference_constants.json).
"""
import json
import math
import os

import numpy as np
from scipy import integrate

from nicam_kokkos_microphysics_amd import fields as F
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "reference_constants.json")) as _fh:
    _REF = json.load(_fh)
C = _REF["constants"]
CVdry, Rdry, Rvap = C["CONST_CVdry"], C["CONST_Rdry"], C["CONST_Rvap"]
LHV0, LHF0, GRAV, TEM00 = C["CONST_LHV0"], C["CONST_LHF0"], C["CONST_GRAV"], C["CONST_TEM00"]
PI = math.pi
DENS00 = 1.28                       # src/mod_mp_nsw6.cpp:8
KA0, DKA_DT = 2.428e-2, 7.47e-5     # src/mod_mp_nsw6.cpp:120-121
KD0, DKD_DT = 2.222e-5, 1.37e-7     # :122-123
NU0, DNU_DT = 1.718e-5, 5.28e-8     # :124-125
PRE00 = C["CONST_Pstd"]             # :471

def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# --------------------------------------------------------------------- THRMDYN
def check_thrmdyn(L):
    rng = np.random.default_rng(1)
    ij = 64
    q = np.zeros((6, 96, ij))
    q[0] = rng.uniform(0, 0.02, (96, ij))
    q[1:] = rng.uniform(0, 2e-3, (5, 96, ij))
    ein = rng.uniform(1.5e5, 2.3e5, (96, ij))
    rho = rng.uniform(0.2, 1.3, (96, ij))
    qd, cv, tem, pre = O.thrmdyn(q, ein, rho, L=L)
    # qd + condensates = 1, whatever qv is
    assert np.abs(qd + q[1:].sum(axis=0) - 1.0).max() < 1e-15, "qd is not 1 - (qc+qr+qi+qs+qg)"
    q2 = q.copy()
    q2[0] *= 0.5
    assert np.array_equal(O.thrmdyn(q2, ein, rho, L=L)[0], qd), "qd depends on qv"
    assert np.abs(cv / CVdry - 1.0).max() < 1e-14, "cv != CVdry under SIMPLE"
    assert np.abs(tem * cv / ein - 1.0).max() < 1e-14
    p_dry, p_vap = rho * qd * Rdry * tem, rho * q[0] * Rvap * tem
    assert np.abs(pre / (p_dry + p_vap) - 1.0).max() < 1e-14, "pressure is not p_dry + p_vap"


# --------------------------------------------------------------------- negative_filter
def _filter_state(seed=2, ij=400):
    gz, gzh, st = F.synth_state(1, ij, seed=seed)
    s = {k: np.ascontiguousarray(st[k][0]) for k in ("rhog", "rhoge", "rhogq", "rho", "tem", "pre", "q",
                                                      "gsgam2")}
    return s


def check_negative_filter(L):
    rng = np.random.default_rng(3)
    s = _filter_state()
    rq = s["rhogq"]
    # negative hydrometeors (QR..QG) at a third of the interior points, negative QC too
    for n in range(1, 6):
        m = rng.uniform(size=rq[n].shape) < 0.3
        rq[n][m] = -rng.uniform(0, 2e-5, m.sum())
    # vapour pushed below the compensation at a few points
    deep = rng.uniform(size=rq[0].shape) < 0.02
    rq[0][deep] = rng.uniform(0, 1e-5, deep.sum())
    s["q"][:] = rq / s["rhog"]
    before = {k: v.copy() for k, v in s.items()}
    O.negative_filter(s["rhog"], s["rhoge"], s["rhogq"], s["rho"], s["tem"], s["pre"], s["q"], s["gsgam2"], L=L)
    I = slice(1, 95)
    b, a = before["rhogq"][:, I], s["rhogq"][:, I]
    assert (a[2:] >= 0).all(), "QR..QG not filtered"
    assert np.array_equal(a[1], b[1]), "QC must pass unfiltered (negative QC kept)"
    comp = b[0] + (b[2:] - np.maximum(b[2:], 0)).sum(axis=0)    # vapour after compensation
    ok = comp >= 0
    tot_b, tot_a = b.sum(axis=0), a.sum(axis=0)
    assert (np.abs(tot_a - tot_b)[ok] <= 1e-15 * np.abs(b).sum(axis=0)[ok]).all(), "water not conserved"
    assert (a[0][~ok] == 0).all()
    # rhog changes by the vapour deficit removed: rhog' = rhog (1 - (comp - max(comp, 0)))
    d = comp - np.maximum(comp, 0)
    assert np.abs(s["rhog"][I] / (before["rhog"][I] * (1 - d)) - 1).max() < 1e-15, "rhog change"
    assert (~ok).sum() > 0 and (d < 0).any()
    for k in ("rhog", "rhogq", "rhoge"):   # halo levels untouched
        assert np.array_equal(s[k][..., 0, :], before[k][..., 0, :]) and \
            np.array_equal(s[k][..., 95, :], before[k][..., 95, :]), k
    # identity on a non-negative state (the filtered one): QR..QG unchanged; QV and rhog
    # to the round-off of the filter's own sum (it adds and subtracts the hydrometeors
    # term by term)
    again = {k: v.copy() for k, v in s.items()}
    O.negative_filter(again["rhog"], again["rhoge"], again["rhogq"], again["rho"], again["tem"], again["pre"],
                      again["q"], again["gsgam2"], L=L)
    assert np.array_equal(again["rhogq"][1:], s["rhogq"][1:]), "not idempotent: QC..QG"
    hsum = np.abs(s["rhogq"]).sum(axis=0)
    assert (np.abs(again["rhogq"][0] - s["rhogq"][0]) <= 4e-16 * hsum).all(), "not idempotent: QV"
    assert np.abs(again["rhog"] / s["rhog"] - 1).max() < 1e-15, "not idempotent: rhog"


# --------------------------------------------------------------------- rhogkin
def check_rhogkin(L):
    # a grid whose cell centres sit at 30 % of each layer (not mid-layer), so that the
    # two interpolation weights of every wall and cell differ
    _, gzh = F.synth_grid()
    dz = np.diff(gzh, append=2 * gzh[-1] - gzh[-2])
    gz = gzh + 0.3 * dz
    ij = 8
    k = np.arange(96)[:, None] * np.ones((1, ij))
    rhog = 1.2 * np.exp(-gz[:, None] / 8000.0) * (1 + 0.01 * np.arange(ij))[None, :]
    u, v = 12.0, -7.0
    w = 1.0 + 0.5 * np.sin(0.37 * k) + 0.1 * np.arange(ij)[None, :]
    one = np.ones((96, ij))
    # wall density by linear interpolation in height between the cell centres k-1 and k
    a = ((gzh[1:] - gz[:-1]) / (gz[1:] - gz[:-1]))[:, None]
    rhoh = np.zeros((96, ij))
    rhoh[1:] = a * rhog[1:] + (1 - a) * rhog[:-1]
    rhogw = w * rhoh
    rhogkin, kh, kv = O.rhogkin(gz, gzh, rhog, rhog * u, rhog * v, np.zeros_like(rhog), rhogw, one, one, L=L)
    I = slice(1, 95)
    assert np.abs(kh[I] / (0.5 * rhog[I] * (u * u + v * v)) - 1).max() < 1e-14
    kv_want = 0.5 * rhoh * w * w
    W = slice(2, 95)
    assert np.abs(kv[W] / kv_want[W] - 1).max() < 1e-13, "vertical kinetic energy at the walls"
    assert (kv[[0, 1, 95]] == 0).all()
    # the cell's vertical kinetic energy: linear interpolation from its walls to gz
    c = ((gz[:-1] - gzh[:-1]) / (gzh[1:] - gzh[:-1]))[:, None]
    kv_cell = c[1:95] * kv_want[2:96] + (1 - c[1:95]) * kv_want[1:95]
    kv_cell[-1] = c[94] * 0.0 + (1 - c[94]) * kv_want[94]   # wall 95 carries none
    kv_cell[0] = c[1] * kv_want[2]                          # wall 1 carries none
    assert np.abs(rhogkin[I] - (kh[I] + kv_cell)).max() < 1e-13 * rhogkin[I].max(), "W2C weights"


# --------------------------------------------------------------------- transport budgets
def _transport_case(uniform_winds, seed=4, ij=300):
    rng = np.random.default_rng(seed)
    gz, gzh, st = F.synth_state(1, ij, seed=seed)
    if uniform_winds:
        for n in ("vx", "vy", "vz"):
            st[n][0] = st[n][0][40:41, :]          # the wind of level 40, at every level
    for n, c in (("rhogvx", "vx"), ("rhogvy", "vy"), ("rhogvz", "vz")):
        st[n][0] = st["rhog"][0] * st[c][0]
    st["rhogw"][:] = 0.0
    # snow and graupel down to the ground, so that both precipitation rows are nonzero
    for n, qq in ((4, 2e-4), (5, 3e-4)):
        st["rhogq"][0, n, 1:6] = st["rhog"][0, 1:6] * qq
        st["q"][0, n, 1:6] = qq
    V = np.zeros((6, 96, ij))
    V[2, 1:95] = -rng.uniform(2.0, 5.0, (94, ij))   # QR
    V[4, 1:95] = -rng.uniform(0.5, 1.5, (94, ij))   # QS
    V[5, 1:95] = -rng.uniform(1.0, 4.0, (94, ij))   # QG
    return gz, gzh, st, V


def _column(x, dgz):
    return (x[:, 1:95, :] * dgz[None, 1:95, None]).sum(axis=1)


def check_transport_budgets(L, dt=15.0):
    gz, gzh = F.synth_grid()
    dgz = np.diff(gzh, append=gzh[-1] + (gzh[-1] - gzh[-2]))
    for uniform in (True, False):
        _, _, st, V = _transport_case(uniform)
        b = F.copy_state(st)
        O.precip_transport(st, gz, gzh, V, dt, L=L)
        p0, p1 = st["precip"][:, 0], st["precip"][:, 1]
        scale = _column(np.abs(b["rhogq"][:, 2]) + np.abs(b["rhogq"][:, 4]) + np.abs(b["rhogq"][:, 5]), dgz)
        dmass = _column(st["rhog"] - b["rhog"], dgz)
        assert np.abs(dmass + dt * (p0 + p1)).max() < 1e-12 * scale.max(), "column mass budget"
        dqr = _column(st["rhogq"][:, 2] - b["rhogq"][:, 2], dgz)
        assert np.abs(dqr + dt * p0).max() < 1e-12 * scale.max(), "rain budget"
        dsol = _column(st["rhogq"][:, 4] + st["rhogq"][:, 5] - b["rhogq"][:, 4] - b["rhogq"][:, 5], dgz)
        assert np.abs(dsol + dt * p1).max() < 1e-12 * scale.max(), "snow + graupel budget"
        assert (p0 > 0).any() and (p1 > 0).any()
        if uniform:
            ix, jx = st["ix"], st["jx"]
            up = st["vx"][:, 40] * st["ix"] + st["vy"][:, 40] * st["iy"] + st["vz"][:, 40] * st["iz"]
            vp = st["vx"][:, 40] * st["jx"] + st["vy"][:, 40] * st["jy"] + st["vz"][:, 40] * st["jz"]
            for n, i_, j_ in (("rhogvx", "ix", "jx"), ("rhogvy", "iy", "jy"), ("rhogvz", "iz", "jz")):
                dmom = _column(st[n] - b[n], dgz)
                want = -dt * (p0 + p1) * (st[i_] * up + st[j_] * vp)
                mscale = _column(np.abs(b[n]), dgz).max()
                assert np.abs(dmom - want).max() < 1e-13 * mscale, f"column momentum budget {n}"
        else:
            def etot(s):
                kin = 0.5 * (s["rhogvx"] ** 2 + s["rhogvy"] ** 2 + s["rhogvz"] ** 2) / s["rhog"]
                return s["rhoge"] + kin + s["rhog"] * GRAV * s["z"]
            de = _column(etot(st) - etot(b), dgz)
            want = -dt * (st["precip_rhoe"] + st["precip_rhophi"] + st["precip_rhokin"])
            escale = _column(np.abs(etot(b)), dgz).max()
            assert np.abs(de - want).max() < 1e-13 * escale, "column total energy budget"
            # the budget is not trivial: KIN2EIN moves kinetic energy into rhoge here
            kin_loss = _column(0.5 * (b["rhogvx"] ** 2 + b["rhogvy"] ** 2) / b["rhog"], dgz)
            assert np.abs(want).max() > 1e-10 * escale and kin_loss.max() > 0


# --------------------------------------------------------------------- saturation adjustment
def check_saturation_wrapper(L):
    gz, gzh, st = F.synth_state(1, 300, seed=6, frac_supersat=0.3)
    s = {k: np.ascontiguousarray(st[k][0]) for k in ("rhog", "rhoge", "rhogq", "tem", "q", "qd", "gsgam2")}
    q = s["q"]
    q[1:] = s["rhogq"][1:] / s["rhog"]
    q[1, [0, 95]] = 1e-4                    # condensate on the halo levels
    s["rhogq"][1, [0, 95]] = s["rhog"][[0, 95]] * 1e-4
    s["qd"][:] = 1.0 - q[1:].sum(axis=0)
    s["rhoge"][:] = s["rhog"] * s["tem"] * CVdry
    b = {k: v.copy() for k, v in s.items()}
    O.saturation_adjustment(s["rhog"], s["rhoge"], s["rhogq"], s["tem"], s["q"], s["qd"], s["gsgam2"], L=L)
    e0 = b["rhoge"] / b["rhog"] + b["q"][0] * LHV0
    e1 = s["rhoge"] / s["rhog"] + s["q"][0] * LHV0
    # Moist internal energy e + qv LHV: where no Newton runs it is conserved.  Where the
    # Newton runs it solves T CVdry (1 + qc) + qsat LHV = e + qv LHV (its heat capacity
    # counts the new condensate on top of a total without the input qc), while the
    # rebuild rhoge = rhog T cv uses the caller's qd, which excludes the input qc: one
    # call loses exactly T CVdry qc_in.  The Newton stops at |dtemp| < 1e-8 K with q(QV)
    # one iterate behind, so the rest agrees within (cv + LHV dqsat/dT) 1e-8 << 1e-3.
    newton = s["q"][1] > 0
    want = e0 - np.where(newton, s["tem"] * CVdry * b["q"][1], 0.0)
    assert np.abs(e1 - want).max() < 1e-3, "moist internal energy relation"
    assert (newton & (b["q"][1] > 0)).sum() > 100
    rho = b["rhog"] / b["gsgam2"]
    qsat = np.array([[L.ora_psat_liq(t) for t in row] for row in b["tem"]]) / (rho * Rvap * b["tem"])
    unsat = (b["q"][0] + b["q"][1] < qsat) & (b["q"][1] == 0)
    unsat[[0, 95]] = False
    assert unsat.sum() > 100 and (~unsat[1:95]).sum() > 10
    assert np.abs(s["tem"] - b["tem"])[unsat].max() < 1e-9, "unsaturated point changed"
    assert np.abs(s["q"][0] - (b["q"][0] + b["q"][1]))[unsat].max() < 1e-18
    for k in (0, 95):   # halo levels: the condensate evaporated, no Newton
        assert np.array_equal(s["q"][1, k], np.zeros(300)), "halo condensate not evaporated"
        assert np.abs(s["q"][0, k] - (b["q"][0, k] + b["q"][1, k])).max() < 1e-18
    cv = CVdry * (s["qd"] + s["q"][1:].sum(axis=0))
    assert np.abs(s["rhoge"] / (s["rhog"] * s["tem"] * cv) - 1).max() < 1e-14


# --------------------------------------------------------------------- process rates
def _point(L, P, T, rho, q6, dt=1e-3, sliq=None):
    qsl = L.ora_psat_liq(T) / (rho * Rvap * T)
    qsi = L.ora_psat_ice(T) / (rho * Rvap * T)
    q6 = list(q6)
    if sliq is not None:
        q6[0] = sliq * qsl
    qd = 1 - sum(q6[1:])
    pre = rho * (qd * Rdry + q6[0] * Rvap) * T
    o = O.bigloop_point(P, dt, rho, T, np.array(q6), qsl, qsi, pre, rho, 50.0, 1.0, 0.5, 0.5, L=L)
    return o, qsl, pre


def _slope(rho, q, A, N0, B):
    """1 / lambda of an exponential distribution N0 exp(-D lambda) whose mass
    integral of A D^B equals rho q."""
    return (rho * q / (A * N0 * math.gamma(1 + B))) ** (1.0 / (1 + B))


def _quad(f):
    v, err = integrate.quad(f, 0.0, np.inf, epsabs=0.0, epsrel=1e-13, limit=400)
    return v


def _field_moment(n, Tc, M2):
    """M_n = a(n, Tc) M2^b(n, Tc) of Field et al. (2005) with Tc clipped to <= -0.1 C:
    log10 a = a1 + a2 T + a3 n + a4 T n + a5 T^2 + a6 n^2 + a7 T^2 n + a8 T n^2 + a9 T^3
    + a10 n^3 (b alike), coefficient arrays of the reference."""
    ca, cb = _REF["roh_coefficients"]["coef_a"], _REF["roh_coefficients"]["coef_b"]
    T = min(-0.1, Tc)

    def poly(c):
        return (c[0] + c[1] * T + c[2] * n + c[3] * T * n + c[4] * T * T + c[5] * n * n + c[6] * T * T * n +
                c[7] * T * n * n + c[8] * T ** 3 + c[9] * n ** 3)
    return 10.0 ** poly(ca) * M2 ** poly(cb)


def check_praci(L):
    P = O.params()
    T, rho, qi, qr = 260.0, 0.8, 1e-4, 5e-5
    o, _, _ = _point(L, P, T, rho, [0, 0, qr, qi, 0, 0])
    R = _slope(rho, qr, P.Ar, P.N0r, P.Br)
    rf = math.sqrt(DENS00 / rho)
    sweep = _quad(lambda D: 0.25 * PI * D * D * P.Cr * D ** P.Dr * rf * P.N0r * math.exp(-D / R))
    assert _rel(o.wk[I_Praci], qi * P.Eri * sweep) < 1e-10, ("Praci", o.wk[I_Praci], qi * P.Eri * sweep)


def check_pgacr(L):
    P = O.params()
    T, rho, qr, qg = 265.0, 0.9, 5e-5, 2e-4
    o, _, _ = _point(L, P, T, rho, [0, 0, qr, 0, 0, qg])
    Rr, Rg = _slope(rho, qr, P.Ar, P.N0r, P.Br), _slope(rho, qg, P.Ag, P.N0g, P.Bg)
    rf = math.sqrt(DENS00 / rho)

    def mass_mean_speed(C_, D_, A_, B_, R_):
        num = _quad(lambda D: A_ * D ** B_ * C_ * D ** D_ * rf * math.exp(-D / R_))
        return num / _quad(lambda D: A_ * D ** B_ * math.exp(-D / R_))
    Vtr = -mass_mean_speed(P.Cr, P.Dr, P.Ar, P.Br, Rr)
    Vtg = -mass_mean_speed(P.Cg, P.Dg, P.Ag, P.Bg, Rg)
    mr = [_quad(lambda D, j=j: D ** j * P.Ar * D ** P.Br * P.N0r * math.exp(-D / Rr)) for j in range(3)]
    ng = [_quad(lambda D, j=j: D ** j * P.N0g * math.exp(-D / Rg)) for j in range(3)]
    # (Dg + Dr)^2 = Dg^2 + 2 Dg Dr + Dr^2 over the product of the two distributions
    sweep = ng[2] * mr[0] + 2 * ng[1] * mr[1] + ng[0] * mr[2]
    want = (1 / rho) * 0.25 * PI * P.Egr * abs(Vtg - Vtr) * sweep
    assert _rel(o.wk[I_Pgacr], want) < 1e-10, ("Pgacr", o.wk[I_Pgacr], want)


def check_pgfrz(L):
    P = O.params()
    T, rho, qr = 263.0, 0.9, 2e-4
    o, _, _ = _point(L, P, T, rho, [0, 0, qr, 0, 0, 0], sliq=1.0)
    R = _slope(rho, qr, P.Ar, P.N0r, P.Br)
    temc = T - TEM00
    # Bigg freezing: each drop freezes at B (exp(-A Tc) - 1) per unit volume and time
    J = P.B_frz * (math.exp(-P.A_frz * temc) - 1.0)
    want = (1 / rho) * _quad(lambda D: P.Ar * D ** P.Br * (PI / 6) * D ** 3 * J * P.N0r * math.exp(-D / R))
    assert _rel(o.wk[I_Pgfrz], want) < 1e-10, ("Pgfrz", o.wk[I_Pgfrz], want)


def check_prevp(L):
    P = O.params()
    T, rho, qr, S = 285.0, 1.0, 3e-4, 0.7
    o, qsl, pre = _point(L, P, T, rho, [0, 0, qr, 0, 0, 0], sliq=S)
    R = _slope(rho, qr, P.Ar, P.N0r, P.Br)
    temc = T - TEM00
    rf = math.sqrt(DENS00 / rho)
    Ka, Kd, nu = KA0 + DKA_DT * temc, (KD0 + DKD_DT * temc) * PRE00 / pre, (NU0 + DNU_DT * temc) / rho
    G = 1.0 / (LHV0 / (Ka * T) * (LHV0 / (Rvap * T) - 1.0) + 1.0 / (Kd * rho * qsl))
    # ventilated diffusional growth of each drop: 2 pi D (f1 + f2 sqrt(V D / nu))
    vent = _quad(lambda D: D * (P.f1r + P.f2r * math.sqrt(P.Cr * D ** P.Dr * rf * D / nu)) * P.N0r *
                 math.exp(-D / R))
    want = 2 * PI / rho * (1 - S) * G * vent
    assert _rel(o.wk[I_Prevp], want) < 1e-10, ("Prevp", o.wk[I_Prevp], want)


def check_psacw(L):
    P = O.params()
    T, rho, qc, qs = 263.0, 0.9, 2e-4, 3e-4
    o, _, _ = _point(L, P, T, rho, [0, qc, 0, 0, qs, 0], sliq=1.0)
    M = _field_moment(2.0 + P.Ds, T - TEM00, rho * qs / P.As)
    want = qc * 0.25 * PI * P.Esw * P.Cs * M * math.sqrt(DENS00 / rho)
    assert _rel(o.wk[I_Psacw], want) < 1e-10, ("Psacw", o.wk[I_Psacw], want)


def check_psmlt(L):
    P = O.params()
    T, rho, qs = 275.0, 1.0, 2e-4
    o, _, _ = _point(L, P, T, rho, [0, 0, 0, 0, qs, 0], sliq=0.9)
    temc = T - TEM00
    M2 = rho * qs / P.As
    nu = (NU0 + DNU_DT * temc) / rho
    vents = P.f1s * _field_moment(1.0, temc, M2) + \
        P.f2s * math.sqrt(P.Cs * math.sqrt(DENS00 / rho) / nu) * _field_moment((3.0 + P.Ds) / 2, temc, M2)
    Ka = KA0 + DKA_DT * temc
    # melting by conduction: heat flux Ka Tc over the fusion heat
    want = 2 * PI / rho * (Ka * temc / LHF0) * vents
    assert _rel(o.wk[I_Psmlt], want) < 1e-10, ("Psmlt", o.wk[I_Psmlt], want)


def _wk_index(name):
    """Position of a rate in the oracle's wk[] enum (parsed from its source)."""
    import re
    src = open(os.path.join(os.path.dirname(O.__file__), "nsw6_oracle.c")).read()
    body = re.search(r"enum \{(.*?)\};", src, re.S).group(1)
    names = [t.strip() for t in re.sub(r"/\*.*?\*/", "", body, flags=re.S).replace("\n", " ").split(",")]
    return names.index(name)


I_Psmlt, I_Pgfrz, I_Pgacr, I_Praci, I_Psacw, I_Prevp = (_wk_index(n) for n in ("I_Psmlt", "I_Pgfrz", "I_Pgacr",
                                                                               "I_Praci", "I_Psacw", "I_Prevp"))

# pin name -> (check, the plants it must catch)
PINS = {
    "thrmdyn": (check_thrmdyn, (4, 5)),
    "negative_filter": (check_negative_filter, (1, 2, 3)),
    "rhogkin": (check_rhogkin, (6, 7)),
    "transport_budgets": (check_transport_budgets, (8, 9, 10, 11)),
    "saturation_wrapper": (check_saturation_wrapper, (12, 13)),
    "praci": (check_praci, (14,)),
    "psacw": (check_psacw, (15,)),
    "pgacr": (check_pgacr, (16,)),
    "prevp": (check_prevp, (17,)),
    "psmlt": (check_psmlt, (18,)),
    "pgfrz": (check_pgfrz, (19,)),
}
