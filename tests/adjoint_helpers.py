"""Shared helpers of the reverse-mode tests: random output cotangents scaled per
field, random input/coefficient directions, and the directional product of a
gradient with a direction."""
import numpy as np

from nicam_kokkos_microphysics_amd import fields as F
from oracle import oracle as O

DT = 15.0
STATE_KEYS = ["rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rhogq", "tem", "q", "vx", "vy", "vz"]


def all_species(st, seed):
    """All five hydrometeors at every level (lognormal, median 3e-5), consistent rhogq."""
    rng = np.random.default_rng(seed)
    s = F.copy_state(st)
    for n in (1, 2, 3, 4, 5):
        s["q"][:, n] += np.exp(rng.normal(np.log(3e-5), 1.5, s["q"][:, n].shape)).clip(0, 2e-3)
    s["rhogq"][:] = s["q"] * s["rhog"][:, None]
    return s


def non_roh(P):
    """The parameters as the reference sets them without Roh & Satoh
    (src/mod_mp_nsw6.cpp:276-356 with Roh_flag false): sw_roh2014 = 0, N0g = 4e6,
    As = pi rho_s / 6, Bs = 3 (and the Gamma functions of it), Esi = 1, Egi = 0.1,
    Egs = 1."""
    import math
    P.sw_roh2014 = 0.0
    P.N0g = 4.0e6
    P.As = math.pi * P.rho_s / 6.0
    P.Bs = 3.0
    P.GAM_1bs = math.gamma(1.0 + P.Bs)
    P.GAM_2bs = math.gamma(2.0 + P.Bs)
    P.GAM_3bs = math.gamma(3.0 + P.Bs)
    P.GAM_1bsds = math.gamma(1.0 + P.Bs + P.Ds)
    P.Esi = 1.0
    P.Egi = 0.1
    P.Egs = 1.0
    return P


def coefs():
    P = O.params()
    return np.array([P.kk_aut_rate, P.qicrt_saut, P.Esi, P.Esr, P.Cr, P.Cs, P.Cg, P.icenuc_coef])


def random_cotangents(st, gz, gzh, seed):
    """Cotangents of every differentiable output, each field scaled by the inverse of
    its typical magnitude after one step, and of the surface mass fluxes."""
    rng = np.random.default_rng(seed)
    out = F.copy_state(st)
    assert O.step_fp64(out, gz, gzh, DT) == 0
    cot = {}
    for k in F.ADJ_OUTPUTS:
        scale = float(np.sqrt(np.mean(out[k] ** 2))) or 1.0
        cot[k] = rng.uniform(-1, 1, out[k].shape) / scale
    lall, ijdim = st["rhog"].shape[0], st["rhog"].shape[-1]
    cs = rng.uniform(-1, 1, (lall, 3, ijdim)) / 1e-3
    return cot, cs


def random_directions(st, seed, n):
    rng = np.random.default_rng(seed)
    c0 = coefs()
    dirs = [{k: st[k] * rng.uniform(-1, 1, st[k].shape) for k in STATE_KEYS} for _ in range(n)]
    dcoef = np.array([c0 * rng.uniform(-1, 1, 8) for _ in range(n)])
    return dirs, dcoef


def directional(grad, gcoef, d, dc):
    """<grad, d> + <gcoef, dc> and the sum of the absolute terms (the scale that
    round-off in the dot product is measured against)."""
    s = float(np.dot(gcoef, dc))
    a = float(np.sum(np.abs(gcoef * dc)))
    for k, g in grad.items():
        if k in d:
            s += float(np.sum(g * d[k]))
            a += float(np.sum(np.abs(g * d[k])))
    return s, a


def output_cotangents(st, gz, gzh, dt, P, seed):
    """random_cotangents for any dt and parameters (the surface fluxes' cotangents
    have 4 rows under four_water: QR, QI, QS, QG)."""
    rng = np.random.default_rng(seed)
    out = F.copy_state(st)
    assert O.step_fp64(out, gz, gzh, dt, P=P) == 0
    cot = {k: rng.uniform(-1, 1, out[k].shape) / (float(np.sqrt(np.mean(out[k] ** 2))) or 1.0)
           for k in F.ADJ_OUTPUTS}
    nt = 4 if P is not None and P.four_water != 0.0 else 3
    return cot, rng.uniform(-1, 1, (st["rhog"].shape[0], nt, st["rhog"].shape[-1])) / 1e-3


def compare_with_host(st, gz, gzh, dt, P, cot, cs, g_gpu, gc_gpu, g_host, gc_host, max_kinks=32):
    """The GPU reverse mode against the host run of the same passes, elementwise
    within 1e-9 of each field's max-abs, and against the host step's dual numbers,
    directionally within 1e-9 of the absolute terms.

    Two kinds of points have no derivative, and there the two builds may differ:
    where a tracer's input rhogq is exactly 0 (the rain fall speed goes as
    (rho q)^(1/8) behind the zero switch and the value after the rates is round-off),
    left out; and the kinks of the rates' limiters, where the last bits of each
    build's arithmetic decide the side: at each (at most max_kinks entries) the GPU's
    value must be one of the host step's one-sided difference quotients and the
    host's the other, and the directions vanish there.  Returns the kink count."""
    def J(s):
        j, _ = O.dual_vjp(s, gz, gzh, dt, cot, cs, [{}], np.zeros((1, 8)), P=P)
        return j

    kinks = []
    for k in F.ADJ_INPUTS:
        keep = (st[k] != 0) if k == "rhogq" else np.ones(g_host[k].shape, dtype=bool)
        scale = np.abs(g_host[k][keep]).max()
        off = (np.abs(g_gpu[k] - g_host[k]) > 1e-9 * scale) & keep
        kinks += [(k, tuple(int(v) for v in i)) for i in np.argwhere(off)]
    assert len(kinks) <= max_kinks, len(kinks)
    j0 = J(st) if kinks else None
    for k, i in kinks:
        x0 = st[k][i]
        h = 1e-7 * abs(x0) or 1e-12
        sp, sm = F.copy_state(st), F.copy_state(st)
        sp[k][i] = x0 + h
        sm[k][i] = x0 - h
        right, left = (J(sp) - j0) / h, (j0 - J(sm)) / h
        # a kink: the one-sided quotients differ well beyond their rounding error
        # (about 1e-6 relative at this h), and each build sits on one side, within a
        # tenth of the gap (a quotient also carries h times the curvature, up to a few
        # 1e-3 relative next to a kink)
        gap = abs(right - left)
        assert gap > 1e-5 * max(abs(right), abs(left)), (k, i, right, left)
        near = lambda a, b: abs(a - b) <= 0.1 * gap
        assert (near(g_gpu[k][i], right) and near(g_host[k][i], left)) or \
               (near(g_gpu[k][i], left) and near(g_host[k][i], right)), \
            (k, i, g_gpu[k][i], g_host[k][i], left, right)
    np.testing.assert_allclose(gc_gpu, gc_host, rtol=1e-9, atol=0)
    dirs, dcoef = random_directions(st, 17, 4)
    rng = np.random.default_rng(4)
    for key in STATE_KEYS[:4]:
        dirs.append({key: st[key] * rng.uniform(-1, 1, st[key].shape)})
    dcoef = np.concatenate([dcoef, np.zeros((4, 8))])
    pts = {(i[0],) + i[-2:] for _, i in kinks}
    for d in dirs:
        for v in d.values():
            for p in pts:
                if v.ndim == 3:
                    v[p] = 0.0
                else:
                    v[p[0], :, p[1], p[2]] = 0.0
    _, dj = O.dual_vjp(st, gz, gzh, dt, cot, cs, dirs, dcoef, P=P)
    for i in range(len(dirs)):
        s, a = directional(g_gpu, gc_gpu, dirs[i], dcoef[i])
        assert abs(s - dj[i]) <= 1e-9 * a, (i, s, dj[i], a)
    return len(kinks)
