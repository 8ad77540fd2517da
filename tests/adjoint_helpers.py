"""Shared helpers of the reverse-mode tests: random output cotangents scaled per
field, random input/coefficient directions, and the directional product of a
gradient with a direction."""
import numpy as np

from nicam_kokkos_microphysics_amd import fields as F
from oracle import oracle as O

DT = 15.0
STATE_KEYS = ["rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rhogq", "tem", "q", "vx", "vy", "vz"]


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
