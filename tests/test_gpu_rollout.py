"""Multi-step rollout with step-level checkpoints and the reverse mode over all steps
(nicam_kokkos_microphysics_amd.rollout).

* Checkpointing changes nothing: the gradient with a checkpoint every 3 steps is
  bitwise that with a checkpoint at every step, and region chunking of the reverse
  mode (1 region at a time vs all) is bitwise neutral.
* The FP64 rollout gradient matches central differences of the FP64 rollout loss
  (h = 1e-8 relative) along random directions through every input field and the
  coefficients, for the directions whose stencil crosses no limiter switch (the
  same quotient at h and 2h); the host chain of the same reverse passes gives the
  same gradient, and at h = 1e-7 one of these directions crosses a switch.
* The mixed-precision rollout recomputes its own trajectory exactly (the backward
  checks every recomputed step against the forward) and gives a finite gradient
  within a few percent of the FP64 rollout's in direction.
"""
import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from nicam_kokkos_microphysics_amd import fields as F

from adjoint_helpers import coefs, random_directions

pytestmark = pytest.mark.gpu
DT = 15.0


def _run(st, gz, gzh, steps, every, chunk, precision=N.NSW6_FP64, params=None):
    import torch
    from nicam_kokkos_microphysics_amd.rollout import Rollout
    lall, ijdim = st["rhog"].shape[0], st["rhog"].shape[-1]
    ro = Rollout(gz, gzh, lall, ijdim, precision=precision, chunk_regions=chunk, params=params)
    dev = F.to_device(st)
    rhoge0 = dev["rhoge"].clone()
    print(f"rollout lall={lall} ijdim={ijdim} steps={steps} every={every} chunk={chunk}", flush=True)
    ckpt, precip = ro.forward(dev, DT, steps, every)
    L = ro.loss(rhoge0, dev, precip, DT)
    grad, gc = ro.backward(dev, ckpt, precip, DT, steps, every)
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in grad.items()}
    g = gc.cpu().numpy()
    ro.close()
    return L, out, g


def test_checkpointing_and_chunking_are_exact():
    gz, gzh, st = F.synth_state(3, 200, seed=9, frac_towers=0.3)
    _, a, ga = _run(st, gz, gzh, 6, 1, 3)
    _, b, gb = _run(st, gz, gzh, 6, 3, 3)
    _, c, gc = _run(st, gz, gzh, 6, 3, 1)
    for k in F.ADJ_INPUTS:
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a[k], c[k]), k
    assert np.array_equal(ga, gb)
    np.testing.assert_allclose(gc, ga, rtol=1e-13)   # per-chunk coefficient sums differ in order


def test_fp64_rollout_gradient_matches_central_differences():
    """Directions whose difference stencil crosses a limiter switch somewhere along the
    6-step trajectory show it as a difference quotient that changes with h; those are
    skipped (at least 2 of 4 must be switch-free) and the rest must agree to 1e-5."""
    from nicam_kokkos_microphysics_amd.driver import default_params
    gz, gzh, st = F.synth_state(2, 128, seed=5, frac_supersat=0.2)
    steps, every = 6, 3
    _, g, gcoef = _run(st, gz, gzh, steps, every, 2)
    dirs, dcoef = random_directions(st, 41, 4)
    c0 = coefs()

    def fd(i, h):
        vals = []
        for sgn in (1.0, -1.0):
            P = default_params()
            for c, name in enumerate(N.TUNABLE_COEFFS):
                setattr(P, name, c0[c] + sgn * h * dcoef[i][c])
            sp = {k: (st[k] + sgn * h * dirs[i][k] if k in dirs[i] else st[k]) for k in st}
            vals.append(_loss_only(sp, gz, gzh, steps, P))
        return (vals[0] - vals[1]) / (2 * h)

    checked = 0
    for i in range(len(dirs)):
        f1, f2 = fd(i, 1e-8), fd(i, 2e-8)
        if abs(f1 - f2) > 1e-4 * abs(f1):
            continue
        s = float(np.dot(gcoef, dcoef[i])) + sum(float(np.sum(g[k] * dirs[i][k])) for k in dirs[i])
        assert abs(s - f1) <= 1e-5 * abs(f1), (i, s, f1)
        checked += 1
    assert checked >= 2, checked


def _loss_only(st, gz, gzh, steps, P):
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6
    lall, ijdim = st["rhog"].shape[0], st["rhog"].shape[-1]
    drv = NSW6(gz, gzh, params=P, max_columns=lall * ijdim)
    dev = F.to_device(st)
    rhoge0 = dev["rhoge"].clone()
    s = 0.0
    for _ in range(steps):
        drv.step(dev, DT)
        s += float((dev["precip"] ** 2).sum())
    s += float(((dev["rhoge"] - rhoge0) / (steps * DT)).sum())
    torch.cuda.synchronize()
    drv.close()
    return s


def test_mixed_rollout_recomputes_exactly_and_tracks_fp64_gradient():
    gz, gzh, st = F.synth_state(2, 256, seed=9, frac_towers=0.1)
    L64, g64, c64 = _run(st, gz, gzh, 6, 3, 2, precision=N.NSW6_FP64)
    Lmx, gmx, cmx = _run(st, gz, gzh, 6, 3, 2, precision=N.NSW6_MIXED)
    assert abs(Lmx - L64) <= 1e-4 * abs(L64)
    for k in F.ADJ_INPUTS:
        assert np.all(np.isfinite(gmx[k])), k
    v64 = np.concatenate([g64[k].ravel() for k in F.ADJ_INPUTS])
    vmx = np.concatenate([gmx[k].ravel() for k in F.ADJ_INPUTS])
    cos = float(np.dot(v64, vmx) / (np.linalg.norm(v64) * np.linalg.norm(vmx)))
    assert cos >= 0.99, cos
