"""The reverse mode at the sizes of the adjoint and rollout workloads.

Adjoint workload (one region of 16,900 columns x 96 levels, one FP64 step; loss
L = sum precip^2 + sum (rhoge_out - rhoge_in) / dt, the single-step adjoint check):

* along 3 random directions through the 8 tunable coefficients, the GPU gradient
  against central differences of the GPU step at h = 1e-6 relative: 1e-6 relative;
* along 3 random directions through every input field restricted to 64 columns
  scattered over the region, against central differences, for the directions whose
  difference quotient is the same at h and 2h (no switch crossed): 1e-6 relative,
  at least 2 such directions;
* along 3 random directions through every input field (all columns) and the
  coefficients, against the forward-mode dual-number derivative of the host step
  (oracle.dual_vjp, the same cotangents including the surface mass fluxes): 1e-6
  relative.

A difference stencil through every field of all 16,900 columns crosses some
discontinuous switch of the scheme at every h (threshold switches of the slope
parameters and rate partitions, the region-wide kcell bounds, the Newton's stopping
rule): the quotients then differ by 1e-2 .. 2e-1 between h = 1e-5 and 1e-7 while the
coefficient directions agree to ~1e-8, so the field check uses a column subset.

Rollout workload (40 regions x 16,900 columns = 676,000 x 96, 20 steps of 15 s,
checkpoints every 5 steps, 10 % convective towers):

* the FP64 rollout gradient (reverse mode over all 20 steps) against central
  differences of the FP64 rollout loss along directions through every input field of
  4 scattered columns whose quotient is the same at h and 2h to 1e-6: 1e-5 relative, at
  least 2 such directions of 6 (coefficient directions, which move all 676,000 columns,
  cross switches at every h: their quotients move by 0.5-12 % between h values);
* the mixed-precision rollout (FP32 rates forward, FP64 reverse along its trajectory)
  recomputes its trajectory exactly, its loss is within 1e-4 of the FP64 rollout's,
  every input field's gradient but rhogq's is within cosine 0.999 of the FP64
  rollout's and the coefficient gradients within 1e-2 relative; rhogq's gradient is
  dominated by points near q = 0 where the slope parameters have unbounded derivatives
  (its cosine is printed, 0.64 when measured).
"""
import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from nicam_kokkos_microphysics_amd import fields as F

from adjoint_helpers import coefs, directional, random_cotangents

pytestmark = pytest.mark.gpu
DT = 15.0


def _loss(drv, dev):
    import torch
    s = F.copy_state(dev)
    r0 = s["rhoge"].clone()
    drv.step(s, DT)
    torch.cuda.synchronize()
    p = s["precip"]
    return float((p * p).sum() + ((s["rhoge"] - r0) / DT).sum()), s


def _fd(drv, dev, d, dc, h):
    """Central difference of the loss along fields `d` (device tensors) and
    coefficient direction `dc`."""
    from nicam_kokkos_microphysics_amd.driver import default_params
    c0 = coefs()
    vals = []
    for sgn in (1.0, -1.0):
        P = default_params()
        for c, name in enumerate(N.TUNABLE_COEFFS):
            setattr(P, name, c0[c] + sgn * h * dc[c])
        drv.set_params(P)
        sp = dict(dev)
        for k, v in d.items():
            sp[k] = dev[k] + sgn * h * v
        vals.append(_loss(drv, sp)[0])
    drv.set_params(default_params())
    return (vals[0] - vals[1]) / (2 * h)


@pytest.fixture(scope="module")
def adjoint_case():
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6
    w = F.WORKLOADS["nsw6_16900x96_adjoint"]
    gz, gzh, st = F.workload_state(w)
    drv = NSW6(gz, gzh, max_columns=w.ijdim)
    dev = F.to_device(st)
    print(f"adjoint workload: lall={w.lall} ijdim={w.ijdim}", flush=True)
    L0, out = _loss(drv, dev)
    cot = {"precip": 2.0 * out["precip"], "rhoge": torch.full_like(out["rhoge"], 1.0 / DT)}
    grad, gc = drv.adjoint(dev, cot, DT)
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in grad.items()}
    g["rhoge"] = g["rhoge"] - 1.0 / DT      # the explicit -rhoge_in/dt term of the loss
    yield gz, gzh, st, drv, dev, g, gc.cpu().numpy()
    drv.close()


def test_adjoint_16900_coefficient_directions_against_central_differences(adjoint_case):
    gz, gzh, st, drv, dev, g, gcoef = adjoint_case
    rng = np.random.default_rng(61)
    c0 = coefs()
    for i in range(3):
        dc = c0 * rng.uniform(-1, 1, 8)
        s = float(np.dot(gcoef, dc))
        fd = _fd(drv, dev, {}, dc, 1e-6)
        print(f"coefficient direction {i}: adjoint {s:.10e} fd {fd:.10e}", flush=True)
        assert abs(s - fd) <= 1e-6 * abs(fd), (i, s, fd)


def test_adjoint_16900_field_directions_on_a_column_subset(adjoint_case):
    import torch
    gz, gzh, st, drv, dev, g, gcoef = adjoint_case
    rng = np.random.default_rng(62)
    keys = ["rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge", "rhogq", "tem", "q", "vx", "vy", "vz"]
    checked = 0
    for i in range(3):
        cols = rng.choice(st["rhog"].shape[-1], 64, replace=False)
        mask = np.zeros(st["rhog"].shape[-1])
        mask[cols] = 1.0
        d = {k: st[k] * rng.uniform(-1, 1, st[k].shape) * mask for k in keys}
        dd = {k: torch.from_numpy(v).cuda() for k, v in d.items()}
        s, _ = directional(g, gcoef, d, np.zeros(8))
        f1, f2 = _fd(drv, dev, dd, np.zeros(8), 1e-6), _fd(drv, dev, dd, np.zeros(8), 2e-6)
        print(f"field direction {i}: adjoint {s:.10e} fd(h) {f1:.10e} fd(2h) {f2:.10e}", flush=True)
        if abs(f1 - f2) > 1e-5 * abs(f1):
            continue   # the stencil crosses a switch
        assert abs(s - f1) <= 1e-6 * abs(f1), (i, s, f1)
        checked += 1
    assert checked >= 2, checked


def test_adjoint_16900_against_dual_numbers(adjoint_case):
    """The same reverse mode on the general cotangents (every output, the surface mass
    fluxes) against the host dual-number derivative of the templated step."""
    import torch
    from oracle import oracle as O
    from adjoint_helpers import random_directions
    gz, gzh, st, drv, dev, _, _ = adjoint_case
    cot, cs = random_cotangents(st, gz, gzh, 63)
    dcot = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in cot.items()}
    grad, gc = drv.adjoint(dev, dcot, DT, cot_sflux=torch.from_numpy(np.ascontiguousarray(cs)).cuda())
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in grad.items()}
    gcoef = gc.cpu().numpy()
    dirs, dcoef = random_directions(st, 64, 3)
    _, dj = O.dual_vjp(st, gz, gzh, DT, cot, cs, dirs, dcoef)
    for i in range(3):
        s, _ = directional(g, gcoef, dirs[i], dcoef[i])
        print(f"dual direction {i}: adjoint {s:.12e} dual {dj[i]:.12e}", flush=True)
        assert abs(s - dj[i]) <= 1e-6 * abs(dj[i]), (i, s, dj[i])


# ----------------------------------------------------------------------------- rollout
def _rollout_state(w):
    """The rollout workload's state on the device, generated 10 regions at a time."""
    import torch
    parts = []
    gz = gzh = None
    for l0 in range(0, w.lall, 10):
        gz, gzh, part = F.workload_state(w, lall=min(10, w.lall - l0), l_offset=l0)
        parts.append(F.to_device(part))
        del part
    dev = {k: torch.cat([p[k] for p in parts], dim=0) for k in parts[0]}
    del parts
    return gz, gzh, dev


def _rollout_loss(fwd, dev, steps):
    import torch
    s = F.copy_state(dev)
    r0 = s["rhoge"].clone()
    L = 0.0
    for _ in range(steps):
        fwd.step(s, DT)
        L += float((s["precip"] ** 2).sum())
    L += float(((s["rhoge"] - r0) / (steps * DT)).sum())
    torch.cuda.synchronize()
    return L


def test_rollout_676000_gradients():
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6
    from nicam_kokkos_microphysics_amd.rollout import Rollout
    w = F.WORKLOADS["nsw6_676000x96_rollout"]
    steps, every = w.steps, 5
    gz, gzh, dev = _rollout_state(w)
    print(f"rollout workload: lall={w.lall} ijdim={w.ijdim} steps={steps} every={every}", flush=True)
    grads = {}
    for prec in (N.NSW6_FP64, N.NSW6_MIXED):
        ro = Rollout(gz, gzh, w.lall, w.ijdim, precision=prec, chunk_regions=10)
        s = F.copy_state(dev)
        r0 = s["rhoge"].clone()
        ckpt, precip = ro.forward(s, DT, steps, every)
        g, gc = ro.backward(s, ckpt, precip, DT, steps, every)   # checks every recomputed step
        torch.cuda.synchronize()
        grads[prec] = (g, gc.cpu().numpy(), ro.loss(r0, s, precip, DT))
        print(f"precision {prec}: loss {grads[prec][2]:.12e}", flush=True)
        ro.close()
        del s, ckpt, precip, r0
    g64, c64, L64 = grads[N.NSW6_FP64]
    gmx, cmx, Lmx = grads[N.NSW6_MIXED]
    assert abs(Lmx - L64) <= 1e-4 * abs(L64)
    # every input field's gradient along the mixed trajectory within cosine 0.999 of the
    # FP64 one, except rhogq: its norm is dominated by a few points where q ~ 0, at
    # which the slope parameters (q^(1/4)) and the ice fall speed ((rho q)^0.16) have
    # unbounded derivatives, so the two trajectories' values there differ
    for k in F.ADJ_INPUTS:
        a, b = g64[k], gmx[k]
        assert torch.isfinite(b).all(), k
        cos = float((a * b).sum() / (a.norm() * b.norm()))
        print(f"mixed vs FP64 rollout gradient, {k}: cosine {cos:.6f}", flush=True)
        if k != "rhogq":
            assert cos >= 0.999, (k, cos)
    print("coefficient gradients FP64", c64, "mixed", cmx, flush=True)
    np.testing.assert_allclose(cmx, c64, rtol=1e-2)
    del grads, gmx
    # the FP64 rollout gradient against central differences of the FP64 rollout loss along
    # directions through every input field of 4 columns scattered over the regions
    # (directions through the coefficients, or through all columns, cross switches of the
    # scheme somewhere in the 1.3e9 point-steps at every h)
    fwd = NSW6(gz, gzh, max_columns=w.lall * w.ijdim)
    gen = torch.Generator(device="cuda").manual_seed(66)
    rng = np.random.default_rng(66)
    checked = 0
    try:
        for i in range(6):
            cols = torch.zeros(w.lall, w.ijdim, dtype=torch.float64, device="cuda")
            pick = rng.choice(w.lall * w.ijdim, 4, replace=False)
            cols.view(-1)[torch.from_numpy(pick).cuda()] = 1.0
            d = {}
            for k in F.ADJ_INPUTS:
                m = cols.view(w.lall, *([1] * (dev[k].dim() - 2)), w.ijdim)
                u = torch.rand(dev[k].shape, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
                d[k] = dev[k] * u * m

            def fd(h):
                vals = []
                for sgn in (1.0, -1.0):
                    sp = dict(dev)
                    for k, v in d.items():
                        sp[k] = dev[k] + sgn * h * v
                    vals.append(_rollout_loss(fwd, sp, steps))
                return (vals[0] - vals[1]) / (2 * h)
            f1, f2 = fd(1e-6), fd(2e-6)
            s = sum(float((g64[k] * d[k]).sum()) for k in F.ADJ_INPUTS)
            print(f"rollout field direction {i}: adjoint {s:.10e} fd(h) {f1:.10e} fd(2h) {f2:.10e}", flush=True)
            if abs(f1 - f2) > 1e-6 * abs(f1):
                continue   # the stencil crosses a switch (or a kink) along the 20 steps
            assert abs(s - f1) <= 1e-5 * abs(f1), (i, s, f1)
            checked += 1
    finally:
        fwd.close()
    assert checked >= 2, checked
