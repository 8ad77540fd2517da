"""Parity of the HIP step with the CPU oracle at the production region width.

One `k_extremes` block reduces 4,096 columns of one level, so a region of more than
4,096 columns splits the per-level departure-distance extremes of vadv1d_prep
(src/mod_vadv1d.cpp:317-327) over several blocks, which `k_bounds` merges before its
kcell search (:329-352).  These tests run that multi-block path:

* the device's kcell_min/kcell_max (nsw6_kcell_bounds) equal the oracle's
  (ora_record_kcell_bounds) at ijdim = 4,097, 8,193 and 16,900, with the extreme
  column planted in the LAST block of one region (a heavy-rain column at ij =
  ijdim - 1 that moves kcell_max at the lower levels), and the step agrees with the
  oracle within 1e-10 of each field's max-abs;
* the 10 regions x 16,900 columns x 96 levels workload (dt = 15 s): from each of 10
  consecutive states of the oracle's trajectory the GPU step agrees with the
  oracle's within 1e-10 of each field's max-abs, except at the points where the
  oracle's own Newton stopping test lies within 1e-4 of its criterion (a handful per
  step), where q and rhogq of QV and QC may differ by the reference's own one-
  iteration jump; a free-running GPU trajectory is reported beside it with the
  reference's check table of (GPU - oracle);
* one step at the reference's own shape, one region of ijdim = 16,641
  (include/problemsize.h:63);
* a ragged 2 x 9,001 state in which one region takes the general (multi-cell)
  sedimentation and the other the single-cell path, over 3 steps;
* the reverse mode at 2 x 9,001 against the host run of the same reverse passes;
* 2 steps of a 2 x 9,001 state with deep convection on the reference's own vgrid
  (72 m bottom layers) at dt = 60 s, where rain falls through up to 7 layers and
  some departures point upward: both regions take the general wall fluxes (from the
  third step on the reference's own saturation adjustment leaves points
  unconverged in that state).

Tolerance: 1e-10 of each field's max-abs (FP64 form); kcell bounds bit-exact.
"""
import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import fields as F
from nicam_kokkos_microphysics_amd.check import check_table, field_errors

pytestmark = pytest.mark.gpu

TOL = 1e-10
DT = 15.0
COMPARE = [k for k in F.OUTPUTS if k not in ("rhogq", "q")] + \
          [f"rhogq:{t}" for t in F.TRACERS] + [f"q:{t}" for t in F.TRACERS]


def _plant_heavy_rain(st, l, qr=5e-3, ktop=40):
    """Heavy rain (q(QR) = qr) over levels 1..ktop-1 of the last column of region l."""
    j = st["rhog"].shape[-1] - 1
    st["rhogq"][l, 2, 1:ktop, j] = st["rhog"][l, 1:ktop, j] * qr
    st["q"][l, 2, 1:ktop, j] = qr


def _general_regions(b):
    """Regions whose kcell bounds leave the single-cell range [k, k+1] somewhere."""
    k = np.arange(b.shape[-2])
    g = (b[..., 0] != k) | (b[..., 1] > k + 1)
    return g.any(axis=(1, 2, 3))


@pytest.mark.parametrize("ijdim", [4097, 8193, 16900])
def test_kcell_bounds_multiblock_last_block_extreme(ijdim):
    gz, gzh, st = F.synth_state(2, ijdim, seed=31)
    _plant_heavy_rain(st, 1)
    reps, bounds = _run(st, gz, gzh, 1)
    # the planted column sits in the last reduction block and moves the region's
    # bounds: region 1 differs from region 0 in kcell_max at the lower levels
    assert (bounds[0][1] != bounds[0][0]).any()
    assert reps[-1].n_general_regions == 0


# Points where the oracle's saturation-adjustment Newton stopped within this relative
# distance of its criterion dtemp_criteria = 1e-8 K (over its last two steps).  Its
# residual Emoist_new - Emoist is a difference of two ~2.7e5 J/kg values, so a
# last-bit change anywhere upstream moves the final |dtemp| by ~1e-6 relative; at
# such a point one more or one fewer iteration runs, and because q(QV) = qsat of the
# previous iterate (src/mod_satadjust.cpp:838-878) the reference's own result jumps
# by dqsat/dT * |dtemp| (about 1e-11 in q).  The threshold is 100x that noise.
NEWTON_MARGIN = 1e-4
NEWTON_FIELDS = ("q:QV", "q:QC", "rhogq:QV", "rhogq:QC")
LOVR_ICE = 2834000.0 / 461.46   # the steeper of the two Clausius-Clapeyron slopes


def _newton_jump_bound(ref, key):
    """|jump| of `key` from one Newton iteration more or fewer: dqsat/dT * 1e-8 K
    (x 2 margin), times rhog for rhogq."""
    qv, tem = ref["q"][:, 0], ref["tem"]
    b = 2.0 * qv * LOVR_ICE / tem ** 2 * 1e-8
    return b * ref["rhog"] if key.startswith("rhogq") else b


def _assert_step_parity(out, ref, margin):
    """Every field within TOL of its max-abs at every point, except that at the
    oracle's Newton-borderline points (margin < NEWTON_MARGIN) q and rhogq of QV and
    QC may differ by up to the reference's own iteration jump.  Returns the number of
    borderline points."""
    border = margin < NEWTON_MARGIN
    errs = field_errors(out, ref, COMPARE)
    bad = {}
    for key in COMPARE:
        err, scale, ratio = errs[key]
        if ratio <= TOL:
            continue
        if key not in NEWTON_FIELDS:
            bad[key] = (err, scale, ratio)
            continue
        d = np.abs(F.field_view(out, key) - F.field_view(ref, key))
        off = d > TOL * scale
        if (off & ~border).any():
            i = tuple(np.argwhere(off & ~border)[0])
            bad[key] = ("outside the borderline points", i, float(d[i]), scale)
        elif (d[off] > TOL * scale + _newton_jump_bound(ref, key)[off]).any():
            bad[key] = ("beyond the Newton jump bound", err, scale)
    assert not bad, "fields beyond tolerance: " + repr(bad)
    return int(border.sum())


def _run(st, gz, gzh, steps, P_edit=None, dt=DT):
    """`steps` consecutive steps along the oracle's trajectory: at each the GPU steps
    from the oracle's state and must agree with the oracle's step
    (_assert_step_parity), kcell bounds bit-exact.  Returns the GPU reports and the
    oracle's kcell bounds of each step."""
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6, default_params

    lall, ijdim = st["rhog"].shape[0], st["rhog"].shape[-1]
    P_ora, P_gpu = O.params(), default_params()
    for k, v in (P_edit or {}).items():
        setattr(P_ora, k, v)
        setattr(P_gpu, k, v)
    ref = F.copy_state(st)
    drv = NSW6(gz, gzh, params=P_gpu, max_columns=lall * ijdim)
    reps, bounds = [], []
    try:
        for s in range(steps):
            print(f"step {s + 1}/{steps}: lall={lall} ijdim={ijdim}", flush=True)
            one = F.to_device(ref)
            drv.step(one, dt)
            torch.cuda.synchronize()
            reps.append(drv.report())
            b_gpu = drv.kcell_bounds(lall)
            r, b_ora, margin = O.mp_driver_traced(ref, gz, gzh, dt, P=P_ora)
            assert r.status == 0
            out = F.to_host(one)
            del one
            mism = np.argwhere(b_gpu != b_ora)
            assert mism.size == 0, f"step {s + 1}: kcell bounds differ at [l, t, type, k, min/max]: " \
                                   f"{mism[:10].tolist()}"
            n = _assert_step_parity(out, ref, margin)
            print(f"   Newton-borderline points: {n}", flush=True)
            bounds.append(b_ora)
    finally:
        drv.close()
    return reps, bounds


def test_ten_steps_10x16900():
    """The FP64 throughput workload, 10 regions x 16,900 columns x 96 levels, along 10
    consecutive steps of the oracle's trajectory: at each step the GPU takes the
    oracle's state and its step must agree with the oracle's (_assert_step_parity);
    the kcell bounds bit-exact.  A free-running GPU trajectory beside it is reported
    with the check table of (GPU - oracle) after the 10 steps."""
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6

    w = F.WORKLOADS["nsw6_169000x96_fp64"]
    gz, gzh, st = F.workload_state(w)
    lall, ijdim = w.lall, w.ijdim
    ref = F.copy_state(st)
    free = F.to_device(st)
    drv = NSW6(gz, gzh, max_columns=lall * ijdim)
    n_border = []
    try:
        for s in range(w.steps):
            print(f"step {s + 1}/{w.steps}: lall={lall} ijdim={ijdim}", flush=True)
            one = F.to_device(ref)
            drv.step(one, DT)
            torch.cuda.synchronize()
            b_gpu = drv.kcell_bounds(lall)
            r, b_ora, margin = O.mp_driver_traced(ref, gz, gzh, DT)
            assert r.status == 0
            out = F.to_host(one)
            del one
            assert np.array_equal(b_gpu, b_ora), f"kcell bounds differ at step {s + 1}"
            n_border.append(_assert_step_parity(out, ref, margin))
            drv.step(free, DT)
        torch.cuda.synchronize()
        out_free = F.to_host(free)
    finally:
        drv.close()
    print("Newton-borderline points per step:", n_border)
    print("free-running, after 10 steps (GPU - oracle):")
    print("\n".join(check_table(out_free, ref)))
    errs = field_errors(out_free, ref, COMPARE)
    worst = max(errs.items(), key=lambda kv: kv[1][2])
    print(f"free-running worst field {worst[0]}: {worst[1][2]:.3e} of max-abs")


def test_reference_shape_16641():
    """One region of ADM_gall_in = 16,641 columns (include/problemsize.h:63)."""
    gz, gzh, st = F.synth_state(1, 16641, seed=16641)
    _run(st, gz, gzh, 1)


def test_ragged_general_and_single_cell_regions():
    """2 x 9,001 columns (a ragged last block), rain 1.5x faster and a heavy-rain
    column in region 1: region 1 departs beyond one cell (general sweep), region 0
    stays on the single-cell path; 3 steps."""
    gz, gzh, st = F.synth_state(2, 9001, seed=41, frac_towers=0.1)
    _plant_heavy_rain(st, 1, qr=1e-2)
    reps, bounds = _run(st, gz, gzh, 3, P_edit={"Cr": 130.0 * 1.5})
    assert reps[0].n_general_regions == 1
    assert list(_general_regions(bounds[0])) == [False, True]


def test_reverse_mode_multiblock_against_host_passes():
    """The reverse mode's forward part (the same k_extremes/k_bounds) and passes at
    2 x 9,001 with the last-block plant, against the host run of the same passes:
    every input cotangent within 1e-9 of its field's max-abs."""
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6

    from adjoint_helpers import random_cotangents

    gz, gzh, st = F.synth_state(2, 9001, seed=43)
    _plant_heavy_rain(st, 1)
    cot, cs = random_cotangents(st, gz, gzh, 3)
    g_host, gc_host, status = O.adjoint_host(st, gz, gzh, DT, cot, cs)
    assert status == 0
    drv = NSW6(gz, gzh, max_columns=2 * 9001)
    try:
        dev = F.to_device(st)
        dcot = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in cot.items()}
        dcs = torch.from_numpy(np.ascontiguousarray(cs)).cuda()
        print("adjoint lall=2 ijdim=9001", flush=True)
        grad, gc = drv.adjoint(dev, dcot, DT, cot_sflux=dcs)
        torch.cuda.synchronize()
        rep = drv.report()
    finally:
        drv.close()
    assert rep.status == 0 and rep.n_general_regions == 0
    for k in F.ADJ_INPUTS:
        g = grad[k].cpu().numpy()
        scale = np.abs(g_host[k]).max()
        err = np.abs(g - g_host[k]).max()
        assert err <= 1e-9 * scale + 1e-300, (k, err, scale)
    np.testing.assert_allclose(gc.cpu().numpy(), gc_host, rtol=1e-9, atol=0)


def test_general_walls_reference_grid_dt60():
    """Deep convection (10 % towers, 1 % rain shafts) on the reference vgrid at
    dt = 60 s: kcell ranges of several cells and upward departures, so every region
    takes k_flux_general; 2 steps along the oracle's trajectory, kcell bounds
    bit-exact."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    gz = np.fromfile(os.path.join(here, "GRD_gz.dat"), dtype="<f8")
    gzh = np.fromfile(os.path.join(here, "GRD_gzh.dat"), dtype="<f8")
    gz, gzh, st = F.synth_state(2, 9001, seed=41, gz=gz, gzh=gzh, frac_towers=0.1, frac_rainshafts=0.01)
    reps, bounds = _run(st, gz, gzh, 2, dt=60.0)
    k = np.arange(96)
    for r, b in zip(reps, bounds):
        assert r.n_general_regions == 2
        assert _general_regions(b).all()
        assert (b[:, :, 0, 1:95, 1] - k[1:95]).max() >= 3      # departures 3+ cells up
        assert (b[:, :, 0, 1:95, 0] - k[1:95]).min() <= -1     # and upward ones
