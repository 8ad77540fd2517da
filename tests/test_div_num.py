"""MP_DIV_NUM sub-cycling of mp_driver (src/mod_mp_driver.cpp:420-612): n NSW6 steps
of dt_mp = dt * (1/n); precip, precip_rhoe/lh_heat/rhophi/rhokin, GPREC, GDCLW,
GDCFRC and the ISO precipitations are the sub-step sums times 1/n.

* CPU: the oracle's sub-cycled call equals n single oracle steps of dt_mp with the
  sums formed here, bit for bit (a wrong dt_mp, a missing zeroing of an output
  between sub-steps or a wrong fraction breaks it).
* GPU: the HIP call equals the oracle's within 1e-10 of each field's max-abs, and
  equals n single HIP steps with the sums formed in torch, bit for bit.
"""
import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from nicam_kokkos_microphysics_amd import fields as F
from nicam_kokkos_microphysics_amd.check import compare_states

DT = 15.0
SUMMED = ("precip", "ISO1_precip", "ISO2_precip", "precip_rhoe", "precip_lh_heat", "precip_rhophi",
          "precip_rhokin", "GDCLW", "GDCFRC", "GPREC")
COMPARE = [k for k in F.OUTPUTS if k not in ("rhogq", "q")] + \
          [f"rhogq:{t}" for t in F.TRACERS] + [f"q:{t}" for t in F.TRACERS]


def _state(lall=2, ijdim=48, seed=11):
    """Near-freezing surface (275 K): both rain and solid precipitation reach it
    (precip row 1 accumulates across sub-steps, src/mod_precip_transport.cpp:933)."""
    return F.synth_state(lall, ijdim, seed=seed, frac_supersat=0.05, frac_towers=0.25, tsfc=275.0)


def _composed(step, st, n, zeros):
    """n single steps of dt * (1/n) via `step(state, dt)`, then the outputs replaced by
    their sums (0 + x at the first sub-step) times 1/n."""
    frac = 1.0 / float(n)
    sums = {}
    for m in range(n):
        step(st, DT * frac)
        for k in SUMMED:
            sums[k] = (zeros(st[k]) + st[k]) if m == 0 else sums[k] + st[k]
    for k in SUMMED:
        st[k][...] = sums[k] * frac


@pytest.mark.parametrize("n", [2, 3])
def test_oracle_subcycling_is_the_composition(n):
    from oracle import oracle as O
    gz, gzh, st = _state()
    a, b = F.copy_state(st), F.copy_state(st)
    assert O.mp_driver(a, gz, gzh, DT, div_num=n).status == 0
    _composed(lambda s, dt: O.mp_driver(s, gz, gzh, dt), b, n, np.zeros_like)
    for k in F.OUTPUTS:
        assert np.array_equal(a[k], b[k]), k
    # both precipitation rows are exercised, and sub-cycling is not a no-op
    assert np.count_nonzero(a["precip"][:, 0]) and np.count_nonzero(a["precip"][:, 1])
    c = F.copy_state(st)
    assert O.mp_driver(c, gz, gzh, DT).status == 0
    assert not np.array_equal(a["precip"], c["precip"])


def test_oracle_div_one_is_the_plain_step():
    from oracle import oracle as O
    gz, gzh, st = _state(lall=1, ijdim=32)
    a, b = F.copy_state(st), F.copy_state(st)
    O.mp_driver(a, gz, gzh, DT)
    assert O.mp_driver(b, gz, gzh, DT, div_num=1).status == 0
    for k in F.OUTPUTS:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("n", [2, 3])
def test_gpu_subcycling_matches_oracle(n):
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = _state(lall=2, ijdim=700)
    ref = F.copy_state(st)
    dev = F.to_device(st)
    drv = NSW6(gz, gzh, max_columns=2 * 700, div_num=n)
    for _ in range(2):
        assert O.mp_driver(ref, gz, gzh, DT, div_num=n).status == 0
        drv.step(dev, DT)
    bad = compare_states(F.to_host(dev), ref, COMPARE, 1e-10)
    drv.close()
    assert not bad, bad


@pytest.mark.gpu
def test_gpu_subcycling_is_the_composition():
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = _state(lall=3, ijdim=1000)
    a, b = F.to_device(st), F.to_device(st)
    sub = NSW6(gz, gzh, max_columns=3000, div_num=3)
    one = NSW6(gz, gzh, max_columns=3000)
    sub.step(a, DT)
    _composed(lambda s, dt: one.step(s, dt), b, 3, torch.zeros_like)
    torch.cuda.synchronize()
    for k in F.OUTPUTS:
        assert torch.equal(a[k], b[k]), k
    # a later call with div_num = 1 on the same context is the plain step again
    sub.set_div_num(1)
    c, d = F.to_device(st), F.to_device(st)
    sub.step(c, DT)
    one.step(d, DT)
    for k in F.OUTPUTS:
        assert torch.equal(c[k], d[k]), k
    sub.close()
    one.close()


@pytest.mark.gpu
def test_gpu_subcycling_limits():
    from nicam_kokkos_microphysics_amd.driver import NSW6, NSW6Error
    gz, gzh, st = _state(lall=1, ijdim=64)
    drv = NSW6(gz, gzh, max_columns=64)
    for bad in (0, -1, N.NSW6_MAX_DIV_NUM + 1):
        with pytest.raises(NSW6Error):
            drv.set_div_num(bad)
    drv.set_div_num(N.NSW6_MAX_DIV_NUM)
    drv.set_div_num(2)
    drv.close()
