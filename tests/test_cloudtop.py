"""Cloud-top diagnostics rctop, tctop, rwtop (src/mod_mp_nsw6.cpp:710-726, 1566-1592).

* Default (the reference as written): the switch [qc+qr > q_min]*zerosw is 0 wherever
  a particle moment exists, and the EPS term of rf_qc alone keeps r2 above r2_min, so
  every level writes UNDEF and the three rows are UNDEF everywhere.
* cloudtop_intended = 1: the highest level with qc+qr > q_min gives rctop = rceff
  (the step's re_liquid there) and tctop = T (the input temperature there); UNDEF
  when no level qualifies; rwtop = rctop where tctop > TEM00, else UNDEF.  The
  expectation is formed here in numpy from the inputs and outputs (the rate
  kernel's q of QC/QR is the input q: the negative filter only rescales them by
  rhog, far from the threshold in these states), so it pins the oracle's serial
  update and the GPU's per-column maximum independently.
"""
import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import fields as F

UNDEF = -9.9999e30
TEM00 = 273.15
DT = 15.0


def _state(lall=2, ijdim=96, seed=3):
    gz, gzh, st = F.synth_state(lall, ijdim, seed=seed, frac_supersat=0.05, frac_towers=0.3)
    # column classes: no cloud at all, cloud only below level 20, supercooled cloud
    # water added at level 60 (a cold top), untouched
    for l in range(lall):
        for ij in range(ijdim):
            if ij % 3 == 0:
                ks = slice(0, 96) if ij % 2 == 0 else slice(20, 96)
                for t in (1, 2):   # QC, QR
                    st["rhogq"][l, t, ks, ij] = 0.0
                    st["q"][l, t, ks, ij] = 0.0
            elif ij % 3 == 1 and ij % 2 == 0:
                st["q"][l, 1, 60, ij] = 3.0e-5
                st["rhogq"][l, 1, 60, ij] = 3.0e-5 * st["rhog"][l, 60, ij]
    return gz, gzh, st


def _expected(q_in, tem_in, re_liquid):
    lall, _, ijdim = tem_in.shape
    rct = np.full((lall, 1, ijdim), UNDEF)
    tct = np.full((lall, 1, ijdim), UNDEF)
    cloudy = (q_in[:, 1] + q_in[:, 2]) > 1.0e-5
    for l in range(lall):
        for ij in range(ijdim):
            ks = np.nonzero(cloudy[l, 1:95, ij])[0]
            if ks.size:
                k = 1 + ks[-1]
                rct[l, 0, ij] = re_liquid[l, k, ij]
                tct[l, 0, ij] = tem_in[l, k, ij]
    rwt = np.where(tct > TEM00, rct, UNDEF)
    return rct, tct, rwt


def _margin_ok(q_in):
    s = q_in[:, 1, 1:95] + q_in[:, 2, 1:95]
    return np.abs(s - 1.0e-5).min() > 1.0e-8 * 1.0e-5


def test_oracle_default_mode_is_undef_everywhere():
    from oracle import oracle as O
    gz, gzh, st = _state()
    O.mp_driver(st, gz, gzh, DT)
    for k in ("rctop", "tctop", "rwtop"):
        assert np.all(st[k] == UNDEF), k


def test_oracle_intended_mode_is_the_highest_cloudy_level():
    from oracle import oracle as O
    gz, gzh, st = _state()
    assert _margin_ok(st["q"])
    q_in, tem_in = st["q"].copy(), st["tem"].copy()
    P = O.params()
    P.cloudtop_intended = 1.0
    assert O.mp_driver(st, gz, gzh, DT, P=P).status == 0
    rct, tct, rwt = _expected(q_in, tem_in, st["re_liquid"])
    np.testing.assert_array_equal(st["rctop"], rct)
    np.testing.assert_array_equal(st["tctop"], tct)
    np.testing.assert_array_equal(st["rwtop"], rwt)
    # every class is present: no cloud, a warm low top, a cold high top
    assert (rct == UNDEF).any() and (rwt != UNDEF).any() and ((tct != UNDEF) & (tct < TEM00)).any()


@pytest.mark.gpu
@pytest.mark.parametrize("intended", [0.0, 1.0])
def test_gpu_cloud_top_matches_oracle(intended):
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6, default_params
    gz, gzh, st = _state(lall=3, ijdim=900)
    assert _margin_ok(st["q"])
    q_in, tem_in = st["q"].copy(), st["tem"].copy()
    P_ora, P_gpu = O.params(), default_params()
    P_ora.cloudtop_intended = intended
    P_gpu.cloudtop_intended = intended
    ref = F.copy_state(st)
    assert O.mp_driver(ref, gz, gzh, DT, P=P_ora).status == 0
    dev = F.to_device(st)
    drv = NSW6(gz, gzh, params=P_gpu, max_columns=3 * 900)
    outs = []
    for _ in range(2):                     # the second call starts from the cleared levels
        d = F.to_device(st)
        drv.step(d, DT)
        torch.cuda.synchronize()
        outs.append(F.to_host(d))
    drv.close()
    for k in ("rctop", "tctop", "rwtop", "re_liquid"):
        # the same cells are UNDEF; values within the FP64 parity bar (re_liquid goes
        # through exp/log, whose last bits differ between the host and device libraries)
        np.testing.assert_array_equal(outs[0][k] == UNDEF, ref[k] == UNDEF, err_msg=k)
        a, b = outs[0][k][ref[k] != UNDEF], ref[k][ref[k] != UNDEF]
        if b.size:
            assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max(), k
        np.testing.assert_array_equal(outs[1][k], outs[0][k], err_msg=k)
    np.testing.assert_array_equal(outs[0]["tctop"], ref["tctop"])
    if intended:
        rct, tct, rwt = _expected(q_in, tem_in, outs[0]["re_liquid"])
        np.testing.assert_array_equal(outs[0]["rctop"], rct)
        np.testing.assert_array_equal(outs[0]["tctop"], tct)
        np.testing.assert_array_equal(outs[0]["rwtop"], rwt)
    del dev
