"""Each pin of tests/pin_checks.py catches a mistake planted in the oracle.

The oracle source carries numbered plant switches (ORA_PLANT in oracle/nsw6_oracle.c);
this test builds the oracle once per plant with that switch (oracle/Makefile target
`planted`, gcc, CPU only), loads the build, and requires the pin that covers the planted
function to fail on it, while the same pin passes on the oracle
(tests/test_oracle_pins_step.py).  The plants:

  1  negative_filter does not add the removed hydrometeor mass to QV
  2  negative_filter clips QC too
  3  negative_filter changes rhog with the wrong sign of the vapour deficit
  4  THRMDYN_qd subtracts QV
  5  THRMDYN_tempre swaps Rdry and Rvap
  6  cnvvar_rhogkin_in swaps the two cell-to-wall weights
  7  cnvvar_rhogkin_in swaps the two wall-to-cell weights
  8  precip_transport_new drops the potential-energy term of the falling mass
  9  KIN2EIN with the wrong sign
 10  the sedimented horizontal momentum with the unit vectors i and j swapped
 11  the mass flux divergence over the half-level thickness
 12  SATURATION_adjustment's moist energy without the vapour's latent heat
 13  SATURATION_adjustment leaves the halo levels' condensate
 14  Praci with lambda^-3 instead of lambda^-(3+Dr)
 15  Psacw with the snow moment 2 instead of 2 + Ds
 16  Pgacr without the factor 2 of the cross term
 17  Prevp's ventilation with lambda^3 instead of lambda^2
 18  Psmlt's conduction term over LHV0 instead of LHF0
 19  Pgfrz with lambda^6 instead of lambda^7
 20  Pgacw with Gamma(3) instead of Gamma(3 + Dg)
 21  Pgaci with the cloud-water efficiency Egw instead of Egi
 22  Psaci without the temperature factor of its efficiency
 23  Piacr with Gamma(3 + Dr) instead of Gamma(6 + Dr)
 24  Psacr without the factor 2 of the cross term
 25  Pracs with the snow moment 3 instead of 4
 26  Pgacs with the rain speed instead of the snow speed
 27  the ice diffusion factor with LHV0 instead of LHS0
 28  the graupel ventilation with lambda^(3+dg) instead of lambda^(5+dg)
 29  the cloud-ice number with (rho qi)^0.7 instead of ^0.75
 30  the nucleated ice mass with Ni0^1.3 instead of ^1.33
 31  Pihtr linear in qc
 32  Pihom from -38 C instead of -40 C
 33  Pgmlt without the heat of the collected cloud water
 34  Pgaut with the ice threshold instead of the snow threshold
 35  Psfw's riming with Ri50 instead of Ri50^2
 36  the rain partition at qr = 1e-5 instead of 1e-4
 37  the graupel slope with the rain intercept N0r
 38  the original Pracw with Gamma(3) instead of Gamma(3 + Dr)
 39  Pimlt melting the snow mass instead of the ice mass
 40  KK2000's accretion with exponent 1.5 instead of 1.15
 41  the graupel fall speed with lambda instead of lambda^(dg)
 42  satadjust_liq's slope without the heat capacity
 43  satadjust_all's slope without the fusion term
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import pin_checks as pc
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")
PLANTS = sorted(n for _, (_, ns) in pc.PINS.items() for n in ns)


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    d = tmp_path_factory.mktemp("planted")

    def build(n):
        out = str(d / f"liboracle_plant{n}.so")
        r = subprocess.run(["make", "-s", "-C", ORACLE, "planted", f"PLANT={n}", f"OUT={out}"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert r.returncode == 0, r.stdout
        return n, out
    with ThreadPoolExecutor(max_workers=8) as pool:
        return dict(pool.map(build, PLANTS))


@pytest.mark.parametrize("plant", PLANTS)
def test_pin_catches_planted_mistake(planted, plant):
    name = next(k for k, (_, ns) in pc.PINS.items() if plant in ns)
    check, _ = pc.PINS[name]
    L = O.load(planted[plant])
    with pytest.raises(AssertionError):
        check(L)
