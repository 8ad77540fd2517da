"""The step-phase pins of tests/pin_checks.py on the oracle itself: THRMDYN,
negative_filter, cnvvar_rhogkin_in, the transport's column budgets (mass, momentum,
total energy under KIN2EIN), the saturation-adjustment wrapper and its Newton slopes,
and every process rate of TagBigLoop (collection, vapour growth, nucleation, freezing,
melting, conversions, the partitions between snow and graupel, the slope parameters
and fall speeds) against quadratures of their size integrals, the published snow-moment
relation and the published parameterisations they follow.  tests/test_oracle_plants.py
shows that each pin catches a planted mistake."""
import pytest

import pin_checks as pc
from oracle import oracle as O


@pytest.mark.parametrize("name", sorted(pc.PINS))
def test_pin_holds_on_the_oracle(name):
    check, _ = pc.PINS[name]
    check(O.lib())
