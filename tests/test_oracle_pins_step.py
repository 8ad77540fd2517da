"""The step-phase pins of tests/pin_checks.py on the oracle itself: THRMDYN,
negative_filter, cnvvar_rhogkin_in, the transport's column budgets (mass, momentum,
total energy under KIN2EIN), the saturation-adjustment wrapper, and the process rates
Praci, Psacw, Pgacr, Prevp, Psmlt and Pgfrz against quadratures of their size
integrals and the published snow-moment relation.  tests/test_oracle_plants.py shows
that each pin catches a planted mistake."""
import pytest

import pin_checks as pc
from oracle import oracle as O


@pytest.mark.parametrize("name", sorted(pc.PINS))
def test_pin_holds_on_the_oracle(name):
    check, _ = pc.PINS[name]
    check(O.lib())
