"""Pins the oracle's shock detector (FluidFeatures_P % Detect) against flag
sets written out by hand from the reference rule (tests/shock_cases.py): the
0.1 threshold on both sides of it, the sqrt(tiny) floor of P_Min, the sign
test on dV of the swept velocity component, the cyclic transverse spread and
the boundary clearing."""
import pytest

import oracle

from . import shock_cases
from ._helpers import oracle_flag_bits

CASES = shock_cases.cases()


@pytest.mark.parametrize("name,chart,prim,expected", CASES, ids=[c[0] for c in CASES])
def test_flag_set(name, chart, prim, expected):
    o = oracle.OracleFluid(chart)
    o.set_primitive(*prim)
    o.step(0.0)
    got = oracle_flag_bits(o, chart)
    assert (got == expected).all(), (name, int((got != expected).sum()))


def test_cases_are_not_trivial():
    fired = [c for c in CASES if c[3].any()]
    assert 10 < len(fired) < len(CASES)
