"""The FP64 stage kernels' power-of-two geometry form: where every centre
spacing, slope denominator and cell width of the chart is a power of two,
the divisions by the geometry are multiplications by the exact reciprocal
(gdiv).  They must give the oracle's bits like the general form, for both
tile shapes, on a chart off the unit cube with every boundary condition."""
import numpy as np
import pytest

from genasis_amd.chart import Chart
from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC

from .test_gpu_parity import _pair, _random_state, _run_lockstep

pytestmark = pytest.mark.gpu

MIXED = (("REFLECTING", "OUTFLOW"), ("PERIODIC", "PERIODIC"), ("OUTFLOW", "REFLECTING"))


@pytest.mark.parametrize("n,lo,hi,expect", [
    ((32, 16, 8), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), True),
    ((32, 16, 8), (-1.0, 0.0, 0.5), (1.0, 0.5, 0.75), True),   # widths 1/16, 1/32, 1/32
    ((24, 16, 8), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), False),    # 1/24
    ((32, 16, 8), (0.0, 0.0, 0.0), (1.0, 1.0, 0.3), False),    # 0.3/8
])
def test_geometry_class(n, lo, hi, expect):
    chart = Chart(n, lo=lo, hi=hi, bc=MIXED)
    g = Step_RK2_C_ASC(chart, FluidParameters())
    assert g.geometry_pow2() == expect


@pytest.mark.parametrize("tile", [(16, 16), (32, 8)])
def test_power_of_two_geometry_bitwise(tile):
    chart = Chart((32, 16, 8), lo=(-1.0, 0.0, 0.5), hi=(1.0, 0.5, 0.75), bc=MIXED)
    prim = _random_state(chart, 11)
    o, g = _pair(chart, prim)
    assert g.geometry_pow2()
    g.set_tile(*tile)
    _run_lockstep(o, g, chart, 6, f"pow2 {tile}")


def test_general_geometry_bitwise():
    """The same chart with one non-power-of-two extent takes the corrected
    quotient and equals the oracle too."""
    chart = Chart((32, 16, 8), lo=(-1.0, 0.0, 0.5), hi=(1.0, 0.5, 0.8), bc=MIXED)
    prim = _random_state(chart, 11)
    o, g = _pair(chart, prim)
    assert not g.geometry_pow2()
    _run_lockstep(o, g, chart, 6, "general")


@pytest.mark.parametrize("riemann,use_limiter", [("HLL", True), ("HLLC", False)])
def test_power_of_two_geometry_hll_and_unlimited_bitwise(riemann, use_limiter):
    """The centred slope without the limiter, and HLL everywhere, in the
    power-of-two form."""
    chart = Chart((16, 32, 8), lo=(0.0, -0.25, 0.0), hi=(0.5, 0.25, 0.25), bc=MIXED)
    prim = _random_state(chart, 12)
    o, g = _pair(chart, prim, riemann=riemann, use_limiter=use_limiter)
    assert g.geometry_pow2()
    _run_lockstep(o, g, chart, 5, f"pow2 {riemann} limiter={use_limiter}")
