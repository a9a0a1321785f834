"""A chart split into bricks in the GPU's memory (gx_create_bricks, the
reference's Chart_SLD with nBricks: Chart_SLD__Form.f90:131-201, 324-554,
557-1007) steps bit for bit as the one-brick chart: the same time-step
candidates, conserved fields and shock flags after every step, the same
boundary fluence, and tallies equal to 1e-13 (their sums run in another
order).  Shocks cross the brick faces, so flags set next to a face between
bricks -- which only the chart's faces clear (FluidFeatures_P__Form.f90:
333-338) -- and the flag exchange matter; periodic faces split into bricks
read their ghosts from the far brick."""
import numpy as np
import pytest

from genasis_amd import _capi, problems
from genasis_amd.chart import Chart
from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC

from ._helpers import same_bits

pytestmark = pytest.mark.gpu

PER = (("PERIODIC", "PERIODIC"),) * 3
OUT = (("OUTFLOW", "OUTFLOW"),) * 3
MIXED = (("OUTFLOW", "REFLECTING"), ("PERIODIC", "PERIODIC"), ("REFLECTING", "OUTFLOW"))


def _shocky(chart, seed):
    """Random states with density and pressure jumps in every dimension."""
    rng = np.random.default_rng(seed)
    shape = chart.m[::-1]
    x, y, z = chart.padded_centers()
    N = np.exp(rng.uniform(np.log(0.3), np.log(3.0), shape))
    N *= 1.0 + 2.0 * (x > 0.45) + 1.0 * (y > 0.6) + 0.5 * (z < 0.35)
    P = np.exp(rng.uniform(np.log(0.3), np.log(3.0), shape)) * (1.0 + 4.0 * (x < 0.55))
    V = [rng.uniform(-0.5, 0.5, shape) for _ in range(3)]
    return N, V[0], V[1], V[2], P / 0.4


def _lockstep(chart, prim, n_bricks, params, n_steps):
    one = Step_RK2_C_ASC(chart, params)
    many = Step_RK2_C_ASC(chart, params, n_bricks=n_bricks)
    assert many.n_bricks == tuple(n_bricks)
    one.set_primitive(*prim)
    many.set_primitive(*prim)
    flagged = 0
    for s in range(n_steps):
        c1, cb = one.time_step_candidate(), many.time_step_candidate()
        assert c1 == cb, (n_bricks, s, c1, cb)
        ramp = min((s + 1) / 100.0, 1.0)
        dt = ramp * c1 if ramp < 1.0 else c1
        one.step(0.0, dt)
        many.step(0.0, dt)
        u1, ub = one.conserved(), many.conserved()
        assert same_bits(u1, ub), (n_bricks, s, np.argwhere(u1 != ub)[:3].tolist())
        f1, fb = one.flags(), many.flags()
        assert np.array_equal(f1, fb), (n_bricks, s, np.argwhere(f1 != fb)[:3].tolist())
        flagged += int((f1 != 0).sum())
        for b, (a1, ab) in enumerate(zip(one.boundary_fluence(), many.boundary_fluence())):
            assert same_bits(a1, ab), (n_bricks, s, "fluence", b)
    # the sums run in another order: equal to 1e-13 of the tallies' size (a
    # periodic chart's boundary tally cancels to round-off)
    t1, tb = one.compute_tally(), many.compute_tally()
    scale = np.abs(t1["interior"]).max()
    for key in ("interior", "boundary", "total"):
        assert np.allclose(t1[key], tb[key], rtol=1e-13, atol=1e-13 * scale), key
    assert same_bits(one.primitive(), many.primitive())
    T1, Tb = one.totals(), many.totals()
    assert np.allclose(T1, Tb, rtol=1e-13, atol=1e-13 * np.abs(T1).max()), (T1, Tb)
    one.close()
    many.close()
    return flagged


@pytest.mark.parametrize("n_bricks", [(2, 2, 1), (1, 2, 2), (2, 1, 3), (4, 2, 2)])
def test_bricks_kelvin_helmholtz_periodic(n_bricks):
    chart = Chart((64, 48, 24), bc=PER)
    prim = problems.kelvin_helmholtz(chart, seed=9)
    _lockstep(chart, prim, n_bricks, FluidParameters(CourantFactor=0.8), 8)


@pytest.mark.parametrize("riemann", ["HLLC", "HLL"])
@pytest.mark.parametrize("n_bricks,bc", [((2, 3, 2), MIXED), ((3, 2, 2), OUT),
                                         ((2, 2, 2), PER)])
def test_bricks_shocks_bitwise(n_bricks, bc, riemann):
    chart = Chart((36, 30, 24), bc=bc)
    flagged = _lockstep(chart, _shocky(chart, 4), n_bricks,
                        FluidParameters(CourantFactor=0.8, RiemannSolverType=riemann), 10)
    assert flagged > 0


def test_bricks_sod_across_bricks():
    chart = Chart((128, 16, 16), bc=OUT)
    _lockstep(chart, problems.sod(chart), (8, 2, 1), FluidParameters(CourantFactor=0.8), 40)


def test_bricks_tile_32x8():
    chart = Chart((64, 32, 24), bc=MIXED)
    one = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8))
    many = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8), n_bricks=(2, 2, 2))
    many.set_tile(32, 8)
    prim = _shocky(chart, 11)
    one.set_primitive(*prim)
    many.set_primitive(*prim)
    for s in range(6):
        c = one.time_step_candidate()
        assert c == many.time_step_candidate()
        one.step(0.0, c)
        many.step(0.0, c)
        assert same_bits(one.conserved(), many.conserved()), s
        assert np.array_equal(one.flags(), many.flags()), s


def test_bricks_fp32():
    chart = Chart((48, 32, 24), bc=MIXED)
    _lockstep(chart, _shocky(chart, 7), (2, 2, 2),
              FluidParameters(CourantFactor=0.8, Precision="FP32"), 6)


def test_bricks_reference_layout_and_refusals():
    chart = Chart((24, 20, 16), bc=MIXED)
    prim = _shocky(chart, 3)
    nv = chart.n_values
    value = np.zeros((23, nv))
    for k, f in enumerate((6, 9, 10, 11, 15)):
        value[f, :] = prim[k].ravel()
    one = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8))
    many = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8), n_bricks=(2, 2, 2))
    for g in (one, many):
        g.upload(value)
        for _ in range(3):
            g.step(0.0, g.time_step_candidate())
    o1, ob = np.full((23, nv), -1.0), np.full((23, nv), -1.0)
    one.download(o1)
    many.download(ob)
    assert same_bits(o1, ob)
    d1, db = np.full((3, nv), -1.0), np.full((3, nv), -1.0)
    one.download_features(d1)
    many.download_features(db)
    assert np.array_equal(d1, db)
    with pytest.raises(_capi.GxError, match="bricks"):
        many.write_checkpoint("/tmp/never_written.chk", 0.0, 0)
    with pytest.raises(_capi.GxError, match="bricks"):
        many.record_sources(True)
    one.close()
    many.close()
