"""Boundary fluence and conservation tallies on the GPU against the oracle:
the fluence of every boundary face is the oracle's bit for bit (the same face
fluxes, summed over the two stages in the same order); the tallies agree to
round-off (sums in a different order); the totals of baryon number, momentum
and fluid energy are conserved with the boundary counted, also at a size the
oracle is not run at; repeated runs give identical bits."""
import numpy as np
import pytest

from genasis_amd.chart import Chart
from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC
from genasis_amd import problems

from ._helpers import make_oracle, same_bits

pytestmark = pytest.mark.gpu

BN, M1, M2, M3, KE, AM1, AM2, AM3, GE, TE, FE, IE = range(12)
CONSERVED_TALLIES = [BN, M1, M2, M3, FE]
MIXED = (("OUTFLOW", "REFLECTING"), ("REFLECTING", "OUTFLOW"), ("OUTFLOW", "OUTFLOW"))
WALLS = (("REFLECTING", "REFLECTING"), ("PERIODIC", "PERIODIC"), ("OUTFLOW", "REFLECTING"))


def _state(chart, seed):
    rng = np.random.default_rng(seed)
    shape = chart.m[::-1]
    N = np.exp(rng.uniform(np.log(0.3), np.log(3.0), shape))
    V = [rng.uniform(-0.5, 0.5, shape) for _ in range(3)]
    E = np.exp(rng.uniform(np.log(0.3), np.log(3.0), shape)) / 0.4
    x, _, _ = chart.padded_centers()
    mid = 0.5 * (chart.lo[0] + chart.hi[0])
    N = N * (1 + 2 * (x > mid))
    E = E * (1 + 5 * (x > mid))  # a shock: flags switch HLLC to HLL
    return (N, V[0], V[1], V[2], E)


def _scale(tally):
    return np.maximum(np.abs(tally["interior"]), np.abs(tally["boundary"])).max()


def _assert_close(a, b, tag, rel=1e-13):
    for key in ("interior", "boundary", "total", "change"):
        s = max(np.abs(a[key]).max(), np.abs(b[key]).max(), 1e-300)
        err = np.abs(a[key] - b[key]).max()
        assert err <= rel * max(s, _scale(a)), (tag, key, err, a[key], b[key])


@pytest.mark.parametrize("bc,riemann", [(MIXED, "HLLC"), (WALLS, "HLL"),
                                        ((("PERIODIC", "PERIODIC"),) * 3, "HLLC")])
def test_fluence_and_tallies_match_the_oracle(bc, riemann):
    chart = Chart((19, 12, 10), lo=(-0.5, 0.25, 1.0), hi=(0.5, 1.0, 1.5), bc=bc)
    prim = _state(chart, 7)
    o = make_oracle(chart, prim, courant=0.8, riemann=riemann)
    g = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8, RiemannSolverType=riemann))
    g.set_primitive(*prim)
    # the uploaded state: the Current holds the uploaded primitives
    _assert_close(g.compute_tally(False), o.compute_tally(False), "upload")
    for cycle in range(3):
        for s in range(4):
            dt = o.time_step_candidate()[0]
            assert g.time_step_candidate() == dt
            o.step(dt)
            g.step(0.0, dt)
            faces = g.boundary_fluence()
            for b in range(6):
                ref = o.boundary_fluence(b)
                assert same_bits(faces[b], ref), (cycle, s, b, np.abs(faces[b] - ref).max())
        _assert_close(g.compute_tally(), o.compute_tally(), f"cycle {cycle}")


def test_uniform_flow_closed_form():
    """A uniform state keeps its exact flux on every face: the fluence of a
    step is dt times that flux, and the interior tallies are the closed forms."""
    lo, hi = (0.5, -1.0, 2.0), (1.5, 0.5, 2.75)
    chart = Chart((24, 16, 12), lo=lo, hi=hi, bc=(("OUTFLOW", "OUTFLOW"),) * 3)
    N0, v, E0, gamma = 1.3, np.array([0.2, -0.15, 0.1]), 2.1, 1.4
    shape = chart.n[::-1]
    g = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8))
    g.set_primitive(np.full(shape, N0), *[np.full(shape, vi) for vi in v], np.full(shape, E0))
    t0 = g.compute_tally(False)
    vol = np.prod(np.subtract(hi, lo))
    S, G, P = N0 * v, E0 + 0.5 * N0 * float(v @ v), (gamma - 1.0) * E0
    np.testing.assert_allclose(t0["interior"][[BN, M1, M2, M3, FE, IE, KE]],
                               np.array([N0, *S, G, E0, 0.5 * N0 * float(v @ v)]) * vol,
                               rtol=1e-13)
    np.testing.assert_allclose(t0["interior"][AM1:AM3 + 1],
                               np.cross(0.5 * np.add(lo, hi), S) * vol, rtol=1e-13)
    dt = 0.01
    g.step(0.0, dt)
    faces = g.boundary_fluence()
    for b in range(6):
        d = b // 2
        flux = np.array([N0 * v[d], *(S * v[d] + P * np.eye(3)[d]), (G + P) * v[d]])
        for f in range(5):
            np.testing.assert_allclose(faces[b][f], dt * flux[f], rtol=1e-13, atol=1e-17)


def test_conservation_at_size_and_repeatability():
    """128 x 64 x 64 shock tube drifting through outflow ends and into
    reflecting side walls: the totals with the boundary are conserved to
    round-off, and a second run from the same upload gives the same bits."""
    bc = (("OUTFLOW", "OUTFLOW"), ("REFLECTING", "REFLECTING"), ("REFLECTING", "OUTFLOW"))
    chart = Chart((128, 64, 64), bc=bc)
    shape = chart.n[::-1]
    x = np.broadcast_to(chart.center[0][2:-2], shape)
    N = np.where(x < 0.5, 1.0, 0.125)
    E = np.where(x < 0.5, 1.0, 0.1) / 0.4
    prim = (N, np.full(shape, 0.3), np.full(shape, 0.1), np.full(shape, -0.05), E)
    results = []
    for run in range(2):
        g = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8))
        g.set_primitive(*prim)
        t0 = g.compute_tally(False)
        for _ in range(40):
            g.step(0.0, g.time_step_candidate())
        t1 = g.compute_tally()
        results.append((t1, g.boundary_fluence()))
        scale = abs(t0["total"][FE])
        for t in CONSERVED_TALLIES:
            assert abs(t1["change"][t]) <= 1e-13 * scale, (t, t1["change"][t])
        assert abs(t1["interior"][BN] - t0["interior"][BN]) > 1e-3 * t0["total"][BN]
    for key in ("interior", "boundary", "total", "change"):
        assert np.array_equal(results[0][0][key], results[1][0][key]), key
    for b in range(6):
        assert np.array_equal(results[0][1][b], results[1][1][b])


def test_fp32_tallies_conserve():
    """FP32 state and sweeps, FP64 sums: fluence close to the FP64 run's, the
    totals with the boundary conserved to single-precision round-off."""
    chart = Chart((32, 24, 16), bc=MIXED)
    prim = tuple(chart.proper(a) for a in _state(chart, 3))
    runs, dts = {}, []
    for prec in ("FP64", "FP32"):
        g = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8, Precision=prec))
        g.set_primitive(*prim)
        t0 = g.compute_tally(False)
        for s in range(10):
            if prec == "FP64":
                dts.append(g.time_step_candidate())
            g.step(0.0, dts[s])
        runs[prec] = (t0, g.compute_tally(), g.boundary_fluence())
    t0, t1, faces = runs["FP32"]
    for t in CONSERVED_TALLIES:
        assert abs(t1["change"][t]) <= 1e-5 * abs(t0["total"][FE]), (t, t1["change"][t])
    for b in range(6):
        ref = runs["FP64"][2][b]
        assert np.abs(faces[b] - ref).max() <= 1e-4 * np.abs(ref).max() + 1e-12
