"""Pinning the oracle (CPU restatement of the reference step).

The reference ships no golden vectors for this path, so each check below
fixes the restatement by mathematics: closed forms, textbook solutions,
invariants and special cases, chosen so that a plausible slip (a dropped
term, a swapped width, a wrong sign, a transposed index) fails one of them.
"""
import math

import numpy as np
import pytest

import oracle
from genasis_amd import problems
from genasis_amd.chart import Chart

from . import exact_riemann
from ._helpers import OracleStepper, evolve, make_oracle, oracle_conserved

OUT = (("OUTFLOW", "OUTFLOW"),) * 3
PER = (("PERIODIC", "PERIODIC"),) * 3
REFL = (("REFLECTING", "REFLECTING"),) * 3


# --- limiter (Gradient_Kernel.f90:166-190) -------------------------------

def _slope_reference_form(dvl, dvr, dxl, dxr, theta):
    """The theta-MC slope written independently: zero unless both one-sided
    differences agree in sign, else the smallest of theta times the left and
    right slopes and the width-weighted centred slope."""
    if dvl * dvr <= 0.0:
        return 0.0
    sl, sr = dvl / dxl, dvr / dxr
    central = (dxr * sl + dxl * sr) / (dxl + dxr)
    return math.copysign(min(abs(theta * sl), abs(theta * sr), abs(central)), dvl)


def test_limiter_matches_independent_form():
    rng = np.random.default_rng(1)
    for _ in range(4000):
        dvl, dvr = rng.normal(size=2)
        dxl, dxr = rng.uniform(0.2, 3.0, size=2)
        theta = rng.uniform(1.0, 2.0)
        got = oracle.limited_slope(dvl, dvr, dxl, dxr, theta)
        ref = _slope_reference_form(dvl, dvr, dxl, dxr, theta)
        assert got == pytest.approx(ref, rel=1e-13, abs=1e-300)


def test_limiter_special_cases():
    # linear data on a uniform grid: the exact gradient for theta >= 1
    assert oracle.limited_slope(0.3, 0.3, 0.1, 0.1, 1.4) == pytest.approx(3.0, rel=1e-15)
    # extremum: zero slope
    assert oracle.limited_slope(0.3, -0.2, 0.1, 0.1, 1.4) == 0.0
    # one-sided zero difference: zero slope (sign(1/2, 0) = +1/2)
    assert oracle.limited_slope(0.0, 0.5, 0.1, 0.1, 1.4) == 0.0
    assert oracle.limited_slope(0.0, -0.5, 0.1, 0.1, 1.4) == 0.0
    # steep right jump: theta times the left slope limits
    assert oracle.limited_slope(0.1, 10.0, 1.0, 1.0, 1.4) == pytest.approx(0.14, rel=1e-15)
    # non-uniform widths: the centred slope weights the left slope by dxr
    dxl, dxr = 1.0, 3.0
    got = oracle.limited_slope(1.0, 2.0, dxl, dxr, 100.0)
    assert got == pytest.approx((dxr * 1.0 / dxl + dxl * 2.0 / dxr) / (dxl + dxr), rel=1e-15)


# --- one face (Current_Kernel.f90, Fluid_P__Kernel.f90) -------------------

def _euler_flux(prim, d, gamma=1.4):
    n, v1, v2, v3, e = prim
    v = [v1, v2, v3]
    p = (gamma - 1.0) * e
    s = [n * vi for vi in v]
    g = e + 0.5 * n * (v1 * v1 + v2 * v2 + v3 * v3)
    f = [n * v[d], s[0] * v[d], s[1] * v[d], s[2] * v[d], (g + p) * v[d]]
    f[1 + d] += p
    return np.array(f)


@pytest.mark.parametrize("riemann", ["HLL", "HLLC"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_face_flux_consistency(riemann, d):
    # equal states: the physical flux
    rng = np.random.default_rng(d)
    for _ in range(50):
        prim = [rng.uniform(0.1, 3), *rng.uniform(-1, 1, 3), rng.uniform(0.1, 5)]
        got = oracle.face_flux(prim, prim, d, riemann=riemann)
        np.testing.assert_allclose(got, _euler_flux(prim, d), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_face_flux_supersonic_upwind(d):
    # every wave moves right: F = F(left) for HLL and HLLC
    left = [1.0, 0.0, 0.0, 0.0, 2.5]
    left[1 + d] = 5.0
    right = [0.5, 0.0, 0.0, 0.0, 1.0]
    right[1 + d] = 4.0
    for riemann in ("HLL", "HLLC"):
        got = oracle.face_flux(left, right, d, riemann=riemann)
        np.testing.assert_allclose(got, _euler_flux(left, d), rtol=1e-14)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_resolves_contact(d):
    # isolated contact moving at v: HLLC gives the upwind flux exactly (to
    # rounding), HLL does not
    v = 0.3
    left = [2.0, 0.0, 0.0, 0.0, 2.5]
    right = [1.0, 0.0, 0.0, 0.0, 2.5]
    left[1 + d] = v
    right[1 + d] = v
    hllc = oracle.face_flux(left, right, d, riemann="HLLC")
    np.testing.assert_allclose(hllc, _euler_flux(left, d), rtol=1e-13)
    hll = oracle.face_flux(left, right, d, riemann="HLL")
    assert abs(hll[0] - 2.0 * v) > 1e-3
    # a shock flag on the face restores HLL
    flagged = oracle.face_flux(left, right, d, riemann="HLLC", df=1.0)
    assert np.array_equal(flagged, hll)
    # stationary contact: zero mass and energy flux, pressure in the momentum
    left[1 + d] = right[1 + d] = 0.0
    f = oracle.face_flux(left, right, d, riemann="HLLC")
    p = 0.4 * 2.5
    assert abs(f[0]) < 1e-15 and abs(f[4]) < 1e-15
    assert f[1 + d] == pytest.approx(p, rel=1e-14)


# --- geometry (Chart_Template.f90:463-593) -------------------------------

def test_geometry_closed_form_and_single_precision_ghosts():
    chart = Chart((10, 7, 3), lo=(0.0, -1.0, 2.0), hi=(1.0, 1.0, 5.0))
    o = oracle.OracleFluid(chart)
    for d in range(3):
        c, wl, wr = o.geometry(d)
        assert np.array_equal(c, chart.center[d])
        assert np.array_equal(wl, chart.width_left[d])
        assert np.array_equal(wr, chart.width_right[d])
        h = (chart.hi[d] - chart.lo[d]) / chart.n[d]
        interior = c[2:-2]
        np.testing.assert_allclose(interior, chart.lo[d] + (np.arange(chart.n[d]) + 0.5) * h,
                                   rtol=0, atol=1e-14 * max(1.0, abs(chart.hi[d])))
        # the first ghost width is the first interior width rounded to float32
        e = chart.edge[d]
        w_ig = float(np.float32(e[3] - e[2]))
        assert e[1] == e[2] - w_ig
        assert wl[1] + wr[1] == pytest.approx(w_ig, rel=1e-15)


# --- whole steps ---------------------------------------------------------

class _Recording(OracleStepper):
    def __init__(self, o):
        super().__init__(o)
        self.candidates = []

    def time_step_candidate(self):
        c = super().time_step_candidate()
        self.candidates.append(c)
        return c


def _sod_density_l1(n):
    chart = Chart((n, 4, 4), bc=OUT)
    o = make_oracle(chart, problems.sod(chart), courant=0.8)
    rec = _Recording(o)
    it = evolve(rec, 10**9, finish_time=0.2)
    it.candidates = rec.candidates
    assert it.s.Time == 0.2
    rho = chart.proper(o.field("N"))[0, 0, :]
    x = chart.center[0][2:-2]
    exact = exact_riemann.sample(x, 0.2, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))[0]
    return np.mean(np.abs(rho - exact)), it


def test_sod_against_exact_solution_and_converges():
    e128, _ = _sod_density_l1(128)
    e256, it = _sod_density_l1(256)
    assert e256 < 3.0e-3, e256
    assert e256 < 0.7 * e128, (e128, e256)
    # ramp: time step k is (k+1)/100 of the CFL candidate for k < 99, and
    # the last step lands on FinishTime exactly
    for k in range(3):
        assert it.s.history[k][2] == pytest.approx((k + 1) / 100.0 * it.candidates[k],
                                                   rel=1e-15)
    assert it.s.history[150][2] == pytest.approx(it.candidates[150], rel=1e-15)
    assert it.s.history[-1][1] == 0.2


def test_uniform_state_is_preserved_exactly():
    for bc in (OUT, PER, REFL):
        chart = Chart((8, 6, 5), bc=bc)
        shape = chart.m[::-1]
        prim = (np.full(shape, 1.3), np.full(shape, 0.0), np.full(shape, 0.0),
                np.full(shape, 0.0), np.full(shape, 2.0))
        if bc is not REFL:
            prim = (prim[0], np.full(shape, 0.4), np.full(shape, -0.2),
                    np.full(shape, 0.1), prim[4])
        o = make_oracle(chart, prim, courant=0.8)
        u0 = oracle_conserved(o, chart)
        for _ in range(3):
            o.step(0.5 * o.time_step_candidate()[0])
        assert np.array_equal(oracle_conserved(o, chart), u0)


def test_time_step_closed_form():
    chart = Chart((16, 16, 16), bc=PER)
    shape = chart.m[::-1]
    gamma, rho, p, v = 1.4, 1.0, 1.0, 0.3
    prim = (np.full(shape, rho), np.full(shape, v), np.zeros(shape), np.zeros(shape),
            np.full(shape, p / (gamma - 1.0)))
    o = make_oracle(chart, prim, courant=0.8)
    cand, inv = o.time_step_candidate()
    c = math.sqrt(gamma * p / rho)
    h = 1.0 / 16
    assert inv == pytest.approx((abs(v) + c) / h + 2 * c / h, rel=1e-14)
    assert cand == pytest.approx(0.8 / inv, rel=1e-15)


def test_periodic_conservation_to_round_off():
    chart = Chart((16, 16, 16), bc=PER)
    o = make_oracle(chart, problems.kelvin_helmholtz(chart, seed=3), courant=0.8)
    vol = 1.0 / 16**3
    t0 = [np.sum(u) * vol for u in oracle_conserved(o, chart)]
    for s in range(10):
        o.step(0.5 * o.time_step_candidate()[0])
    t1 = [np.sum(u) * vol for u in oracle_conserved(o, chart)]
    for f in range(5):
        assert abs(t1[f] - t0[f]) <= 1e-13 * max(1.0, abs(t0[f])), (f, t0[f], t1[f])


def test_reflecting_walls_conserve_mass_and_energy():
    chart = Chart((32, 4, 4), bc=REFL)
    o = make_oracle(chart, problems.sod(chart), courant=0.8)
    u0 = oracle_conserved(o, chart)
    for s in range(40):
        o.step(0.8 * o.time_step_candidate()[0])
    u1 = oracle_conserved(o, chart)
    for f in (0, 4):
        assert abs(u1[f].sum() - u0[f].sum()) <= 1e-13 * abs(u0[f].sum())


def _tube(axis, riemann, dts=None):
    n = [4, 4, 4]
    n[axis] = 48
    chart = Chart(tuple(n), bc=OUT)
    o = make_oracle(chart, problems.sod(chart, axis=axis), courant=0.8, riemann=riemann)
    used = []
    for s in range(15):
        dt = (s + 1) / 100.0 * o.time_step_candidate()[0] if dts is None else dts[s]
        o.step(dt)
        used.append(dt)
    u = oracle_conserved(o, chart)
    # bring the tube axis to the x position of a (z, y, x) array
    perm = {0: (0, 1, 2), 1: (0, 2, 1), 2: (2, 1, 0)}[axis]
    return [np.transpose(u[f], perm) for f in (0, 1 + axis, 4)], used


def test_sod_direction_symmetry():
    """The same tube along x, y and z, driven with the same time steps, gives
    transposed fields: bit-identical with HLL.  With HLLC the shock-flag
    clearing at the chart faces adds round-off transverse fluxes, so x and y
    (whose increments add in either order) agree bit for bit and z to
    round-off.  (The CFL sum over dimensions is ordered, so each tube's own
    time step differs in the last bits; hence the shared time steps.)"""
    for riemann in ("HLL", "HLLC"):
        ux, dts = _tube(0, riemann)
        uy, _ = _tube(1, riemann, dts)
        uz, _ = _tube(2, riemann, dts)
        for f in range(3):
            assert np.array_equal(ux[f], uy[f])
            if riemann == "HLL":
                assert np.array_equal(ux[f], uz[f])
            else:
                np.testing.assert_allclose(uz[f], ux[f], rtol=1e-13, atol=1e-15)


def test_shock_flags_mark_transverse_faces_only():
    chart = Chart((48, 8, 8), bc=OUT)
    o = make_oracle(chart, problems.sod(chart), courant=0.8)
    for s in range(20):
        o.step(0.8 * o.time_step_candidate()[0])
    df = [chart.proper(o.flags(d)) for d in range(3)]
    assert not df[0].any()
    assert df[1].any() and df[2].any()
    # cleared in the first and last two layers of every dimension
    for d in (1, 2):
        a = df[d]
        assert not a[:, :, :1].any() and not a[:, :, -1:].any()
        assert not a[:, :1, :].any() and not a[:, -1:, :].any()
        assert not a[:1].any() and not a[-1:].any()
    # near the shock only
    xs = np.where(df[1].any(axis=(0, 1)))[0]
    assert xs.min() > 24
