"""The oracle's PARABOLIC reconstruction (EV/Increments/IncrementDivergence_
FV__Kernel.f90:296-596) pinned by an independent derivation: each parabola is
built in vertex form (Lagrange interpolation for the three-point parabolas,
(x - x0)^2 for the ones flat at a neighbour or a face) instead of the
reference's Cramer-rule coefficients, on random non-uniform stencils that
visit every branch; quadratic data are reproduced exactly; and whole steps
with PARABOLIC keep uniform states, conserve, and converge to the exact Sod
solution."""
import numpy as np
import pytest

import oracle
from genasis_amd.chart import Chart
from genasis_amd import problems

from ._helpers import OracleStepper, make_oracle
from . import exact_riemann


def _lagrange(xs, vs):
    """Coefficients (a, b, c) of a x^2 + b x + c through three points."""
    return np.polyfit(np.array(xs), np.array(vs), 2)


def _vertex(coef):
    a, b, _ = coef
    return -b / (2.0 * a) if a != 0.0 else np.inf


def _independent(vm, vc, vp, xm, xc, xp, dxl, dxr, margin=1e-7):
    """Face values by vertex forms; None when a branch decision lies within
    `margin` of its boundary (then rounding may pick either side)."""
    xi, xo = xc - dxl, xc + dxr
    if (vc - vm) * (vp - vc) <= 0.0:
        return vc, vc
    scale = xp - xm

    def near(a, b):
        return abs(a - b) < margin * scale

    p1 = _lagrange((xm, xc, xp), (vm, vc, vp))
    xe = _vertex(p1)
    if any(near(xe, b) for b in (xm, xc, xp)):
        return None
    vi, vo = np.polyval(p1, xi), np.polyval(p1, xo)
    if xm < xe <= xc:  # flat at xm, through vc
        vi = vm + (vc - vm) * ((xi - xm) / (xc - xm)) ** 2
    elif xc < xe < xp:  # flat at xp, through vc
        vo = vp + (vc - vp) * ((xo - xp) / (xc - xp)) ** 2
    p2 = _lagrange((xi, xc, xo), (vi, vc, vo))
    xe = _vertex(p2)
    if any(near(xe, b) for b in (xi, xc, xo)):
        return None
    vi, vo = np.polyval(p2, xi), np.polyval(p2, xo)
    if xi < xe <= xc:  # flat at xi through (xc, vc) and (xo, vo)
        B = (vc - vo) / ((xc - xi) ** 2 - (xo - xi) ** 2)
        vi = vc - B * (xc - xi) ** 2
    elif xc < xe < xo:  # flat at xo through (xc, vc) and (xi, vi)
        B = (vc - vi) / ((xc - xo) ** 2 - (xi - xo) ** 2)
        vo = vc - B * (xc - xo) ** 2
    return vi, vo


def test_matches_an_independent_vertex_form_derivation():
    rng = np.random.default_rng(11)
    seen = set()
    checked = 0
    for _ in range(20000):
        w = rng.uniform(0.5, 2.0, 3)
        xc = rng.uniform(-3.0, 3.0)
        dxl = 0.5 * w[1] * rng.uniform(0.8, 1.2)
        dxr = w[1] - dxl
        xm, xp = xc - 0.5 * (w[0] + w[1]), xc + 0.5 * (w[1] + w[2])
        v = rng.normal(size=3) * rng.choice([1e-3, 1.0, 1e3])
        ref = _independent(v[0], v[1], v[2], xm, xc, xp, dxl, dxr)
        if ref is None:
            continue
        got = oracle.parabolic_face_values(v[0], v[1], v[2], xm, xc, xp, dxl, dxr)
        s = np.abs(v).max()
        assert abs(got[0] - ref[0]) <= 1e-9 * s and abs(got[1] - ref[1]) <= 1e-9 * s, \
            (v, xm, xc, xp, dxl, dxr, got, ref)
        checked += 1
        # which branches this sample took
        p1 = _lagrange((xm, xc, xp), v)
        xe = _vertex(p1)
        seen.add(("extremum" if (v[1] - v[0]) * (v[2] - v[1]) <= 0 else
                  "left1" if xm < xe <= xc else "right1" if xc < xe < xp else "plain1"))
    assert checked > 15000
    assert seen == {"extremum", "left1", "right1", "plain1"}, seen


def test_second_parabola_branches_are_visited():
    """Monotone data whose face values make the second parabola turn inside
    the cell: the vertex-form result (flat at xI or xO) is the oracle's."""
    rng = np.random.default_rng(5)
    hits = {"left2": 0, "right2": 0}
    for _ in range(20000):
        xm, xc, xp = -1.0, 0.0, 1.0
        dxl = dxr = 0.5
        v = np.sort(rng.normal(size=3))
        if rng.random() < 0.5:
            v = v[::-1]
        ref = _independent(*v, xm, xc, xp, dxl, dxr)
        if ref is None or (v[1] - v[0]) * (v[2] - v[1]) <= 0:
            continue
        p1 = _lagrange((xm, xc, xp), v)
        xe1 = _vertex(p1)
        vi, vo = np.polyval(p1, -0.5), np.polyval(p1, 0.5)
        if -1.0 < xe1 <= 0.0:
            vi = v[0] + (v[1] - v[0]) * 0.25
        elif 0.0 < xe1 < 1.0:
            vo = v[2] + (v[1] - v[2]) * 0.25
        xe2 = _vertex(_lagrange((-0.5, 0.0, 0.5), (vi, v[1], vo)))
        key = "left2" if -0.5 < xe2 <= 0.0 else "right2" if 0.0 < xe2 < 0.5 else None
        if key is None:
            continue
        got = oracle.parabolic_face_values(*v, xm, xc, xp, dxl, dxr)
        assert np.allclose(got, ref, rtol=0, atol=1e-9 * np.abs(v).max()), (v, got, ref)
        hits[key] += 1
    assert hits["left2"] > 10 and hits["right2"] > 10, hits


@pytest.mark.parametrize("shift", [0.0, 7.25])
def test_quadratic_data_are_reproduced(shift):
    """Cell values of a monotone quadratic at the centres (the reference
    passes XA = X, X2A = X**2, so the values are point values): the faces get
    the quadratic's values."""
    rng = np.random.default_rng(2)
    for _ in range(500):
        w = rng.uniform(0.5, 2.0, 3)
        xc = shift + rng.uniform(-1, 1)
        xm, xp = xc - 0.5 * (w[0] + w[1]), xc + 0.5 * (w[1] + w[2])
        dxl, dxr = 0.5 * w[1], 0.5 * w[1]
        # vertex beyond the stencil: the data are monotone over it
        x0 = xp + rng.uniform(0.5, 3.0) if rng.random() < 0.5 else xm - rng.uniform(0.5, 3.0)
        a = rng.uniform(-2, 2)
        q = lambda x: a * (x - x0) ** 2 + 1.0
        vi, vo = oracle.parabolic_face_values(q(xm), q(xc), q(xp), xm, xc, xp, dxl, dxr)
        assert abs(vi - q(xc - dxl)) <= 1e-11 * (1 + abs(q(xc))), (vi, q(xc - dxl))
        assert abs(vo - q(xc + dxr)) <= 1e-11 * (1 + abs(q(xc))), (vo, q(xc + dxr))


def test_linear_data_and_extrema():
    vi, vo = oracle.parabolic_face_values(1.0, 2.0, 3.0, 0.0, 1.0, 2.0, 0.5, 0.5)
    assert (vi, vo) == (1.5, 2.5)
    assert oracle.parabolic_face_values(1.0, 2.0, 1.0, 0.0, 1.0, 2.0, 0.5, 0.5) == (2.0, 2.0)
    assert oracle.parabolic_face_values(2.0, 2.0, 3.0, 0.0, 1.0, 2.0, 0.5, 0.5) == (2.0, 2.0)


def test_whole_step_uniform_states_and_conservation():
    for bc in ((("PERIODIC", "PERIODIC"),) * 3,
               (("OUTFLOW", "REFLECTING"), ("REFLECTING", "OUTFLOW"), ("OUTFLOW", "OUTFLOW"))):
        chart = Chart((12, 10, 9), bc=bc)
        shape = chart.m[::-1]
        v = (0.0, 0.0, 0.0) if "REFLECTING" in str(bc) else (0.2, -0.1, 0.3)
        prim = (np.full(shape, 1.3), *[np.full(shape, a) for a in v], np.full(shape, 2.0))
        o = make_oracle(chart, prim, courant=0.8, reconstruction="PARABOLIC")
        u0 = [chart.proper(a).copy() for a in o.conserved()]
        for _ in range(3):
            o.step(o.time_step_candidate()[0])
        for a, b in zip(u0, o.conserved()):
            assert np.array_equal(a, chart.proper(b))
    chart = Chart((16, 8, 8))
    rng = np.random.default_rng(3)
    shape = chart.m[::-1]
    prim = (np.exp(rng.uniform(-0.5, 0.5, shape)), *[rng.uniform(-0.3, 0.3, shape) for _ in range(3)],
            np.exp(rng.uniform(-0.5, 0.5, shape)) / 0.4)
    o = make_oracle(chart, prim, courant=0.8, reconstruction="PARABOLIC")
    t0 = o.compute_tally(False)
    for _ in range(5):
        o.step(o.time_step_candidate()[0])
    t1 = o.compute_tally()
    for k in (0, 1, 2, 3, 10):
        assert abs(t1["change"][k]) <= 1e-14 * abs(t0["total"][10]), k


def _sod_l1(n, reconstruction):
    from ._helpers import evolve
    chart = Chart((n, 4, 4), bc=(("OUTFLOW", "OUTFLOW"),) * 3)
    o = make_oracle(chart, problems.sod(chart), courant=0.8, reconstruction=reconstruction)
    it = evolve(OracleStepper(o), 10**9, finish_time=0.2)
    assert it.s.Time == 0.2
    x = chart.center[0][2:-2]
    rho = chart.proper(o.field("N"))[0, 0, :]
    exact = exact_riemann.sample(x, 0.2, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))[0]
    return float(np.mean(np.abs(rho - exact)))


def test_sod_converges_to_the_exact_solution():
    e128, e256 = _sod_l1(128, "PARABOLIC"), _sod_l1(256, "PARABOLIC")
    assert e256 < 3e-3 and e256 < 0.75 * e128, (e128, e256)
