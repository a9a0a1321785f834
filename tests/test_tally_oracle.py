"""The oracle's conservation tallies and boundary fluence (Tally_F_P on one
chart, FLD/Current_ASC__Template.f90:239-349; RecordBoundaryFluence_CSL,
EV/Increments/IncrementDivergence_FV__Form.f90:1002-1069), pinned by what the
mathematics fixes: the totals of baryon number, momentum and fluid energy are
conserved once the boundary fluence is counted, periodic faces cancel, a
uniform state has closed-form fluences and integrals, the surface integral of
a constant fluence obeys the divergence theorem, and a translated chart shifts
the angular momentum by a x momentum."""
import numpy as np
import pytest

from genasis_amd.chart import Chart
from oracle import TALLY_N

from ._helpers import make_oracle

BN, M1, M2, M3, KE, AM1, AM2, AM3, GE, TE, FE, IE = range(TALLY_N)
CONSERVED_TALLIES = [BN, M1, M2, M3, FE]


def _state(chart, seed, shock=True):
    rng = np.random.default_rng(seed)
    shape = chart.m[::-1]
    N = np.exp(rng.uniform(np.log(0.5), np.log(2.0), shape))
    V = [rng.uniform(-0.3, 0.3, shape) for _ in range(3)]
    E = np.exp(rng.uniform(np.log(0.5), np.log(2.0), shape)) / 0.4
    if shock:
        x, _, _ = chart.padded_centers()
        N = N * (1 + 2 * (x > 0.5))
        E = E * (1 + 5 * (x > 0.5))
    return (N, V[0], V[1], V[2], E)


def _abs_scale(o, chart):
    """Volume integrals of |D|, |S|, |G|: the size of the tallied totals."""
    w = [chart.width_left[d][2:-2] + chart.width_right[d][2:-2] for d in range(3)]
    dv = w[2][:, None, None] * w[1][None, :, None] * w[0][None, None, :]
    return {t: float(np.sum(np.abs(chart.proper(o.field(f))) * dv))
            for t, f in ((BN, "D"), (M1, "S_1"), (M2, "S_2"), (M3, "S_3"), (FE, "G"))}


@pytest.mark.parametrize("bc", [
    (("OUTFLOW", "OUTFLOW"), ("REFLECTING", "REFLECTING"), ("OUTFLOW", "REFLECTING")),
    (("REFLECTING", "OUTFLOW"), ("PERIODIC", "PERIODIC"), ("OUTFLOW", "OUTFLOW")),
])
def test_totals_are_conserved_with_the_boundary_fluence(bc):
    chart = Chart((12, 10, 9), bc=bc)
    o = make_oracle(chart, _state(chart, 1), courant=0.8)
    t0 = o.compute_tally(compute_change=False)
    assert np.array_equal(t0["total"], t0["interior"])
    assert not np.any(t0["change"]) and not np.any(t0["boundary"])
    scale = _abs_scale(o, chart)
    for _ in range(10):
        o.step(o.time_step_candidate()[0])
    t1 = o.compute_tally()
    for t in CONSERVED_TALLIES:
        # the interior alone changes (matter leaves); interior + boundary does not
        assert abs(t1["interior"][t] - t0["interior"][t]) > 1e-6 * scale[t]
        assert abs(t1["change"][t]) <= 1e-14 * scale[t], (t, t1["change"][t])
        assert t1["total"][t] == t1["interior"][t] + t1["boundary"][t]
    assert t1["boundary"][KE] == 0.0 and t1["boundary"][IE] == 0.0
    assert not np.any(t1["interior"][[GE, TE]])
    # a second tally adds the next steps' boundary terms to the global one
    o.step(o.time_step_candidate()[0])
    si = o.surface_integral()
    t2 = o.compute_tally()
    np.testing.assert_array_equal(t2["boundary"], t1["boundary"] + si)
    for t in CONSERVED_TALLIES:
        assert abs(t2["change"][t]) <= 2e-14 * scale[t]


def test_periodic_faces_cancel():
    """With cell widths exact in single precision the ghost widths equal the
    proper ones, so each periodic face pair records the same fluence."""
    chart = Chart((16, 8, 8))
    o = make_oracle(chart, _state(chart, 2), courant=0.8)
    o.compute_tally(compute_change=False)
    scale = _abs_scale(o, chart)
    for _ in range(3):
        o.step(o.time_step_candidate()[0])
    for d in range(3):
        inner, outer = o.boundary_fluence(2 * d), o.boundary_fluence(2 * d + 1)
        assert np.array_equal(inner, outer) and np.any(inner)
    t = o.compute_tally()
    for k in CONSERVED_TALLIES:
        assert abs(t["boundary"][k]) <= 1e-15 * scale[k]
        assert abs(t["change"][k]) <= 1e-14 * scale[k]


def test_uniform_state_closed_forms():
    lo, hi = (0.5, -1.0, 2.0), (1.5, 0.5, 2.75)
    bc = (("OUTFLOW", "OUTFLOW"),) * 3
    chart = Chart((8, 6, 5), lo=lo, hi=hi, bc=bc)
    N0, v, E0, gamma = 1.3, np.array([0.2, -0.15, 0.1]), 2.1, 1.4
    shape = chart.m[::-1]
    prim = (np.full(shape, N0), *[np.full(shape, vi) for vi in v], np.full(shape, E0))
    o = make_oracle(chart, prim, gamma=gamma, courant=0.8)
    t0 = o.compute_tally(compute_change=False)
    vol = np.prod(np.subtract(hi, lo))
    D, S = N0, N0 * v
    G = E0 + 0.5 * N0 * float(v @ v)
    P = (gamma - 1.0) * E0
    xbar = 0.5 * (np.add(lo, hi))
    expect = np.zeros(TALLY_N)
    expect[BN], expect[M1:M3 + 1], expect[FE], expect[IE] = D * vol, S * vol, G * vol, E0 * vol
    expect[KE] = 0.5 * N0 * float(v @ v) * vol
    expect[AM1:AM3 + 1] = np.cross(xbar, S) * vol
    np.testing.assert_allclose(t0["interior"], expect, rtol=1e-13, atol=1e-15)
    dt = 0.01
    o.step(dt)
    for d in range(3):
        flux = np.array([D * v[d], *(S * v[d] + P * np.eye(3)[d]), (G + P) * v[d]])
        for side in range(2):
            bf = o.boundary_fluence(2 * d + side)
            for f in range(5):
                np.testing.assert_allclose(bf[f], dt * flux[f], rtol=1e-13, atol=1e-17)
    # a uniform flow has no net boundary term: as much enters as leaves
    t1 = o.compute_tally()
    np.testing.assert_allclose(t1["boundary"], 0.0, atol=1e-15)


def _slab(chart, b, f, value):
    d = b // 2
    fl = np.zeros((5,) + tuple(chart.n[e] for e in (2, 1, 0) if e != d))
    fl[f] = value
    return fl


def test_surface_integral_of_given_fluences():
    """Divergence theorem on the closed box: a constant fluence integrates to
    zero over the faces.  With S_2 = 1 on every face, L_3 = X_1 S_2 - X_2 S_1
    collects X_1 on the x faces, (hi_1 - lo_1) (area of an x face) = +volume;
    L_1 = X_2 S_3 - X_3 S_2 collects -X_3 on the z faces, -volume; L_2 = 0.
    On one face alone, L_3 takes that face's position: hi_1 (outer, +) or
    lo_1 (inner, -) times the face area."""
    lo, hi = (0.25, -0.5, 1.0), (1.25, 0.75, 1.5)
    chart = Chart((8, 5, 4), lo=lo, hi=hi, bc=(("OUTFLOW", "OUTFLOW"),) * 3)
    o = make_oracle(chart, _state(chart, 3, shock=False))
    vol = np.prod(np.subtract(hi, lo))
    for f, value in ((0, 1.0), (2, 1.0), (4, -3.0)):
        for b in range(6):
            o.set_boundary_fluence(b, _slab(chart, b, f, value))
        si = o.surface_integral()
        assert np.all(np.abs(si[[BN, M1, M2, M3, FE]]) <= 1e-15)
        if f == 2:  # S_2 = 1 on every face
            np.testing.assert_allclose(si[[AM1, AM2, AM3]], [-vol, 0.0, vol], rtol=1e-14, atol=1e-15)
        else:
            assert np.all(np.abs(si[[AM1, AM2, AM3]]) <= 1e-14)
        assert si[KE] == 0.0 and si[IE] == 0.0
    area_x = (hi[1] - lo[1]) * (hi[2] - lo[2])
    zbar = 0.5 * (lo[2] + hi[2])
    for b, sign, x1 in ((1, 1.0, hi[0]), (0, -1.0, lo[0])):
        for bb in range(6):
            o.set_boundary_fluence(bb, _slab(chart, bb, 2, 1.0 if bb == b else 0.0))
        si = o.surface_integral()
        np.testing.assert_allclose(si[[M2, AM1, AM2, AM3]],
                                   sign * area_x * np.array([1.0, -zbar, 0.0, x1]),
                                   rtol=1e-14, atol=1e-15)


def test_translated_chart_shifts_angular_momentum():
    """A translation by a (exact in binary) leaves the step's bits unchanged
    and adds a x (momentum) to the angular momentum, interior and boundary."""
    bc = (("OUTFLOW", "REFLECTING"), ("OUTFLOW", "OUTFLOW"), ("REFLECTING", "OUTFLOW"))
    a = np.array([2.0, -3.0, 5.0])
    ca = Chart((16, 8, 8), bc=bc)
    cb = Chart((16, 8, 8), lo=tuple(a), hi=tuple(a + 1.0), bc=bc)
    prim = _state(ca, 4)
    oa, ob = make_oracle(ca, prim, courant=0.8), make_oracle(cb, prim, courant=0.8)
    ta0, tb0 = oa.compute_tally(False), ob.compute_tally(False)
    for _ in range(4):
        dt = oa.time_step_candidate()[0]
        oa.step(dt)
        ob.step(dt)
    for f in ("D", "S_1", "S_2", "S_3", "G"):
        assert np.array_equal(oa.field(f), ob.field(f))
    ta, tb = oa.compute_tally(), ob.compute_tally()
    for key in ("interior", "boundary"):
        A, B = ta[key], tb[key]
        shift = np.cross(a, A[M1:M3 + 1])
        scale = np.abs(a).max() * np.abs(A[M1:M3 + 1]).max() + np.abs(A[AM1:AM3 + 1]).max()
        np.testing.assert_allclose(B[AM1:AM3 + 1], A[AM1:AM3 + 1] + shift, rtol=0,
                                   atol=1e-13 * scale, err_msg=key)
        np.testing.assert_array_equal(B[[BN, M1, M2, M3, FE, KE, IE]], A[[BN, M1, M2, M3, FE, KE, IE]])
    shift0 = np.cross(a, ta0["interior"][M1:M3 + 1])
    np.testing.assert_allclose(tb0["interior"][AM1:AM3 + 1], ta0["interior"][AM1:AM3 + 1] + shift0,
                               rtol=1e-12)
