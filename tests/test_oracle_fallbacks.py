"""Pins the oracle's positivity fallbacks with closed forms.

ComputeFromConserved (F/Fluid_P__Kernel.f90:94-192): a cell with D <= N_Min
becomes vacuum (N = D = N_Min, V = S = E = G = 0); then E < 0 gives E = 0 with
G rebuilt as 0 + S.V/2.  The EOS (F/Fluid_P_I__Kernel.f90:249-264) gives
CS = T = 0 and SB = -sqrt(huge) unless N > 0 and P > 0.  A face state with
N < 0 is clamped to N = 0 after its EOS (F/Fluid_D__Kernel.f90:87-88).  Each
expected value below is written from those statements; a dropped assignment
(S left in a vacuum cell, G not rebuilt, the EOS guard missing) changes one
of them.
"""
import numpy as np
import pytest

import oracle
from genasis_amd.chart import Chart

from ._helpers import cold_rarefaction

SQRT_HUGE = np.sqrt(np.finfo(np.float64).max)
OUT = (("OUTFLOW", "OUTFLOW"),) * 3
GM1 = 1.4 - 1.0  # (gamma - 1) as the reference rounds it


def _set_conserved(cells, n_min):
    """A 4x4x4 outflow chart whose listed proper cells hold the given
    (D, S1, S2, S3, G); every other value holds a quiet state."""
    chart = Chart((4, 4, 4), bc=OUT)
    o = oracle.OracleFluid(chart, n_min=n_min)
    shape = chart.m[::-1]
    D = np.full(shape, 1.0)
    S = [np.zeros(shape) for _ in range(3)]
    G = np.full(shape, 2.5)
    for (x, y, z), (d, s1, s2, s3, g) in cells.items():
        i = (z + 2, y + 2, x + 2)
        D[i], S[0][i], S[1][i], S[2][i], G[i] = d, s1, s2, s3, g
    o.set_conserved(D, S[0], S[1], S[2], G)
    return o


def _at(o, name, cell):
    x, y, z = cell
    return o.field(name)[z + 2, y + 2, x + 2]


@pytest.mark.parametrize("n_min", [0.0, 0.05])
def test_vacuum_cells(n_min):
    cells = {(1, 1, 1): (0.5 * n_min, 0.3, -0.2, 0.1, 5.0),  # D below N_Min
             (2, 1, 1): (n_min, 0.3, -0.2, 0.1, 5.0),        # D = N_Min: not D > N_Min
             (1, 2, 1): (-1.0, 0.3, 0.2, 0.1, 5.0)}          # negative D
    o = _set_conserved(cells, n_min)
    for c in cells:
        assert _at(o, "N", c) == n_min and _at(o, "D", c) == n_min
        for f in ("V_1", "V_2", "V_3", "S_1", "S_2", "S_3", "E", "G", "P", "CS", "T", "MN"):
            assert _at(o, f, c) == 0.0, (c, f)
        for f in ("FEP_1", "FEP_2", "FEP_3", "FEM_1", "FEM_2", "FEM_3"):
            assert _at(o, f, c) == 0.0, (c, f)
        assert _at(o, "SB", c) == -SQRT_HUGE
        assert _at(o, "DS", c) == -SQRT_HUGE * n_min
    counts = o.fallback_counts()
    assert counts["store"] == (3, 0)


def test_negative_internal_energy():
    # D = 2, S = (2, 0, 0): V_1 = 1 and S.V/2 = 1 > G = 0.5, so E = -0.5 -> 0
    # and G = 0 + 0.5 * ((2*1 + 0*0) + 0*0) = 1
    cells = {(1, 1, 1): (2.0, 2.0, 0.0, 0.0, 0.5),
             # S along y and z: V = (0, 0.25, -0.5), S.V/2 = 0.3125 > G = 0.3
             (2, 2, 2): (2.0, 0.0, 0.5, -1.0, 0.3)}
    o = _set_conserved(cells, 0.0)
    c = (1, 1, 1)
    assert _at(o, "N", c) == 2.0 and _at(o, "V_1", c) == 1.0
    assert _at(o, "E", c) == 0.0 and _at(o, "G", c) == 1.0
    assert _at(o, "P", c) == 0.0 and _at(o, "CS", c) == 0.0 and _at(o, "T", c) == 0.0
    assert _at(o, "SB", c) == -SQRT_HUGE
    assert _at(o, "FEP_1", c) == 1.0 and _at(o, "FEM_1", c) == 1.0
    c = (2, 2, 2)
    assert (_at(o, "V_2", c), _at(o, "V_3", c)) == (0.25, -0.5)
    assert _at(o, "E", c) == 0.0
    assert _at(o, "G", c) == 0.0 + 0.5 * ((0.0 * 0.0 + 0.5 * 0.25) + (-1.0) * (-0.5))
    assert o.fallback_counts()["store"] == (0, 2)


def test_regular_cell_untouched():
    # D = 1, S = (0.5, 0, 0), G = 2.625: V_1 = 0.5, E = 2.625 - 0.125 = 2.5
    o = _set_conserved({(1, 1, 1): (1.0, 0.5, 0.0, 0.0, 2.625)}, 0.05)
    c = (1, 1, 1)
    assert _at(o, "E", c) == 2.5 and _at(o, "G", c) == 2.625
    P = GM1 * 2.5
    assert _at(o, "P", c) == P
    assert _at(o, "CS", c) == np.sqrt(1.4 * P / (1.0 * 1.0))
    assert o.fallback_counts()["store"] == (0, 0)


def _hll(FL, FR, UL, UR, AP, AM):
    return np.array([(AP * FL[f] + AM * FR[f] - 1.0 * AP * AM * (UR[f] - UL[f]))
                     / max(AP + AM, np.finfo(float).tiny) for f in range(5)])


@pytest.mark.parametrize("v", [0.5, -0.5])
def test_face_negative_pressure_gives_zero_sound_speed(v):
    # E < 0 on both sides: P < 0 fails the EOS guard, CS = 0 (a missing guard
    # gives sqrt of a negative number, NaN), so AP = max(0, v, v) and
    # AM = max(0, -v, -v): the flux is the upwind raw flux.
    W = np.array([1.0, v, 0.0, 0.0, -0.25])
    F = oracle.face_flux(W, W, 0, riemann="HLL")
    P = GM1 * -0.25
    S1 = 1.0 * 1.0 * v
    G = -0.25 + 0.5 * (S1 * v)
    raw = np.array([1.0 * v, S1 * v + P, 0.0, 0.0, (G + P) * v])
    U = np.array([1.0, S1, 0.0, 0.0, G])
    AP, AM = max(0.0, v), max(0.0, -v)
    assert np.all(np.isfinite(F))
    assert np.array_equal(F, _hll(raw, raw, U, U, AP, AM))


def test_face_negative_density_is_clamped():
    # N = -0.5: the EOS sees N < 0 (CS = 0), then N = 0, so D = S = 0, G = E
    W = np.array([-0.5, 0.5, 0.0, 0.0, 1.0])
    F = oracle.face_flux(W, W, 0, riemann="HLL")
    P = GM1 * 1.0
    raw = np.array([0.0, 0.0 + P, 0.0, 0.0, (1.0 + P) * 0.5])
    U = np.array([0.0, 0.0, 0.0, 0.0, 1.0])
    assert np.array_equal(F, _hll(raw, raw, U, U, 0.5, 0.0))


def test_cold_collision_hllc_star_state():
    # Cold streams v_L = +1/2, v_R = -1/2, N = 1, E = 0: CS = 0, AP = AM = 1/2.
    # MD = AP D_R + AM D_L - F_D,R + F_D,L = 2, SN = AP S_R + AM S_L - F_S,R
    # + F_S,L = 0, so AC = 0 and the left star state holds: AV = AM + v_L = 1,
    # inv = 1/|AM + AC| = 2, D* = 2, S* = 0, P* = 0 + 1*(S_L AV) - 0 = 1/2,
    # G* = (G_L AV + v_L P_L - AC P*) inv = 2 G_L = 1/4.  Flux of the star
    # state with V_d = AC = 0: (0, P*, 0, 0, 0).
    WL = np.array([1.0, 0.5, 0.0, 0.0, 0.0])
    WR = np.array([1.0, -0.5, 0.0, 0.0, 0.0])
    F = oracle.face_flux(WL, WR, 0, riemann="HLLC")
    assert np.array_equal(F, np.array([0.0, 0.5, 0.0, 0.0, 0.0]))
    # HLL instead: (AP F_L + AM F_R - AP AM (U_R - U_L)) / (AP + AM)
    G = 0.0 + 0.5 * (0.5 * 0.5)
    FL = np.array([0.5, 0.25, 0.0, 0.0, G * 0.5])
    FR = np.array([-0.5, 0.25, 0.0, 0.0, G * -0.5])
    UL = np.array([1.0, 0.5, 0.0, 0.0, G])
    UR = np.array([1.0, -0.5, 0.0, 0.0, G])
    assert np.array_equal(oracle.face_flux(WL, WR, 0, riemann="HLL"),
                          _hll(FL, FR, UL, UR, 0.5, 0.5))


def test_cold_rarefaction_enters_every_branch():
    """The workload of the device test: a cold double rarefaction with
    N_Min = 0.05 is repaired in both branches at both call sites (the stage-2
    input and StoreSolution) within 12 full CFL steps."""
    chart, prim = cold_rarefaction()
    o = oracle.OracleFluid(chart, courant=0.8, n_min=0.05)
    o.set_primitive(*prim)
    total = {"stage": np.zeros(2, int), "store": np.zeros(2, int)}
    for _ in range(12):
        o.step(o.time_step_candidate()[0])
        for k, v in o.fallback_counts(reset=True).items():
            total[k] += v
    assert (total["stage"] > 0).all() and (total["store"] > 0).all(), total
    D = chart.proper(o.field("D"))
    assert np.all(D >= 0.05) and np.all(np.isfinite(D))
