"""Shared drivers for the parity tests: the same seeded problem on the oracle
(CPU restatement) and on the HIP library, stepped with the same time steps."""
from __future__ import annotations

import numpy as np

from genasis_amd.chart import Chart
from genasis_amd.integrator import Integrator_C_PS, IntegratorParameters

import oracle


class OracleStepper:
    def __init__(self, o):
        self.o = o

    def time_step_candidate(self):
        return self.o.time_step_candidate()[0]

    def step(self, time, dt):
        self.o.step(dt)


def same_bits(a, b) -> bool:
    """FP64 arrays with identical bit patterns (so +0 and -0 differ)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def oracle_conserved(o, chart: Chart) -> np.ndarray:
    return np.stack([chart.proper(a) for a in o.conserved()])


def oracle_primitive(o, chart: Chart) -> np.ndarray:
    return np.stack([chart.proper(a) for a in o.primitive()])


def oracle_flag_bits(o, chart: Chart) -> np.ndarray:
    bits = np.zeros(chart.n[::-1], dtype=np.uint8)
    for d in range(3):
        bits |= (chart.proper(o.flags(d)) > 0).astype(np.uint8) << d
    return bits


def make_oracle(chart, prim, **kw):
    o = oracle.OracleFluid(chart, **kw)
    o.set_primitive(*prim)
    return o


def evolve(stepper, n_cycles, finish_time=1.0e30, n_write=1):
    it = Integrator_C_PS(stepper, IntegratorParameters(
        FinishTime=finish_time, FinishCycle=n_cycles, nWrite=n_write))
    it.evolve()
    return it


def cold_rarefaction(n=(24, 6, 6), v=2.0, p=1.0e-6):
    """A cold double rarefaction (N = 1, P = 1e-6, v_x = -v left of x = 0.5
    and +v right of it, a small transverse wobble v_y) on an outflow chart:
    with BaryonDensityMin = 0.05 its centre is repaired in both branches of
    ComputeFromConserved, at the stage-2 input and at StoreSolution, within
    a dozen full CFL steps."""
    chart = Chart(n, bc=(("OUTFLOW", "OUTFLOW"),) * 3)
    x, y, z = chart.padded_centers()
    shape = chart.m[::-1]
    N = np.ones(shape)
    E = np.full(shape, p / 0.4)
    V1 = np.where(x < 0.5, -v, v)
    V2 = 0.05 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * z)
    return chart, (N, V1, V2, np.zeros(shape), E)
