"""The torch FP64 transcription of the step (oracle/torch_step.py), the
checker of the discrete adjoint: its forward values are the oracle's, bit for
bit, and its autograd gradient agrees with central differences."""
import sys

import numpy as np
import pytest
import torch

from genasis_amd.chart import Chart
from genasis_amd import problems
from oracle.torch_step import TorchFluid

from ._helpers import make_oracle, oracle_conserved

F64 = torch.float64


def _state(chart, seed):
    rng = np.random.default_rng(seed)
    shape = chart.m[::-1]
    N = np.exp(rng.uniform(np.log(0.5), np.log(2.0), shape))
    V = [rng.uniform(-0.3, 0.3, shape) for _ in range(3)]
    E = np.exp(rng.uniform(np.log(0.5), np.log(2.0), shape)) / 0.4
    x, _, _ = chart.padded_centers()
    N = N * (1 + 2 * (x > 0.5))
    E = E * (1 + 5 * (x > 0.5))  # a shock: flags switch HLLC to HLL
    return (N, V[0], V[1], V[2], E)


@pytest.mark.parametrize("bc", [
    (("PERIODIC", "PERIODIC"),) * 3,
    (("REFLECTING", "OUTFLOW"), ("OUTFLOW", "REFLECTING"), ("PERIODIC", "PERIODIC")),
])
def test_forward_is_the_oracle_bitwise(bc):
    chart = Chart((12, 10, 9), bc=bc)
    prim = _state(chart, 1)
    o = make_oracle(chart, prim, courant=0.8)
    tf = TorchFluid(chart, courant=0.8)
    P0 = torch.tensor(np.stack([chart.proper(a) for a in prim]))
    U, dts = tf.run(P0, torch.tensor(1.4, dtype=F64), 6)
    t = 0.0
    for s in range(6):
        c = o.time_step_candidate()[0]
        ramp = min((s + 1) / 100.0, 1.0)
        ts = min(sys.float_info.max, c)
        ts = ramp * ts if ramp < 1.0 else ts
        dt = (t + ts) - t
        t = t + dt
        assert dt == float(dts[s]), s
        o.step(dt)
    assert np.array_equal(oracle_conserved(o, chart), U.detach().numpy())


def test_autograd_matches_central_differences():
    """At the uniform layers of the Kelvin-Helmholtz state every density slope
    of the first step is at a limiter tie (both differences zero): the step
    is continuous but not differentiable there along directions that vary
    inside a layer.  Directions constant in each layer keep it
    differentiable; the adiabatic index too."""
    chart = Chart((10, 12, 8))
    prim = problems.kelvin_helmholtz(chart, seed=4, padded=False)
    tf = TorchFluid(chart, courant=0.8)
    P0 = torch.tensor(np.stack(prim))
    n_steps = 4

    def loss(N, gamma):
        P = torch.cat([N.unsqueeze(0), P0[1:]])
        U, _ = tf.run(P, gamma, n_steps)
        return tf.kinetic_energy_vy(U)

    N = P0[0].clone().requires_grad_(True)
    gamma = torch.tensor(1.4, dtype=F64, requires_grad=True)
    L = loss(N, gamma)
    gN, gg = torch.autograd.grad(L, (N, gamma))
    g = chart.g
    yc = np.asarray(chart.center[1][g:g + chart.n[1]])
    inner = torch.tensor((np.abs(yc - 0.5) < 0.25).astype(float)).reshape(1, -1, 1)
    inner = inner.expand(chart.n[::-1])
    rng = np.random.default_rng(5)
    h = 1e-6
    with torch.no_grad():
        for _ in range(3):
            a, b = rng.standard_normal(2)
            v = a * inner + b * (1 - inner)
            fd = (loss(P0[0] + h * v, gamma) - loss(P0[0] - h * v, gamma)) / (2 * h)
            ad = torch.sum(gN * v)
            assert abs(float(fd - ad)) <= 1e-6 * abs(float(ad)), (float(fd), float(ad))
        fdg = (loss(P0[0], gamma + h) - loss(P0[0], gamma - h)) / (2 * h)
        assert abs(float(fdg - gg)) <= 1e-6 * abs(float(gg)), (float(fdg), float(gg))
