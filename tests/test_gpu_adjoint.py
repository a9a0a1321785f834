"""The discrete adjoint on the GPU: gradients of the band kinetic energy (and
of a general linear functional of the final state) with respect to the
initial primitives and the adiabatic index, against torch autograd through
the FP64 transcription (oracle/torch_step.py, whose forward values are the
oracle's bit for bit) and against central differences of the HIP forward."""
import numpy as np
import pytest
import torch

from genasis_amd import problems
from genasis_amd.adjoint import DiscreteAdjoint, kinetic_energy_gradient, steps_function
from genasis_amd.chart import Chart
from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC
from oracle.torch_step import TorchFluid

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _torch_grad(chart, prim, n_steps, ubar=None, hllc=True):
    tf = TorchFluid(chart, courant=0.8, hllc=hllc)
    P0 = torch.tensor(np.stack(prim)).requires_grad_(True)
    g = torch.tensor(1.4, dtype=F64, requires_grad=True)
    U, _ = tf.run(P0, g, n_steps)
    J = tf.kinetic_energy_vy(U) if ubar is None else torch.sum(torch.tensor(ubar) * U)
    gP, gg = torch.autograd.grad(J, (P0, g))
    return float(J), gP.numpy(), float(gg), U.detach().numpy()


def test_kh_gradient_matches_autograd():
    chart = Chart((12, 16, 10))
    prim = problems.kelvin_helmholtz(chart, seed=3, padded=False)
    n_steps, every = 7, 3
    st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8))
    st.set_primitive(*prim)
    adj = DiscreteAdjoint(st)
    adj.forward(n_steps, every)
    loss, ub = adj.kinetic_energy(component=1, lo=0.2, hi=0.8)
    pb, gb = adj.backward(ub)
    J, gP, gg, U = _torch_grad(chart, prim, n_steps)
    assert np.array_equal(st.conserved(), U)  # same forward, bit for bit
    assert abs(loss - J) <= 1e-14 * abs(J)
    np.testing.assert_allclose(pb[0], gP[0], rtol=1e-9, atol=1e-12 * np.abs(gP[0]).max())
    assert abs(gb - gg) <= 1e-9 * abs(gg), (gb, gg)
    # the step can be run again from the same upload: the second pass agrees
    loss2, dN, dg = kinetic_energy_gradient(chart, prim, n_steps, every, stepper=st)
    assert loss2 == loss and np.array_equal(dN, pb[0]) and dg == gb


@pytest.mark.parametrize("n", [(11, 9, 7), (16, 8, 8)])
@pytest.mark.parametrize("riemann", ["HLLC", "HLL"])
def test_general_functional_mixed_boundaries(riemann, n):
    """Shock, mixed boundaries, all five primitive bars; (16, 8, 8) on the
    unit cube has a power-of-two geometry, whose line passes replay the
    slope quotients as products."""
    bc = (("REFLECTING", "OUTFLOW"), ("OUTFLOW", "REFLECTING"), ("PERIODIC", "PERIODIC"))
    chart = Chart(n, bc=bc)
    rng = np.random.default_rng(12)
    shape = chart.n[::-1]
    x = chart.center[0][2:-2].reshape(1, 1, -1)
    N = np.exp(rng.uniform(np.log(0.5), np.log(2.0), shape)) * (1 + 2 * (x > 0.5))
    V = [rng.uniform(-0.3, 0.3, shape) for _ in range(3)]
    E = np.exp(rng.uniform(np.log(0.5), np.log(2.0), shape)) / 0.4 * (1 + 5 * (x > 0.5))
    prim = (N, V[0], V[1], V[2], E)  # a shock: flags switch HLLC to HLL
    ubar = rng.standard_normal((5,) + shape)
    n_steps = 5
    st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8, RiemannSolverType=riemann))
    assert st.geometry_pow2() == (n == (16, 8, 8))
    st.set_primitive(*prim)
    adj = DiscreteAdjoint(st)
    adj.forward(n_steps, 2)
    pb, gb = adj.backward(ubar)
    _, gP, gg, U = _torch_grad(chart, prim, n_steps, ubar=ubar, hllc=riemann == "HLLC")
    assert np.array_equal(st.conserved(), U)
    for f in range(5):
        np.testing.assert_allclose(pb[f], gP[f], rtol=1e-9,
                                   atol=1e-11 * np.abs(gP[f]).max(), err_msg=f"field {f}")
    assert abs(gb - gg) <= 1e-9 * abs(gg)


def test_gradient_matches_central_differences():
    """Kelvin-Helmholtz: directions constant in each density layer (along
    which the step is differentiable at the uniform initial state) and the
    adiabatic index, step 1e-6."""
    chart = Chart((16, 16, 12))
    prim = problems.kelvin_helmholtz(chart, seed=3, padded=False)
    n_steps, every = 8, 4
    loss, dN, dg = kinetic_energy_gradient(chart, prim, n_steps, every)
    yc = chart.center[1][2:-2]
    inner = (np.abs(yc - 0.5) < 0.25).reshape(1, -1, 1) * np.ones(chart.n[::-1])
    rng = np.random.default_rng(1)
    h = 1e-6

    def J(N, gamma=1.4):
        st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8, AdiabaticIndex=gamma))
        st.set_primitive(N, *prim[1:])
        a = DiscreteAdjoint(st)
        a.forward(n_steps, every)
        return a.kinetic_energy()[0]

    for _ in range(3):
        a, b = rng.standard_normal(2)
        v = a * inner + b * (1 - inner)
        fd = (J(prim[0] + h * v) - J(prim[0] - h * v)) / (2 * h)
        ad = float(np.sum(dN * v))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)
    fdg = (J(prim[0], 1.4 + h) - J(prim[0], 1.4 - h)) / (2 * h)
    assert abs(fdg - dg) <= 1e-6 * abs(dg), (fdg, dg)


def test_autograd_function_wraps_the_adjoint():
    chart = Chart((10, 12, 8))
    prim = problems.kelvin_helmholtz(chart, seed=9, padded=False)
    Steps = steps_function()
    N0 = torch.tensor(prim[0]).requires_grad_(True)
    g = torch.tensor(1.4, dtype=F64, requires_grad=True)
    U = Steps.apply(N0, g, chart, prim[1:], 4, 2, FluidParameters(CourantFactor=0.8))
    tf = TorchFluid(chart, courant=0.8)
    J = tf.kinetic_energy_vy(U)
    gN, gg = torch.autograd.grad(J, (N0, g))
    Jt, gP, ggt, _ = _torch_grad(chart, prim, 4)
    np.testing.assert_allclose(gN.numpy(), gP[0], rtol=1e-9, atol=1e-12 * np.abs(gP[0]).max())
    assert abs(float(gg) - ggt) <= 1e-9 * abs(ggt)


def test_kelvin_helmholtz_32_50_steps_matches_autograd():
    """The differentiation configuration at 32^3: 50 RK2 steps, checkpoints
    every 10, gradient of the vy kinetic energy in 0.2 < y < 0.8 with respect
    to the initial density and the adiabatic index."""
    chart = Chart((32, 32, 32))
    prim = problems.kelvin_helmholtz(chart, seed=1234, padded=False)
    loss, dN, dg = kinetic_energy_gradient(chart, prim, 50, 10)
    J, gP, gg, _ = _torch_grad(chart, prim, 50)
    assert abs(loss - J) <= 1e-13 * abs(J)
    np.testing.assert_allclose(dN, gP[0], rtol=1e-8, atol=1e-11 * np.abs(gP[0]).max())
    assert abs(dg - gg) <= 1e-8 * abs(gg), (dg, gg)


def test_device_resident_seed_and_result():
    """The seed dJ/dU left on the device by the observable and the result
    kept there give the same gradient as the host round trip; a consumed
    seed is refused."""
    from genasis_amd._capi import GxError
    chart = Chart((12, 10, 8))
    prim = problems.kelvin_helmholtz(chart, seed=5, padded=False)
    st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8))
    st.set_primitive(*prim)
    adj = DiscreteAdjoint(st)
    adj.forward(5, 2)
    loss, ub = adj.kinetic_energy()
    pb, gb = adj.backward(ub)
    loss2, none = adj.kinetic_energy(keep_on_device=True)
    assert none is None and loss2 == loss
    pb2, gb2 = adj.backward(None, fetch=False)
    assert pb2 is None and gb2 == gb
    assert np.array_equal(adj.get_bar(), pb)
    with pytest.raises(GxError):
        adj.backward(None)
    st.close()


@pytest.mark.parametrize("every", [1, 4])
def test_segment_lengths(every):
    """One step per segment (every reverse step starts with the fix kernel,
    buffer sets alternate per step) and one segment holding every step (the
    forward run keeps it; nothing is recomputed), against autograd."""
    chart = Chart((10, 12, 8))
    prim = problems.kelvin_helmholtz(chart, seed=21, padded=False)
    n_steps = 4
    loss, dN, dg = kinetic_energy_gradient(chart, prim, n_steps, every)
    J, gP, gg, _ = _torch_grad(chart, prim, n_steps)
    assert abs(loss - J) <= 1e-14 * abs(J)
    np.testing.assert_allclose(dN, gP[0], rtol=1e-9, atol=1e-12 * np.abs(gP[0]).max())
    assert abs(dg - gg) <= 1e-9 * abs(gg), (dg, gg)


def test_repeated_runs_reuse_or_replace_buffers():
    """Runs of the same shape on one context reuse its checkpoint and segment
    buffers, runs of another shape replace them; each gradient equals that of
    a fresh context bit for bit."""
    chart = Chart((12, 10, 8))
    prim = problems.kelvin_helmholtz(chart, seed=8, padded=False)

    def grad(adj, st, n_steps, every):
        st.set_primitive(*prim)
        adj.forward(n_steps, every)
        loss, ub = adj.kinetic_energy()
        pb, gb = adj.backward(ub)
        return loss, pb, gb

    fresh = {}
    for shape in [(6, 2), (5, 3)]:
        st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8))
        fresh[shape] = grad(DiscreteAdjoint(st), st, *shape)
        st.close()
    st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8))
    adj = DiscreteAdjoint(st)
    for shape in [(6, 2), (6, 2), (5, 3), (6, 2)]:
        loss, pb, gb = grad(adj, st, *shape)
        f_loss, f_pb, f_gb = fresh[shape]
        assert loss == f_loss and gb == f_gb and np.array_equal(pb, f_pb), shape
    st.close()
