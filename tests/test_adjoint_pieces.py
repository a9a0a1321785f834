"""The reverse sweeps of the adjoint (evaluated on the host by the library,
the same code the device runs) against torch autograd of the transcription
(oracle/torch_step.py), on random faces and slopes."""
import ctypes

import numpy as np
import pytest
import torch

from genasis_amd import _capi
from genasis_amd.chart import Chart
from oracle.torch_step import TorchFluid

F64 = torch.float64


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.fixture(scope="module")
def lib():
    return _capi.load()


def _face_states(rng, kind):
    N = rng.uniform(0.3, 2.0, 2)
    V = rng.uniform(-1.0, 1.0, (2, 3))
    E = rng.uniform(0.3, 3.0, 2)
    if kind == "fast":      # supersonic: AM = 0 -> upwind
        V[:, :] = 4.0
    if kind == "contact":   # stationary contact
        V[:, :] = 0.0
        E[1] = E[0]
    if kind == "vacuum":    # N <= 0 on one side
        N[1] = -0.1
    return ([N[0], *V[0], E[0]], [N[1], *V[1], E[1]])


@pytest.mark.parametrize("hllc", [0, 1])
@pytest.mark.parametrize("kind", ["generic", "fast", "contact", "vacuum"])
def test_face_flux_vjp(lib, hllc, kind):
    rng = np.random.default_rng(hash((hllc, kind)) % 2**32)
    tf = TorchFluid(Chart((4, 4, 4)), hllc=bool(hllc))
    for trial in range(20):
        wl, wr = _face_states(rng, kind)
        d = trial % 3
        df = int(trial % 5 == 0)
        fb = rng.standard_normal(5)
        WL = [torch.tensor([x], dtype=F64, requires_grad=True) for x in wl]
        WR = [torch.tensor([x], dtype=F64, requires_grad=True) for x in wr]
        g = torch.tensor(1.4, dtype=F64, requires_grad=True)
        F = tf.face_flux(WL, WR, d, torch.tensor([bool(df)]), g)
        J = sum(float(fb[f]) * F[f].sum() for f in range(5))
        grads = torch.autograd.grad(J, WL + WR + [g], allow_unused=True)
        want = np.array([0.0 if x is None else float(x) for x in grads])
        WLb, WRb, Fo = np.zeros(5), np.zeros(5), np.zeros(5)
        gb = ctypes.c_double()
        lib.gx_adjoint_face_vjp(_pd(np.array(wl)), _pd(np.array(wr)), d, df, 1.4, 1.0,
                                hllc, _pd(fb), _pd(WLb), _pd(WRb), ctypes.byref(gb), _pd(Fo))
        got = np.concatenate([WLb, WRb, [gb.value]])
        assert np.array_equal(Fo, np.array([float(x.detach()) for x in F])), (kind, trial)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max(),
                                   err_msg=f"{kind} trial {trial}")


def test_slope_vjp(lib):
    rng = np.random.default_rng(8)
    tf = TorchFluid(Chart((4, 4, 4)))
    cases = []
    for _ in range(300):
        cases.append(list(rng.uniform(-1, 1, 3)))
    cases += [[1.0, 1.0, 1.0], [1.0, 1.0, 2.0], [0.5, 1.0, 1.0], [1.0, 2.0, 1.0],
              [0.0, 1.0, 1.3], [2.0, 1.0, 1.0]]
    for lim in (1, 0):
        tf.use_limiter = bool(lim)
        for vm, v0, vp in cases:
            dxl, dxr = rng.uniform(0.05, 0.2, 2)
            x = [torch.tensor(v, dtype=F64, requires_grad=True) for v in (vm, v0, vp)]
            s = tf.slope(x[0], x[1], x[2], torch.tensor(dxl, dtype=F64),
                         torch.tensor(dxr, dtype=F64))
            want = np.array([float(t) for t in torch.autograd.grad(s, x)])
            out = np.zeros(3)
            lib.gx_adjoint_slope_vjp(vm, v0, vp, dxl, dxr, 1.4, lim, 1.0, _pd(out))
            np.testing.assert_allclose(out, want, rtol=1e-13, atol=1e-13)
