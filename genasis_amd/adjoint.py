"""Discrete adjoint of the FP64 step (gx_adjoint_* of include/gx_fluid.h).

``DiscreteAdjoint`` drives one ``Step_RK2_C_ASC``: ``forward`` runs n cycles
from the uploaded primitives with the integrator's time steps (CFL candidate
times the nRampCycles ramp), keeping a checkpoint every k steps;
``backward(Ubar)`` returns the bars of the uploaded primitives and of the
adiabatic index for Ubar = dJ/dU_final.  ``StepsFunction`` wraps the pair as
a ``torch.autograd.Function`` (CPU tensors in and out), for checking the
adjoint against other differentiations.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from .fluid import FluidParameters, Step_RK2_C_ASC, _pd


class DiscreteAdjoint:
    def __init__(self, stepper: Step_RK2_C_ASC):
        if stepper.params.Precision != "FP64":
            raise ValueError("the adjoint needs a FP64 stepper")
        self.st = stepper
        self.lib = stepper.lib
        self.shape = stepper.shape

    def forward(self, n_steps: int, checkpoint_every: int = 10, n_ramp: int = 100):
        _capi.check(self.lib.gx_adjoint_forward(self.st._h, int(n_steps),
                                                int(checkpoint_every), int(n_ramp)),
                    "gx_adjoint_forward")

    def backward(self, Ubar: np.ndarray | None = None, fetch: bool = True):
        """(P0bar, gamma bar) for Ubar = dJ/dU_final.  Ubar None takes the seed
        that ``kinetic_energy(keep_on_device=True)`` left on the device;
        fetch False leaves P0bar there (returned as None; ``get_bar`` copies
        it later)."""
        ub = None
        if Ubar is not None:
            ub = np.ascontiguousarray(Ubar, dtype=np.float64).reshape((5,) + self.shape)
        pb = np.empty((5,) + self.shape) if fetch else None
        gb = ctypes.c_double()
        _capi.check(self.lib.gx_adjoint_backward(self.st._h, None if ub is None else _pd(ub),
                                                 None if pb is None else _pd(pb),
                                                 ctypes.byref(gb)), "gx_adjoint_backward")
        return pb, gb.value

    def get_bar(self):
        """P0bar of the last reverse pass."""
        pb = np.empty((5,) + self.shape)
        _capi.check(self.lib.gx_adjoint_get_bar(self.st._h, _pd(pb)), "gx_adjoint_get_bar")
        return pb

    def kinetic_energy(self, component: int = 1, lo: float = 0.2, hi: float = 0.8,
                       keep_on_device: bool = False):
        """(J, dJ/dU) for J = sum over lo < y_c < hi of 1/2 N V_(component+1)^2 dV
        on the current state.  dJ/dU also stays on the device as the seed of
        the next ``backward(None)``; keep_on_device True returns None for it."""
        loss = ctypes.c_double()
        ub = None if keep_on_device else np.empty((5,) + self.shape)
        _capi.check(self.lib.gx_adjoint_kinetic_energy(self.st._h, int(component), float(lo),
                                                       float(hi), ctypes.byref(loss),
                                                       None if ub is None else _pd(ub)),
                    "gx_adjoint_kinetic_energy")
        return loss.value, ub


def kinetic_energy_gradient(chart, prim, n_steps: int, checkpoint_every: int = 10,
                            params: FluidParameters | None = None, component: int = 1,
                            lo: float = 0.2, hi: float = 0.8, stepper=None):
    """J (band kinetic energy after n_steps) and its gradients with respect
    to the initial density field and the adiabatic index."""
    st = stepper or Step_RK2_C_ASC(chart, params or FluidParameters(CourantFactor=0.8))
    st.set_primitive(*prim)
    adj = DiscreteAdjoint(st)
    adj.forward(n_steps, checkpoint_every)
    loss, ub = adj.kinetic_energy(component, lo, hi)
    pb, gb = adj.backward(ub)
    return loss, pb[0], gb


def _torch_function():
    import torch

    class StepsFunction(torch.autograd.Function):
        """U_final = Steps(N0, gamma) with the other primitives fixed; the
        backward is the HIP adjoint."""

        @staticmethod
        def forward(ctx, N0, gamma, chart, rest, n_steps, every, params):
            p = FluidParameters(**{**params.__dict__, "AdiabaticIndex": float(gamma)})
            st = Step_RK2_C_ASC(chart, p)
            prim = [N0.detach().cpu().numpy()] + [np.asarray(r) for r in rest]
            st.set_primitive(*prim)
            adj = DiscreteAdjoint(st)
            adj.forward(n_steps, every)
            ctx.adj = adj
            return torch.from_numpy(st.conserved())

        @staticmethod
        def backward(ctx, Ubar):
            pb, gb = ctx.adj.backward(Ubar.detach().cpu().numpy())
            return (torch.from_numpy(pb[0]), torch.tensor(gb, dtype=torch.float64),
                    None, None, None, None, None)

    return StepsFunction


def steps_function():
    """The torch.autograd.Function of the adjoint (imports torch on use)."""
    return _torch_function()
