"""Host time loop (Integrator_Template / Integrator_C_PS) on a scripted
stepper: ramp, checkpoint clipping, cycle limit, tiny-step guard."""
import pytest

from genasis_amd.integrator import Integrator_C_PS, IntegratorParameters


class Scripted:
    def __init__(self, candidate):
        self.candidate = candidate
        self.calls = []

    def time_step_candidate(self):
        return self.candidate

    def step(self, time, dt):
        self.calls.append((time, dt))


def test_ramp_and_final_clip():
    s = Scripted(0.01)
    it = Integrator_C_PS(s, IntegratorParameters(FinishTime=0.2, nWrite=1))
    it.evolve()
    dts = [c[1] for c in s.calls]
    assert dts[0] == pytest.approx(0.01 * 0.01)
    assert dts[49] == pytest.approx(0.01 * 0.5)
    assert all(d <= 0.01 for d in dts)
    assert it.s.Time == 0.2                      # lands on FinishTime exactly
    assert sum(dts) == pytest.approx(0.2)


def test_finish_cycle_and_checkpoints_not_exact():
    s = Scripted(1e-3)
    it = Integrator_C_PS(s, IntegratorParameters(FinishTime=1.0, FinishCycle=7, nWrite=100))
    it.evolve()
    assert it.s.iCycle == 7 and len(s.calls) == 7
    # not the final interval: no clipping to the checkpoint time
    assert not it.s.CheckpointTimeExact


def test_tiny_time_step_stops_the_loop():
    s = Scripted(1e-12)
    it = Integrator_C_PS(s, IntegratorParameters(FinishTime=1.0, nWrite=10))
    it.evolve()
    assert it.s.iCycle == 1
