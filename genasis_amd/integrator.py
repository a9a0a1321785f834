"""Host time loop of the conserved-current integrator.

Mirrors ``Integrator_Template % Evolve`` and its helpers for one rank
(Modules/Mathematics/Solvers/DifferentialEquations/Evolutions/Integrators/
Integrator_Template.f90:291-346 Evolve, 560-610 checkpoint bookkeeping,
842-893 ComputeNewTime, 896-939 ComputeTimeStep; Integrator_C_PS__Form.f90:
328-371 ComputeCycle_ASC).  The stepper supplies the two calls that cross the
boundary into the device:

* ``time_step_candidate() -> float`` -- ``ComputeTimeStep_C_ASC``
  (CourantFactor * min(huge, 1/max(tiny, invdt)));
* ``step(time, dt)`` -- ``Step_RK_C_ASC % Compute``.

The MPI MIN of the candidate over ranks is the identity on one rank.
"""
from __future__ import annotations

import math
import sys
from dataclasses import dataclass, field

_TINY = sys.float_info.min
_HUGE = sys.float_info.max


@dataclass
class IntegratorParameters:
    """Defaults of IntegratorHeader_Form (EvolutionBasics/IntegratorHeader_Form.
    f90:95-131): FinishTime 1, FinishCycle huge, nWrite 100, nRampCycles 100,
    CheckpointTimeExact false, StartTime 0."""
    FinishTime: float = 1.0
    FinishCycle: int = 2**31 - 1
    nWrite: int = 100
    nRampCycles: int = 100
    CheckpointTimeExact: bool = False
    StartTime: float = 0.0


@dataclass
class IntegratorState:
    Time: float = 0.0
    iCycle: int = 0
    CheckpointTime: float = 0.0
    CheckpointTimeInterval: float = 0.0
    CheckpointTimeExact: bool = False
    IsCheckpointTime: bool = False
    iCheckpoint: int = 0
    history: list = field(default_factory=list)


class Integrator_C_PS:
    """Time loop over a stepper; ``history`` records (iCycle, Time, dt)."""

    def __init__(self, stepper, params: IntegratorParameters | None = None):
        self.stepper = stepper
        self.p = params or IntegratorParameters()
        self.s = IntegratorState(Time=self.p.StartTime,
                                 CheckpointTimeExact=self.p.CheckpointTimeExact)
        self.last_candidate = _HUGE

    # Integrator_Template.f90:595-607 (end of AdministerCheckpoint)
    def _administer_checkpoint(self):
        s, p = self.s, self.p
        s.IsCheckpointTime = False
        if s.Time < p.FinishTime:
            s.CheckpointTimeInterval = (p.FinishTime - p.StartTime) / p.nWrite
            s.CheckpointTime = min(s.Time + s.CheckpointTimeInterval, p.FinishTime)
            if s.CheckpointTime == p.FinishTime:
                s.CheckpointTimeExact = True
        s.iCheckpoint += 1

    # Integrator_Template.f90:476-502 (PrepareInitial with RestartFrom) and
    # 966-996 (ResetInitial): Time, iCycle and iCheckpoint from the file; the
    # next evolve() administers the checkpoint that follows.
    def restart_from(self, path: str, exact: bool = False):
        t, cycle, i_ck = self.stepper.read_checkpoint(path, exact=exact)
        self.s.Time = t
        self.s.iCycle = cycle
        self.s.iCheckpoint = i_ck
        return self.s

    def write_checkpoint(self, path: str):
        """The stepper's checkpoint file with this integrator's Time, iCycle and
        iCheckpoint (the reference writes at AdministerCheckpoint)."""
        self.stepper.write_checkpoint(path, self.s.Time, self.s.iCycle, self.s.iCheckpoint)

    # Integrator_Template.f90:896-939
    def compute_time_step(self) -> float:
        candidate = self.stepper.time_step_candidate()
        self.last_candidate = candidate
        time_step = min(_HUGE, candidate)
        ramp = min(float(self.s.iCycle + 1) / self.p.nRampCycles, 1.0)
        if ramp < 1.0:
            time_step = ramp * time_step
        return time_step

    # Integrator_Template.f90:842-893
    def compute_new_time(self) -> float:
        time_step = self.compute_time_step()
        s = self.s
        if s.CheckpointTimeExact and s.Time + time_step > s.CheckpointTime:
            time_step = s.CheckpointTime - s.Time
        return s.Time + time_step

    # Integrator_C_PS__Form.f90:328-371
    def compute_cycle(self):
        s = self.s
        time_new = self.compute_new_time()
        time_step = time_new - s.Time
        self.stepper.step(s.Time, time_step)
        s.iCycle += 1
        s.Time = s.Time + time_step
        s.history.append((s.iCycle, s.Time, time_step))
        if s.CheckpointTimeExact:
            if s.Time == s.CheckpointTime:
                s.IsCheckpointTime = True
        else:
            if (s.Time > s.CheckpointTime
                    or abs(s.Time - s.CheckpointTime) < 0.5 * time_step):
                s.IsCheckpointTime = True
        return time_step

    # Integrator_Template.f90:291-346
    def evolve(self, max_cycles: int | None = None):
        s, p = self.s, self.p
        self._administer_checkpoint()
        done = 0
        while s.Time < p.FinishTime and s.iCycle < p.FinishCycle:
            time_step = self.compute_cycle()
            done += 1
            ratio = self.last_candidate / max(s.CheckpointTimeInterval,
                                              math.sqrt(_TINY))
            if ratio < 1.0e-6 * p.nWrite:
                self._administer_checkpoint()
                break
            if s.IsCheckpointTime:
                self._administer_checkpoint()
            if max_cycles is not None and done >= max_cycles:
                break
        return s
