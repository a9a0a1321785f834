"""Benchmark: PLM + HLLC RK2 FP64 step of the Fluid_P_I fluid on one MI355X.

Workload: 3-D Kelvin-Helmholtz on a 256^3 periodic unit cube, gamma 1.4,
CourantFactor 0.8, LimiterParameter 1.4, nRampCycles 100, seeded vy
perturbation (genasis_amd.problems.kelvin_helmholtz).  One step = one RK2
cycle of the integrator: CFL time step (device max-reduction, read back),
two fused stage kernels, shock-flag kernels.  The state is resident in HBM
before the timed region; ``warmup`` cycles run first, then ``steps`` cycles
are timed between device synchronisations.

Prints one JSON line.  ``roofline`` prices the dominant kernel (the stage-2
kernel) against HBM: algorithmic bytes per launch (120 B per cell: read Y and
U + K1/2, write U_new, FP64, five fields) over its average duration from HIP
events in the timed region.  ``cpu_baseline`` times the oracle (the CPU
restatement of the reference step) on the host cores: the same
Kelvin-Helmholtz problem at 256^3 for 5 steps, and the Sod workload
(256 x 16 x 16) to FinishTime 0.2, with the host's CPU model, sockets, cores
and OpenMP threads.
``fp32`` runs the same seeded problem at 512^3 for 200 steps with the state
and sweeps in FP32 and in FP64, with the same time steps, and reports the
speed ratio and the per-field L1 / Linf error of FP32 against FP64.
``parabolic`` runs the 256^3 problem with PARABOLIC reconstruction beside
LINEAR.  ``bricks`` runs it on one brick and split into 2x2x1, 2x2x2 and
4x4x2 bricks in the GPU's memory (halo exchange per stage) with the same
time steps, and reports the speed of each and whether the final states are
bit-identical.
``batch`` advances an ensemble of 256 seeded 3-D Riemann problems (cubic
edges log-uniform in [32, 128], outflow, FP64) for 50 steps with the batch
entry points (autotuned tile shapes and z-chunks, HIP graph of the cycles) and, for
comparison, the same problems one after another through single-chart
contexts driven by the host integrator.
``adjoint`` takes the gradient of the vy kinetic energy in 0.2 < y < 0.8
after 50 steps of the 128^3 problem with respect to the initial density and
the adiabatic index (discrete adjoint, checkpoints every 10 steps), reports
its cost over the forward run's, and checks it against central differences
(step 1e-6) along 8 seeded random directions that are constant in each
density layer (the step is not differentiable at the uniform layers along
directions that vary inside them) and along the adiabatic index.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("cell-updates/s, PLM+HLLC RK2 FP64 256^3; %HBM peak; FP32 err; "
          "adjoint/fwd")
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell and launch (FP64, five conserved fields):
# stage 1 reads U and writes U + K1 (80 B; U + K1/2 is an extra 40 B written
# because the reference adds K1/2 and K2/2 to U one after the other), stage 2
# reads U + K1 and U + K1/2 and writes U_new (120 B).
ALG_BYTES_STAGE2 = 120.0
ALG_BYTES_STEP = 200.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage_kernel_kh256_fp64_traffic.txt")
SQ_FILE = os.path.join(ROOT, "profiles", "stage_kernel_kh256_fp64_sq_counters.txt")
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak


def _read_traffic():
    """HBM bytes per stage-2 launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("stage2_hbm_bytes_per_launch"):
                    return float(line.split("=")[1])
    except (OSError, ValueError):
        pass
    return None


def _read_sq(key):
    """A per-launch figure of the stage-2 kernel at 256^3 from the committed
    rocprofv3 SQ counter summary: FP64 FLOP (64 x SQ_INSTS_VALU_FLOPS_FP64)
    or wave VALU instructions (SQ_INSTS_VALU)."""
    try:
        with open(SQ_FILE) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split("=")[1])
    except (OSError, ValueError):
        pass
    return None


SIMDS, CLOCK_HZ = 1024, 2.4e9  # MI355X: 256 CUs x 4 SIMDs, peak engine clock


def host_cpu():
    """The host's CPU as /proc/cpuinfo and the affinity mask give it (the
    mask of the calling thread: read it before OpenMP binds that thread)."""
    info = {"model": None, "sockets": None, "physical_cores": None,
            "logical_cpus": os.cpu_count(),
            "cpus_available": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else None}
    try:
        phys, cores = set(), set()
        with open("/proc/cpuinfo") as f:
            pid = cid = None
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and info["model"] is None:
                    info["model"] = v
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    cid = v
                elif not k and pid is not None:
                    phys.add(pid)
                    cores.add((pid, cid))
                    pid = cid = None
        info["sockets"] = len(phys) or None
        info["physical_cores"] = len(cores) or None
    except OSError:
        pass
    return info


def cpu_baseline(n: int, steps: int, threads: int):
    """The oracle (the C/OpenMP restatement of the reference step, the CPU
    path this build has: the reference's Fortran needs MPI and Silo modules
    the image lacks) on the host's cores: the Kelvin-Helmholtz workload at
    n^3 (steps timed after one untimed step) and the Sod workload (256 x 16 x
    16, outflow, CFL 0.8) to FinishTime 0.2."""
    # OpenMP placement as the plan for the reference's CPU runs sets it; read
    # when the oracle's OpenMP runtime starts
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle
    from genasis_amd import problems
    from genasis_amd.chart import Chart
    from genasis_amd.integrator import Integrator_C_PS, IntegratorParameters

    chart = Chart((n, n, n))
    o = oracle.OracleFluid(chart, courant=0.8, n_threads=threads)
    o.set_primitive(*problems.kelvin_helmholtz(chart, seed=1234))
    c = o.time_step_candidate()[0]
    o.step(0.01 * c)  # untimed: first touch of the oracle's arrays
    t0 = time.perf_counter()
    for s in range(steps):
        c = o.time_step_candidate()[0]
        o.step((s + 2) / 100.0 * c)
    t1 = time.perf_counter()
    o.close()
    kh = chart.n_proper * steps / (t1 - t0)

    class _Stepper:
        def __init__(self, o):
            self.o = o

        def time_step_candidate(self):
            return self.o.time_step_candidate()[0]

        def step(self, time_, dt):
            self.o.step(dt)

    sod_chart = Chart((256, 16, 16), bc=(("OUTFLOW", "OUTFLOW"),) * 3)
    o = oracle.OracleFluid(sod_chart, courant=0.8, n_threads=threads)
    o.set_primitive(*problems.sod(sod_chart))
    it = Integrator_C_PS(_Stepper(o), IntegratorParameters(FinishTime=0.2, nWrite=1))
    t0 = time.perf_counter()
    it.evolve()
    t1 = time.perf_counter()
    o.close()
    sod = sod_chart.n_proper * it.s.iCycle / (t1 - t0)
    return kh, t1 - t0, {"value": sod, "unit": "cell-updates/s", "cycles": it.s.iCycle,
                         "seconds": t1 - t0,
                         "workload": "Sod 256x16x16 outflow, CFL 0.8, to FinishTime 0.2 "
                                     "(integrator with the 100-cycle ramp)"}


def _ramped(candidate: float, cycle: int, n_ramp: int = 100) -> float:
    """Integrator_C_PS time-step ramp (min((iCycle+1)/nRampCycles, 1))."""
    ramp = min(float(cycle + 1) / n_ramp, 1.0)
    return ramp * candidate if ramp < 1.0 else candidate


def parabolic_check(n: int, steps: int, z_chunk: int):
    """ReconstructionType PARABOLIC and LINEAR on the seeded Kelvin-Helmholtz
    problem in FP64 (same time steps): speed of each, per-kernel times."""
    from genasis_amd import problems
    from genasis_amd.chart import Chart
    from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC

    chart = Chart((n, n, n))
    prim = problems.kelvin_helmholtz(chart, seed=1234, padded=False)
    dts, runs = [], {}
    for recon in ("LINEAR", "PARABOLIC"):
        st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8, ReconstructionType=recon))
        if z_chunk > 0:
            st.set_tuning(z_chunk)
        st.set_primitive(*prim)
        st.set_profiling(True)
        st.sync()
        t0 = time.perf_counter()
        for s in range(steps):
            c = st.time_step_candidate()
            if recon == "LINEAR":
                dts.append(_ramped(c, s))
            st.step(0.0, dts[s])
        st.sync()
        t1 = time.perf_counter()
        kt = st.kernel_times()
        runs[recon] = {"ms_per_step": (t1 - t0) / steps * 1e3,
                       "kernel_ms_avg": {k: v[0] / max(v[1], 1) for k, v in kt.items()}}
        st.close()
        print(f"# reconstruction {recon}: {runs[recon]['ms_per_step']:.3f} ms/step",
              file=sys.stderr, flush=True)
    return {
        "workload": f"Kelvin-Helmholtz {n}^3 periodic, seed 1234, {steps} RK2 steps FP64, "
                    "ReconstructionType PARABOLIC (LINEAR with the same time steps beside it)",
        "value": chart.n_proper / (runs["PARABOLIC"]["ms_per_step"] * 1e-3),
        "unit": "cell-updates/s",
        "linear_value": chart.n_proper / (runs["LINEAR"]["ms_per_step"] * 1e-3),
        "parabolic_over_linear_time": runs["PARABOLIC"]["ms_per_step"] / runs["LINEAR"]["ms_per_step"],
        "kernel_ms_avg": runs["PARABOLIC"]["kernel_ms_avg"],
    }


def precision_check(n: int, steps: int, z_chunk: int):
    """FP64 and FP32 GPU runs of the seeded Kelvin-Helmholtz problem with the
    FP64 run's time steps; speed and per-field error of FP32 against FP64."""
    from genasis_amd import problems
    from genasis_amd.chart import Chart
    from genasis_amd.compare import field_errors
    from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC

    chart = Chart((n, n, n))
    prim = problems.kelvin_helmholtz(chart, seed=1234, padded=False)
    dts, runs, states = [], {}, {}
    for prec in ("FP64", "FP32"):
        st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8, Precision=prec))
        if z_chunk > 0:
            st.set_tuning(z_chunk)
        st.set_primitive(*prim)
        st.set_profiling(True)
        st.sync()
        t0 = time.perf_counter()
        for s in range(steps):
            c = st.time_step_candidate()
            if prec == "FP64":
                dts.append(_ramped(c, s))
            st.step(0.0, dts[s])
        st.sync()
        t1 = time.perf_counter()
        kt = st.kernel_times()
        runs[prec] = {"ms_per_step": (t1 - t0) / steps * 1e3,
                      "kernel_ms_per_step": sum(v[0] for v in kt.values()) / steps,
                      "stage2_ms": kt["stage2"][0] / max(kt["stage2"][1], 1)}
        states[prec] = st.conserved()
        st.close()
        print(f"# precision check {prec}: {runs[prec]['ms_per_step']:.3f} ms/step",
              file=sys.stderr, flush=True)
    del prim
    err = field_errors(states["FP32"], states["FP64"])
    return {
        "workload": f"Kelvin-Helmholtz {n}^3 periodic, seed 1234, {steps} RK2 steps, "
                    "same time steps (FP64 CFL) for both runs",
        "value": chart.n_proper / (runs["FP32"]["ms_per_step"] * 1e-3),
        "unit": "cell-updates/s",
        "speedup_vs_fp64": runs["FP64"]["ms_per_step"] / runs["FP32"]["ms_per_step"],
        "kernel_speedup_vs_fp64": runs["FP64"]["kernel_ms_per_step"]
        / runs["FP32"]["kernel_ms_per_step"],
        "fp64": runs["FP64"],
        "fp32": runs["FP32"],
        "density_L1": err["D"]["L1"],
        "errors_vs_fp64": err,
    }


def bricks_check(n: int, steps: int, splits):
    """The seeded Kelvin-Helmholtz problem on one brick and split into bricks
    in the GPU's memory (halo exchange per stage), with the one-brick run's
    time steps: speed of each and whether the final states are identical."""
    import numpy as np
    from genasis_amd import problems
    from genasis_amd.chart import Chart
    from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC

    chart = Chart((n, n, n))
    prim = problems.kelvin_helmholtz(chart, seed=1234, padded=False)
    dts, runs, ref = [], {}, None
    for split in [None] + list(splits):
        st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8), n_bricks=split)
        st.set_primitive(*prim)
        for s in range(2):  # warm-up, then the timed steps from the same state
            c = st.time_step_candidate()
            st.step(0.0, 1e-3 * c)
        st.set_primitive(*prim)
        st.sync()
        t0 = time.perf_counter()
        for s in range(steps):
            c = st.time_step_candidate()
            if split is None:
                dts.append(_ramped(c, s))
            st.step(0.0, dts[s])
        st.sync()
        t1 = time.perf_counter()
        u = st.conserved()
        st.close()
        key = "x".join(map(str, split)) if split else "1x1x1"
        if ref is None:
            ref = u
        runs[key] = {"ms_per_step": (t1 - t0) / steps * 1e3,
                     "bitwise_equal_to_one_brick": bool(np.array_equal(u.view(np.uint64),
                                                                       ref.view(np.uint64)))}
        print(f"# bricks {key}: {runs[key]['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)
    one = runs["1x1x1"]["ms_per_step"]
    for v in runs.values():
        v["time_over_one_brick"] = v["ms_per_step"] / one
    return {
        "workload": f"Kelvin-Helmholtz {n}^3 periodic, seed 1234, {steps} RK2 steps FP64 "
                    "(after an untimed warm-up), one brick and split into bricks in one "
                    "GPU's HBM with a two-layer halo copy per stage; the one-brick run's "
                    "time steps for all",
        "runs": runs,
    }


def batch_check(n_problems: int, steps: int, seed: int, z_chunk: int, sequential: bool):
    """Ensemble of Riemann problems: batch throughput and the sequential one."""
    from genasis_amd import problems
    from genasis_amd.batch import Batch_RK2_C_ASC
    from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC
    from genasis_amd.integrator import Integrator_C_PS, IntegratorParameters

    sizes = problems.riemann_ensemble_sizes(n_problems, seed)
    made = [problems.riemann_ensemble_problem(k, n, seed) for k, n in enumerate(sizes)]
    charts = [m[0] for m in made]
    params = FluidParameters(CourantFactor=0.8)
    batch = Batch_RK2_C_ASC(charts, params)

    def upload():
        for k, (_, prim, _) in enumerate(made):
            batch.set_primitive(k, *prim)

    upload()
    t0 = time.perf_counter()
    batch.autotune(n_repeat=2)
    t_tune = time.perf_counter() - t0
    tuning = {str(c): {"tile": [t[0], t[1]], "z_chunk": t[2]}
              for c, t in batch.tiles().items()}
    per_call = 10
    calls = max(1, steps // per_call)
    batch.step(per_call)  # untimed: graph capture and first launches
    batch.sync()
    upload()              # the timed run starts from the initial data
    batch.step(0)         # first CFL maxima, outside the timed region
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(calls):
        batch.step(per_call)
    batch.sync()
    t1 = time.perf_counter()
    n_steps = calls * per_call
    cells = batch.cells
    _, _, cyc = batch.times()
    assert int(cyc.min()) == n_steps, cyc
    batch.close()
    out = {
        "workload": f"{n_problems} 3-D Riemann problems (seed {seed}), cubic edges "
                    f"{min(sizes)}..{max(sizes)} log-uniform in [32, 128], outflow, "
                    f"PLM+HLLC RK2 FP64, CFL 0.8, {n_steps} steps",
        "cells": cells,
        "value": cells * n_steps / (t1 - t0),
        "unit": "cell-updates/s",
        "ms_per_step": (t1 - t0) / n_steps * 1e3,
        "tuning_by_size_class": tuning,
        "autotune_s": t_tune,
        "sequential_value": None,
        "speedup_vs_sequential": None,
    }
    if sequential:
        elapsed = 0.0
        for chart, prim, _ in made:
            st = Step_RK2_C_ASC(chart, params)
            if z_chunk > 0:
                st.set_tuning(z_chunk)
            st.set_primitive(*prim)
            it = Integrator_C_PS(st, IntegratorParameters(FinishTime=1.0, nWrite=1))
            it._administer_checkpoint()
            st.sync()
            t0 = time.perf_counter()
            for _ in range(n_steps):
                it.compute_cycle()
            st.sync()
            elapsed += time.perf_counter() - t0
            st.close()
        out["sequential_value"] = cells * n_steps / elapsed
        out["speedup_vs_sequential"] = out["value"] / out["sequential_value"]
    print(f"# batch: {out['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    return out


def adjoint_check(n: int, steps: int, every: int, n_dirs: int, seed: int):
    """Discrete adjoint of the Kelvin-Helmholtz run: cost and FD agreement."""
    import numpy as np
    from genasis_amd import problems
    from genasis_amd.adjoint import DiscreteAdjoint
    from genasis_amd.chart import Chart
    from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC
    from genasis_amd.integrator import Integrator_C_PS, IntegratorParameters

    chart = Chart((n, n, n))
    prim = problems.kelvin_helmholtz(chart, seed=1234, padded=False)
    st = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8))

    def plain_forward():
        st.set_primitive(*prim)
        it = Integrator_C_PS(st, IntegratorParameters(FinishTime=1.0e30, nWrite=1))
        st.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            it.compute_cycle()
        st.sync()
        return time.perf_counter() - t0

    plain_forward()  # warm-up
    t_fwd = min(plain_forward() for _ in range(2))
    adj = DiscreteAdjoint(st)
    # warm-up, as for the plain forward: the first run allocates the
    # checkpoints and segment buffers, later runs of the same shape reuse them
    st.set_primitive(*prim)
    adj.forward(steps, every)
    adj.kinetic_energy(component=1, lo=0.2, hi=0.8, keep_on_device=True)
    adj.backward(None, fetch=False)
    st.set_primitive(*prim)
    st.sync()
    t0 = time.perf_counter()
    adj.forward(steps, every)
    t1 = time.perf_counter()
    # the seed dJ/dU and the result stay in HBM, as the forward's state does
    loss, _ = adj.kinetic_energy(component=1, lo=0.2, hi=0.8, keep_on_device=True)
    _, gb = adj.backward(None, fetch=False)
    t2 = time.perf_counter()
    pb = adj.get_bar()
    t3 = time.perf_counter()
    dN = pb[0]

    def J(N, gamma=1.4):
        s2 = Step_RK2_C_ASC(chart, FluidParameters(CourantFactor=0.8, AdiabaticIndex=gamma))
        s2.set_primitive(N, *prim[1:])
        a2 = DiscreteAdjoint(s2)
        a2.forward(steps, every)
        val = a2.kinetic_energy(component=1, lo=0.2, hi=0.8)[0]
        s2.close()
        return val

    g = chart.g
    yc = np.asarray(chart.center[1][g:g + n])
    inner = (np.abs(yc - 0.5) < 0.25).reshape(1, -1, 1) * np.ones(chart.n[::-1])
    rng = np.random.default_rng(seed)
    h = 1e-6
    rel = []
    for _ in range(n_dirs):  # n_dirs = 0: cost only (for kernel traces)
        a, b = rng.standard_normal(2)
        v = a * inner + b * (1.0 - inner)
        fd = (J(prim[0] + h * v) - J(prim[0] - h * v)) / (2 * h)
        ad = float(np.sum(dN * v))
        rel.append(abs(fd - ad) / abs(ad))
    fdg = (J(prim[0], 1.4 + h) - J(prim[0], 1.4 - h)) / (2 * h) if n_dirs else None
    st.close()
    out = {
        "workload": f"Kelvin-Helmholtz {n}^3 periodic, seed 1234, {steps} RK2 steps FP64, "
                    f"checkpoints every {every}; J = vy kinetic energy in 0.2 < y < 0.8",
        "loss": loss,
        "gamma_gradient": gb,
        "forward_ms": t_fwd * 1e3,
        "forward_with_checkpoints_ms": (t1 - t0) * 1e3,
        "backward_ms": (t2 - t1) * 1e3,
        "gradient_to_host_ms": (t3 - t2) * 1e3,
        "adjoint_over_forward": (t2 - t1) / t_fwd,
        "gradient_over_forward": ((t1 - t0) + (t2 - t1)) / t_fwd,
        "fd_step": h,
        "fd_directions": n_dirs,
        "fd_direction_kind": "a*inner + b*(1-inner), a, b standard normal, inner = the "
                             "indicator of the dense layer |y-0.5| < 0.25: the step is not "
                             "differentiable at the limiter's ties inside the uniform layers, "
                             "so a direction that varies cell by cell there has no derivative; "
                             "the tests check arbitrary directions against torch autograd of "
                             "the step's FP64 transcription at 32^3",
        "fd_max_rel_err": max(rel) if rel else None,
        "fd_gamma_rel_err": abs(fdg - gb) / abs(gb) if fdg is not None else None,
    }
    print(f"# adjoint: backward {out['backward_ms']:.1f} ms, forward {out['forward_ms']:.1f} ms",
          file=sys.stderr, flush=True)
    return out


def copy_calibration():
    """A plain 1 GiB device copy (torch), for scaling the PMC byte counters."""
    import torch
    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": a.numel() * 8,
                      "copy_calibration_bytes_written": b.numel() * 8}))


def _tally_summary(t0, t1):
    """Tally_F_P totals after the run and the relative change of the
    conserved ones (baryon number, momentum, fluid energy): interior plus
    boundary fluence, so round-off only."""
    from genasis_amd._capi import TALLY_NAMES
    conserved = ("BaryonNumber", "Momentum_1", "Momentum_2", "Momentum_3", "FluidEnergy")
    scale = max(abs(t0["total"][TALLY_NAMES.index("FluidEnergy")]), 1e-300)
    return {"total": dict(zip(TALLY_NAMES, map(float, t1["total"]))),
            "boundary": dict(zip(TALLY_NAMES, map(float, t1["boundary"]))),
            "max_conserved_change_over_fluid_energy":
                max(abs(float(t1["change"][TALLY_NAMES.index(k)])) for k in conserved) / scale}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=98)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--tile", type=int, choices=(16, 32), default=16,
                    help="stage tile of the main run: 16 (16 x 16) or 32 (32 x 8)")
    ap.add_argument("--z-chunk", type=int, default=0,
                    help="stage z-chunk; 0 keeps the library's default for the chart")
    ap.add_argument("--cpu-baseline", choices=("auto", "off"), default="auto")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--precision", choices=("FP64", "FP32"), default="FP64")
    ap.add_argument("--fp32-check", choices=("auto", "off"), default="auto")
    ap.add_argument("--fp32-n", type=int, default=512)
    ap.add_argument("--fp32-steps", type=int, default=200)
    ap.add_argument("--batch-check", choices=("auto", "off"), default="auto")
    ap.add_argument("--batch-problems", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--batch-sequential", choices=("on", "off"), default="on")
    ap.add_argument("--adjoint-check", choices=("auto", "off"), default="auto")
    ap.add_argument("--parabolic-check", choices=("auto", "off"), default="auto")
    ap.add_argument("--parabolic-steps", type=int, default=30)
    ap.add_argument("--bricks-check", choices=("auto", "off"), default="auto")
    ap.add_argument("--bricks-steps", type=int, default=20)
    ap.add_argument("--adjoint-n", type=int, default=128)
    ap.add_argument("--adjoint-steps", type=int, default=50)
    ap.add_argument("--adjoint-fd-directions", type=int, default=8)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (multi-GPU sharding is a design note only).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    from genasis_amd import problems
    from genasis_amd.chart import Chart
    from genasis_amd.fluid import FluidParameters, Step_RK2_C_ASC
    from genasis_amd.integrator import Integrator_C_PS, IntegratorParameters

    n = args.n
    chart = Chart((n, n, n))
    params = FluidParameters(CourantFactor=0.8, Precision=args.precision)
    esz = 8 if args.precision == "FP64" else 4
    stepper = Step_RK2_C_ASC(chart, params)
    stepper.set_tile(args.tile, 256 // args.tile)
    if args.z_chunk > 0:
        stepper.set_tuning(args.z_chunk)
    z_chunk = stepper.get_tuning()
    geometry_pow2 = stepper.geometry_pow2()
    stepper.set_primitive(*problems.kelvin_helmholtz(chart, seed=1234, padded=False))
    total0 = stepper.totals()
    tally0 = stepper.compute_tally(False)
    integ = Integrator_C_PS(stepper, IntegratorParameters(
        FinishTime=1.0e30, FinishCycle=args.warmup + args.steps, nWrite=1))
    integ._administer_checkpoint()

    for _ in range(args.warmup):
        integ.compute_cycle()
    stepper.sync()
    stepper.set_profiling(True)
    stepper.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        integ.compute_cycle()
    stepper.sync()
    t1 = time.perf_counter()
    kt = stepper.kernel_times()
    stepper.set_profiling(False)
    total1 = stepper.totals()
    tally1 = stepper.compute_tally(True)

    elapsed = t1 - t0
    value = chart.n_proper * args.steps / elapsed
    s2_ms, s2_n = kt["stage2"]
    s2_avg = s2_ms / max(s2_n, 1) * 1e-3
    alg_stage2 = ALG_BYTES_STAGE2 * esz / 8
    alg_step = ALG_BYTES_STEP * esz / 8
    achieved = alg_stage2 * chart.n_proper / s2_avg / 1e9
    step_kernel_s = sum(v[0] for v in kt.values()) / max(s2_n, 1) * 1e-3
    traffic = _read_traffic() if args.precision == "FP64" and n == 256 else None
    at256 = args.precision == "FP64" and n == 256
    flop = _read_sq("stage2_fp64_flop_per_launch") if at256 else None
    insts = _read_sq("stage2_valu_wave_insts_per_launch") if at256 else None
    valu = None
    if flop is not None:
        tf = flop / s2_avg / 1e12
        valu = {"bound": "fp64 vector", "achieved": tf, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS,
                "flop_per_launch": flop,
                "source": "FP64 FLOP per stage-2 launch from the committed SQ counter "
                          "summary, over this run's average stage-2 launch time"}
    issue = None
    if insts is not None:
        # a wave64 VALU instruction holds a 16-lane SIMD four cycles
        frac = insts * 4 / (SIMDS * CLOCK_HZ * s2_avg)
        issue = {"bound": "VALU issue", "achieved": frac, "peak": 1.0,
                 "unit": "fraction of the SIMDs' issue cycles", "frac": frac,
                 "wave_insts_per_launch": insts,
                 "source": "stage-2 wave VALU instructions per launch from the committed "
                           "SQ counter summary, x 4 cycles, over 1024 SIMDs x 2.4 GHz x "
                           "this run's average stage-2 launch time"}
    stepper.close()
    print(f"# main run: {elapsed / args.steps * 1e3:.3f} ms/step", file=sys.stderr, flush=True)

    fp32 = None
    if args.fp32_check == "auto":
        fp32 = precision_check(args.fp32_n, args.fp32_steps, args.z_chunk)
    parabolic = None
    if args.parabolic_check == "auto":
        parabolic = parabolic_check(args.n, args.parabolic_steps, args.z_chunk)
    bricks = None
    if args.bricks_check == "auto":
        bricks = bricks_check(args.n, args.bricks_steps, [(2, 2, 1), (2, 2, 2), (4, 4, 2)])
    adjoint = None
    if args.adjoint_check == "auto":
        adjoint = adjoint_check(args.adjoint_n, args.adjoint_steps, 10,
                                args.adjoint_fd_directions, 2024)
    batch = None
    if args.batch_check == "auto":
        batch = batch_check(args.batch_problems, args.batch_steps, 2024, args.z_chunk,
                            args.batch_sequential == "on")

    cpu = None
    if args.cpu_baseline == "auto":
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) \
            or os.cpu_count() or 1
        # read before the oracle's OpenMP runtime binds this thread to one core
        host = host_cpu()
        t0 = time.perf_counter()
        kh_value, _, sod = cpu_baseline(args.cpu_n, args.cpu_steps, threads)
        cpu = {"value": kh_value, "unit": "cell-updates/s", "cores": threads,
               "kind": "port",
               "sample": (f"oracle (C/OpenMP restatement of the reference step, gcc -O2 "
                          f"-ffp-contract=off) on the same Kelvin-Helmholtz problem at "
                          f"{args.cpu_n}^3, {args.cpu_steps} timed RK2 steps after 1 untimed "
                          f"step, {threads} OpenMP threads (OMP_PROC_BIND "
                          f"{os.environ.get('OMP_PROC_BIND')}, OMP_PLACES "
                          f"{os.environ.get('OMP_PLACES')})"),
               "omp_threads": threads,
               "host": host,
               "sod": sod,
               "seconds_total": time.perf_counter() - t0,
               "gpu_over_cpu": value / kh_value if args.cpu_n == n else None}
        print(f"# cpu baseline: {kh_value:.3e} cell-updates/s KH, {sod['value']:.3e} Sod",
              file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "FP64" else "f32",
        "data": "synthetic (seeded Kelvin-Helmholtz initial data)",
        "config": {"workload": f"Kelvin-Helmholtz {n}^3 periodic, PLM(MC theta=1.4)"
                               f"+HLLC, RK2, CFL 0.8, {args.precision}",
                   "cells": chart.n_proper, "z_chunk": z_chunk,
                   # every geometry divisor a power of two (unit cube, 2^k
                   # cells): divisions by it are exact reciprocal products
                   "geometry_pow2": geometry_pow2,
                   "tile": [args.tile, 256 // args.tile]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic},
        "valu_roofline": valu,
        "valu_issue": issue,
        "hbm_measured": None if traffic is None else {
            "bytes_per_launch": traffic,
            "achieved_gbs": traffic / s2_avg / 1e9,
            "frac": traffic / s2_avg / 1e9 / HBM_PEAK_GBS,
            "amplification": traffic / (alg_stage2 * chart.n_proper),
            "source": "stage-2 HBM bytes per launch from the committed FETCH_SIZE (x2, "
                      "copy-calibrated) and WRITE_SIZE summary, over this run's average "
                      "stage-2 launch time"},
        "step_roofline": {"alg_bytes_per_cell_update": alg_step,
                          "kernel_s_per_step": step_kernel_s,
                          "frac": alg_step * chart.n_proper / step_kernel_s
                          / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_avg": {k: v[0] / max(v[1], 1) for k, v in kt.items()},
        "conserved_totals_change": [float(a - b) for a, b in zip(total1, total0)],
        "tally": _tally_summary(tally0, tally1),
        "cpu_baseline": cpu,
        "fp32": fp32,
        "batch": batch,
        "adjoint": adjoint,
        "parabolic": parabolic,
        "bricks": bricks,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
