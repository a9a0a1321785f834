"""Initial data for the Fluid_P_I step: primitive fields on the padded layout.

Every function returns ``(N, V1, V2, V3, E)`` as float64 arrays of shape
``(m2, m1, m0)`` (x fastest), ghost cells included, as the reference's
``SetProblem`` fills every value before ``PrepareEvolution`` exchanges
periodic ghosts (Programs/Examples/Physics/FluidDynamics/RiemannProblem_Form.
f90:237-275).  E is the internal-energy density P/(gamma-1)
(RiemannProblem_Form.f90 SetProblem: ``Energy = Pressure/(AdiabaticIndex-1)``).
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .chart import Chart

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on uint64 counters."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _MASK64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK64
        return z ^ (z >> np.uint64(31))


def uniform_pm1(seed: int, index: np.ndarray) -> np.ndarray:
    """U(-1, 1) per linear index: SplitMix64 of (seed xor index), top 53 bits."""
    key = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ index.astype(np.uint64)
    r = splitmix64(key) >> np.uint64(11)
    u = r.astype(np.float64) * (1.0 / 9007199254740992.0)  # [0, 1)
    return 2.0 * u - 1.0


def sod(chart: Chart, gamma: float = 1.4, axis: int = 0, x0: float = 0.5,
        left=(1.0, 1.0, 0.0), right=(0.125, 0.1, 0.0), padded: bool = True):
    """Sod shock tube along ``axis``: (density, pressure, speed) left/right of
    the plane ``x_axis = x0``; the left state holds where
    ``x_axis - x0 <= 1e-10`` (RiemannProblem_Form.f90 SetFluidKernel)."""
    g = chart.g
    centers = np.asarray(chart.center[axis])
    if padded:
        shape = chart.m[::-1]
    else:
        shape = chart.n[::-1]
        centers = centers[g:g + chart.n[axis]]
    is_left = (centers - x0) <= 1.0e-10
    # centres increase along the axis, so the left cells are a leading run
    n_left = int(is_left.sum())
    assert is_left[:n_left].all()
    rho_l, p_l, v_l = left
    rho_r, p_r, v_r = right
    e_l = p_l / (gamma - 1.0)
    e_r = p_r / (gamma - 1.0)

    def fill(a, b):
        out = np.full(shape, float(b))
        sl = [slice(None)] * 3
        sl[2 - axis] = slice(0, n_left)
        out[tuple(sl)] = a
        return out

    N = fill(rho_l, rho_r)
    E = fill(e_l, e_r)
    V = [np.zeros(shape) for _ in range(3)]
    V[axis] = fill(v_l, v_r)
    return N, V[0], V[1], V[2], E


def kelvin_helmholtz(chart: Chart, seed: int = 1234, gamma: float = 1.4,
                     padded: bool = True):
    """Kelvin-Helmholtz shear layer on the unit cube.

    rho = 2, vx = +0.5 where |y - 0.5| < 0.25, else rho = 1, vx = -0.5;
    p = 2.5 everywhere; vy = 0.01 * U(-1, 1) per cell from SplitMix64 keyed by
    seed xor (linear proper-cell index); vz = 0.  Ghost cells take the value
    of their periodic image through the exchange that follows.  With
    ``padded=False`` the arrays hold the proper cells only, (n2, n1, n0).
    """
    g = chart.g
    n0, n1, n2 = chart.n
    yc = np.asarray(chart.center[1])[g:g + n1]
    inner = (np.abs(yc - 0.5) < 0.25)[None, :, None]
    shape = (n2, n1, n0)
    N = np.broadcast_to(np.where(inner, 2.0, 1.0), shape)
    V1 = np.broadcast_to(np.where(inner, 0.5, -0.5), shape)
    E = np.full(shape, 2.5 / (gamma - 1.0))
    V2 = np.empty(shape)
    plane = n0 * n1
    step = max(1, (1 << 20) // plane)

    def fill(k0):
        k1 = min(n2, k0 + step)
        lin = np.arange(k0 * plane, k1 * plane, dtype=np.uint64)
        V2[k0:k1] = (0.01 * uniform_pm1(seed, lin)).reshape(k1 - k0, n1, n0)

    # numpy releases the GIL in its element-wise loops: slabs in threads
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
        list(pool.map(fill, range(0, n2, step)))
    V3 = np.zeros(shape)
    fields = (N, V1, V2, V3, E)
    if padded:
        return tuple(np.pad(f, g, mode="wrap") for f in fields)
    return tuple(np.ascontiguousarray(f) for f in fields)


def riemann_ensemble_member(chart: Chart, rng: np.random.Generator,
                            gamma: float = 1.4, padded: bool = True):
    """A random 3-D Riemann problem: rho and p log-uniform in [0.1, 10],
    normal velocity uniform in [-1, 1], discontinuity plane through the
    centre with its normal along x, y or z."""
    axis = int(rng.integers(0, 3))
    rho = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
    p = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
    v = rng.uniform(-1.0, 1.0, 2)
    x0 = 0.5 * (chart.lo[axis] + chart.hi[axis])
    return sod(chart, gamma, axis=axis, x0=x0, left=(rho[0], p[0], v[0]),
               right=(rho[1], p[1], v[1]), padded=padded), axis


def riemann_ensemble_sizes(n_problems: int, seed: int, n_min: int = 32,
                           n_max: int = 128) -> list:
    """Edge sizes of a Riemann-problem ensemble: round(exp(U(ln n_min,
    ln n_max))) per problem from a seeded generator (cubic charts)."""
    rng = np.random.default_rng([seed, 0])
    u = rng.uniform(np.log(n_min), np.log(n_max), n_problems)
    return [int(round(float(np.exp(x)))) for x in u]


def riemann_ensemble_problem(k: int, n: int, seed: int, padded: bool = False):
    """Problem k of the ensemble: an n^3 OUTFLOW unit cube with a random
    Riemann problem (riemann_ensemble_member) from the generator seeded by
    (seed, k + 1).  Returns (chart, primitives, axis)."""
    chart = Chart((n, n, n), bc=(("OUTFLOW", "OUTFLOW"),) * 3)
    rng = np.random.default_rng([seed, k + 1])
    prim, axis = riemann_ensemble_member(chart, rng, padded=padded)
    return chart, prim, axis
