"""Many independent Fluid_P_I charts advanced together (gx_batch_* of
include/gx_fluid.h).

Each problem behaves as its own ``Step_RK2_C_ASC`` driven by its own
``Integrator_C_PS`` with ``compute_cycle`` (Integrator_C_PS__Form.f90:328-371),
FinishTime/FinishCycle/nRampCycles shared; the batch computes every time step
on the device, so ``step(n)`` is one HIP graph of n cycles.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _capi
from .chart import Chart
from .fluid import FluidParameters, _pd


class Batch_RK2_C_ASC:
    def __init__(self, charts: Sequence[Chart], params: FluidParameters | None = None):
        self.lib = _capi.load()
        self.charts = list(charts)
        self.params = params or FluidParameters()
        grids = (_capi.GxGrid * len(self.charts))()
        for k, chart in enumerate(self.charts):
            for d in range(3):
                grids[k].nCells[d] = chart.n[d]
                grids[k].MinCoordinate[d] = chart.lo[d]
                grids[k].MaxCoordinate[d] = chart.hi[d]
            for i, code in enumerate(chart.bc_codes):
                grids[k].BoundaryCondition[i] = code
        h = ctypes.c_void_p()
        cp = self.params.to_c()
        _capi.check(self.lib.gx_batch_create(len(self.charts), grids, ctypes.byref(cp),
                                             ctypes.byref(h)), "gx_batch_create")
        self._h = h

    @property
    def n_problems(self) -> int:
        return len(self.charts)

    @property
    def cells(self) -> int:
        return int(self.lib.gx_batch_cells(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.gx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _proper(self, k: int, a) -> np.ndarray:
        chart = self.charts[k]
        a = np.asarray(a, dtype=np.float64)
        if a.shape == chart.m[::-1]:
            a = chart.proper(a)
        if a.shape != chart.n[::-1]:
            raise ValueError(f"field shape {a.shape} does not match chart {k}")
        return a

    def set_primitive(self, k: int, N, V1, V2, V3, E):
        buf = np.ascontiguousarray(np.stack([self._proper(k, a) for a in (N, V1, V2, V3, E)]))
        _capi.check(self.lib.gx_batch_set_primitive(self._h, int(k), _pd(buf)),
                    "gx_batch_set_primitive")

    def set_integrator(self, FinishTime: float = 1.0, FinishCycle: int = 2**31 - 1,
                       nRampCycles: int = 100):
        _capi.check(self.lib.gx_batch_set_integrator(self._h, float(FinishTime),
                                                     int(FinishCycle), int(nRampCycles)),
                    "gx_batch_set_integrator")

    def step(self, n_steps: int):
        _capi.check(self.lib.gx_batch_step(self._h, int(n_steps)), "gx_batch_step")

    def sync(self):
        _capi.check(self.lib.gx_batch_sync(self._h), "gx_batch_sync")

    def time_step_candidates(self) -> np.ndarray:
        out = np.empty(self.n_problems)
        _capi.check(self.lib.gx_batch_cfl(self._h, _pd(out)), "gx_batch_cfl")
        return out

    def times(self):
        """(Time, last time step, cycle) of every problem."""
        t = np.empty(self.n_problems)
        dt = np.empty(self.n_problems)
        cyc = np.empty(self.n_problems, dtype=np.int32)
        _capi.check(self.lib.gx_batch_times(self._h, _pd(t), _pd(dt),
                                            cyc.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
                    "gx_batch_times")
        return t, dt, cyc

    def conserved(self, k: int) -> np.ndarray:
        out = np.empty((5,) + self.charts[k].n[::-1])
        _capi.check(self.lib.gx_batch_get_conserved(self._h, int(k), _pd(out)),
                    "gx_batch_get_conserved")
        return out

    def flags(self, k: int) -> np.ndarray:
        out = np.empty(self.charts[k].n[::-1], dtype=np.uint8)
        _capi.check(self.lib.gx_batch_get_flags(
            self._h, int(k), out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))),
            "gx_batch_get_flags")
        return out

    def set_tuning(self, z_chunks: dict | None = None, use_graph: bool = True):
        """z_chunks: {size class: z-chunk}."""
        z_chunks = z_chunks or {}
        cls = np.array(list(z_chunks.keys()), dtype=np.int32)
        tz = np.array(list(z_chunks.values()), dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        _capi.check(self.lib.gx_batch_set_tuning(self._h, len(cls), cls.ctypes.data_as(ip),
                                                 tz.ctypes.data_as(ip), int(bool(use_graph))),
                    "gx_batch_set_tuning")

    def tuning(self) -> dict:
        n = ctypes.c_int()
        _capi.check(self.lib.gx_batch_get_tuning(self._h, 0, None, None, ctypes.byref(n)),
                    "gx_batch_get_tuning")
        cls = np.empty(n.value, dtype=np.int32)
        tz = np.empty(n.value, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        _capi.check(self.lib.gx_batch_get_tuning(self._h, n.value, cls.ctypes.data_as(ip),
                                                 tz.ctypes.data_as(ip), ctypes.byref(n)),
                    "gx_batch_get_tuning")
        return {int(c): int(z) for c, z in zip(cls, tz)}

    def set_tiles(self, tiles: dict, use_graph: bool = True):
        """tiles: {size class: (tileX, tileY, z-chunk)}, the stage-tile shape
        16 x 16 or 32 x 8."""
        ip = ctypes.POINTER(ctypes.c_int)
        cls = np.array(list(tiles.keys()), dtype=np.int32)
        v = np.array(list(tiles.values()), dtype=np.int32).reshape(-1, 3)
        tx, ty, tz = (np.ascontiguousarray(v[:, i]) for i in range(3))
        _capi.check(self.lib.gx_batch_set_tiles(self._h, len(cls), cls.ctypes.data_as(ip),
                                                tx.ctypes.data_as(ip), ty.ctypes.data_as(ip),
                                                tz.ctypes.data_as(ip), int(bool(use_graph))),
                    "gx_batch_set_tiles")

    def tiles(self) -> dict:
        """{size class: (tileX, tileY, z-chunk)}."""
        n = ctypes.c_int()
        ip = ctypes.POINTER(ctypes.c_int)
        _capi.check(self.lib.gx_batch_get_tiles(self._h, 0, None, None, None, None,
                                                ctypes.byref(n)), "gx_batch_get_tiles")
        a = [np.empty(n.value, dtype=np.int32) for _ in range(4)]
        _capi.check(self.lib.gx_batch_get_tiles(self._h, n.value, *[x.ctypes.data_as(ip) for x in a],
                                                ctypes.byref(n)), "gx_batch_get_tiles")
        return {int(c): (int(x), int(y), int(z)) for c, x, y, z in zip(*a)}

    def autotune(self, n_repeat: int = 2) -> dict:
        """Tile shape and z-chunk of every size class by timing; returns the
        z-chunks ({size class: z-chunk}; tiles() has the shapes)."""
        _capi.check(self.lib.gx_batch_autotune(self._h, int(n_repeat)), "gx_batch_autotune")
        return self.tuning()

    @property
    def stream(self) -> int:
        return self.lib.gx_batch_stream(self._h) or 0
