"""Host mirror of the reference's fluid-step interface on the HIP path.

``FluidParameters`` carries the reference's parameter names and defaults
(Modules/Physics/StressEnergies/Fluids/Fluid_ASC__Form.f90:128-177,
Fluid_P_I__Form.f90:206-223, FluidFeatures_ASC__Form.f90:61-65,
Evolutions/Integrators/Integrator_C_PS__Form.f90:136-140).

``Step_RK2_C_ASC`` owns one device context (gx_create) and exposes the two
calls the integrator makes across the boundary -- ``time_step_candidate``
(Integrator_C_PS % ComputeTimeStep_C_ASC) and ``step``
(Step_RK_C_ASC_Template % Compute) -- plus state upload and download.
All arithmetic runs in the HIP library; this module only moves buffers.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi
from .chart import Chart


@dataclass
class FluidParameters:
    AdiabaticIndex: float = 1.4
    LimiterParameter: float = 1.4
    ShockThreshold: float = 0.1
    CourantFactor: float = 0.7
    BaryonDensityMin: float = 0.0
    BaryonMassReference: float = 1.0
    UseLimiter: bool = True
    RiemannSolverType: str = "HLLC"
    Precision: str = "FP64"
    ReconstructionType: str = "LINEAR"

    def to_c(self) -> _capi.GxParams:
        if self.RiemannSolverType not in ("HLL", "HLLC"):
            raise ValueError(f"RiemannSolverType not recognized: {self.RiemannSolverType}")
        if self.Precision not in ("FP64", "FP32"):
            raise ValueError(f"Precision not recognized: {self.Precision}")
        if self.ReconstructionType not in ("LINEAR", "PARABOLIC"):
            raise ValueError(f"ReconstructionType not recognized: {self.ReconstructionType}")
        p = _capi.GxParams()
        p.AdiabaticIndex = self.AdiabaticIndex
        p.LimiterParameter = self.LimiterParameter
        p.ShockThreshold = self.ShockThreshold
        p.CourantFactor = self.CourantFactor
        p.BaryonDensityMin = self.BaryonDensityMin
        p.BaryonMassReference = self.BaryonMassReference
        p.UseLimiter = int(bool(self.UseLimiter))
        p.RiemannSolverType = _capi.GX_HLLC if self.RiemannSolverType == "HLLC" else _capi.GX_HLL
        p.Precision = _capi.GX_FP64 if self.Precision == "FP64" else _capi.GX_FP32
        p.ReconstructionType = (_capi.GX_PARABOLIC if self.ReconstructionType == "PARABOLIC"
                                else _capi.GX_LINEAR)
        return p


def _pd(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Step_RK2_C_ASC:
    """Device-resident RK2 step of one Fluid_P_I chart."""

    def __init__(self, chart: Chart, params: FluidParameters | None = None,
                 n_bricks=None):
        """n_bricks: (nx, ny, nz) splits the chart into bricks in the GPU's
        memory (Chart_SLD's nBricks; gx_create_bricks); None is one brick."""
        self.lib = _capi.load()
        self.chart = chart
        self.params = params or FluidParameters()
        g = _capi.GxGrid()
        for d in range(3):
            g.nCells[d] = chart.n[d]
            g.MinCoordinate[d] = chart.lo[d]
            g.MaxCoordinate[d] = chart.hi[d]
        for i, code in enumerate(chart.bc_codes):
            g.BoundaryCondition[i] = code
        h = ctypes.c_void_p()
        cp = self.params.to_c()
        if n_bricks is None:
            _capi.check(self.lib.gx_create(ctypes.byref(g), ctypes.byref(cp), ctypes.byref(h)),
                        "gx_create")
        else:
            nb = (ctypes.c_int * 3)(*[int(v) for v in n_bricks])
            _capi.check(self.lib.gx_create_bricks(ctypes.byref(g), ctypes.byref(cp), nb,
                                                  ctypes.byref(h)), "gx_create_bricks")
        self._h = h
        self.shape = chart.n[::-1]  # (n2, n1, n0), x fastest

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.gx_destroy(self._h)
            self._h = None

    @property
    def n_bricks(self):
        nb = (ctypes.c_int * 3)()
        _capi.check(self.lib.gx_get_bricks(self._h, nb), "gx_get_bricks")
        return tuple(nb)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- state -------------------------------------------------------------
    def _proper(self, a: np.ndarray) -> np.ndarray:
        a = np.asarray(a, dtype=np.float64)
        if a.shape == self.chart.m[::-1]:
            a = self.chart.proper(a)
        if a.shape != self.shape:
            raise ValueError(f"field shape {a.shape} does not match the chart {self.shape}")
        return a

    def set_primitive(self, N, V1, V2, V3, E):
        """Primitives on proper or padded arrays; conserved fields follow on
        the device (ComputeFromInitial)."""
        buf = np.ascontiguousarray(np.stack([self._proper(a) for a in (N, V1, V2, V3, E)]))
        _capi.check(self.lib.gx_set_primitive(self._h, _pd(buf)), "gx_set_primitive")

    def set_state(self, primitive, conserved):
        p = np.ascontiguousarray(np.stack([self._proper(a) for a in primitive]))
        c = np.ascontiguousarray(np.stack([self._proper(a) for a in conserved]))
        _capi.check(self.lib.gx_set_state(self._h, _pd(p), _pd(c)), "gx_set_state")

    def conserved(self) -> np.ndarray:
        out = np.empty((5,) + self.shape)
        _capi.check(self.lib.gx_get_conserved(self._h, _pd(out)), "gx_get_conserved")
        return out

    def primitive(self) -> np.ndarray:
        out = np.empty((5,) + self.shape)
        _capi.check(self.lib.gx_get_primitive(self._h, _pd(out)), "gx_get_primitive")
        return out

    def flags(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.uint8)
        _capi.check(self.lib.gx_get_flags(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))), "gx_get_flags")
        return out

    def set_flags(self, flags: np.ndarray):
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        _capi.check(self.lib.gx_set_flags(
            self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))), "gx_set_flags")

    def totals(self) -> np.ndarray:
        out = np.empty(5)
        _capi.check(self.lib.gx_totals(self._h, _pd(out)), "gx_totals")
        return out

    def boundary_fluence(self):
        """Fluence of the last step through the six faces: a list over
        b = 2*d (inner) and 2*d + 1 (outer) of (5,) + slab arrays (D, S_1,
        S_2, S_3, G; the slab is the chart's numpy shape without dimension d)."""
        n = self.chart.n
        slabs = [n[1] * n[2], n[0] * n[2], n[0] * n[1]]
        out = np.empty(10 * sum(slabs))
        _capi.check(self.lib.gx_get_boundary_fluence(self._h, _pd(out)),
                    "gx_get_boundary_fluence")
        faces, o = [], 0
        for b in range(6):
            d = b // 2
            shape = tuple(n[e] for e in (2, 1, 0) if e != d)
            faces.append(out[o:o + 5 * slabs[d]].reshape((5,) + shape))
            o += 5 * slabs[d]
        return faces

    def write_checkpoint(self, path: str, time: float, cycle: int, i_checkpoint: int = 0):
        """Checkpoint file of the context with the integrator's Time, cycle and
        iCheckpoint (gx_write_checkpoint)."""
        _capi.check(self.lib.gx_write_checkpoint(self._h, str(path).encode(), float(time),
                                                 int(cycle), int(i_checkpoint)),
                    "gx_write_checkpoint")

    def read_checkpoint(self, path: str, exact: bool = False):
        """Restore from a checkpoint file: exact=False repeats the reference's
        RestartFrom (primitives, conserved recomputed), exact=True restores the
        whole device state.  Returns (Time, cycle, iCheckpoint)."""
        t, cyc, ick = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_int()
        mode = _capi.GX_RESTART_EXACT if exact else _capi.GX_RESTART_PRIMITIVE
        _capi.check(self.lib.gx_read_checkpoint(self._h, str(path).encode(), mode,
                                                ctypes.byref(t), ctypes.byref(cyc),
                                                ctypes.byref(ick)), "gx_read_checkpoint")
        return t.value, cyc.value, ick.value

    def compute_tally(self, compute_change: bool = True):
        """Current_ASC % ComputeTally: interior, boundary, total and change
        tallies, GX_TALLY_N values each (Tally_F_P % Value order)."""
        outs = [np.zeros(_capi.GX_TALLY_N) for _ in range(4)]
        _capi.check(self.lib.gx_compute_tally(self._h, int(bool(compute_change)),
                                              *[_pd(o) for o in outs]), "gx_compute_tally")
        return dict(zip(("interior", "boundary", "total", "change"), outs))

    def upload(self, value: np.ndarray, field_map=None):
        """Reference layout Value(nValues, nFields) as a (nFields, nValues)
        C-contiguous array (column f is value[f])."""
        v = np.ascontiguousarray(value, dtype=np.float64)
        fm = np.arange(1, _capi.GX_N_FIELDS + 1, dtype=np.int32) if field_map is None \
            else np.ascontiguousarray(field_map, dtype=np.int32)
        _capi.check(self.lib.gx_upload(self._h, _pd(v), v.shape[1],
                                       fm.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
                    "gx_upload")

    def download(self, value: np.ndarray, field_map=None):
        fm = np.arange(1, _capi.GX_N_FIELDS + 1, dtype=np.int32) if field_map is None \
            else np.ascontiguousarray(field_map, dtype=np.int32)
        assert value.flags.c_contiguous and value.dtype == np.float64
        _capi.check(self.lib.gx_download(self._h, _pd(value), value.shape[1],
                                         fm.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
                    "gx_download")

    def upload_features(self, df: np.ndarray):
        """DiffusiveFlux_I_1..3 in the reference layout: a (3, ldValue)
        C-contiguous array (column d is df[d])."""
        v = np.ascontiguousarray(df, dtype=np.float64)
        _capi.check(self.lib.gx_upload_features(self._h, _pd(v), v.shape[1]),
                    "gx_upload_features")

    def download_features(self, df: np.ndarray):
        assert df.flags.c_contiguous and df.dtype == np.float64 and df.shape[0] == 3
        _capi.check(self.lib.gx_download_features(self._h, _pd(df), df.shape[1]),
                    "gx_download_features")

    # --- the two boundary calls of the integrator --------------------------
    def time_step_candidate(self) -> float:
        c = ctypes.c_double()
        _capi.check(self.lib.gx_compute_time_step(self._h, ctypes.byref(c)),
                    "gx_compute_time_step")
        return c.value

    def step(self, time: float, dt: float):
        _capi.check(self.lib.gx_step_rk2(self._h, float(time), float(dt)), "gx_step_rk2")

    def sync(self):
        _capi.check(self.lib.gx_sync(self._h), "gx_sync")

    def set_tuning(self, z_chunk: int):
        _capi.check(self.lib.gx_set_tuning(self._h, int(z_chunk)), "gx_set_tuning")

    def set_tile(self, tile_x: int, tile_y: int):
        """Stage-tile shape: 16 x 16 (default) or 32 x 8."""
        _capi.check(self.lib.gx_set_tile(self._h, int(tile_x), int(tile_y)), "gx_set_tile")

    def get_tile(self):
        x, y = ctypes.c_int(), ctypes.c_int()
        _capi.check(self.lib.gx_get_tile(self._h, ctypes.byref(x), ctypes.byref(y)),
                    "gx_get_tile")
        return x.value, y.value

    def geometry_pow2(self) -> bool:
        """True when every divisor of the chart's geometry is a power of two
        (the FP64 stage kernels then multiply by exact reciprocals)."""
        v = ctypes.c_int()
        _capi.check(self.lib.gx_get_geometry_pow2(self._h, ctypes.byref(v)),
                    "gx_get_geometry_pow2")
        return bool(v.value)

    def record_sources(self, on: bool = True):
        """Record Current % Sources of the divergence in every step
        (RecordDivergence); turning it on clears them."""
        _capi.check(self.lib.gx_set_record_sources(self._h, int(bool(on))),
                    "gx_set_record_sources")

    def sources(self) -> np.ndarray:
        """Sources of the divergence of the last step: D, S_1..3, G."""
        out = np.empty((5,) + self.shape)
        _capi.check(self.lib.gx_get_sources(self._h, _pd(out)), "gx_get_sources")
        return out

    def get_tuning(self) -> int:
        """The stage kernels' z-chunk (planes per block)."""
        v = ctypes.c_int()
        _capi.check(self.lib.gx_get_tuning(self._h, ctypes.byref(v)), "gx_get_tuning")
        return v.value

    def set_profiling(self, on: bool):
        _capi.check(self.lib.gx_set_profiling(self._h, int(bool(on))), "gx_set_profiling")

    def kernel_times(self):
        """{name: (total_ms, steps)} for stage 1, stage 2, the shock flags and
        the boundary fluence with the boundary tally."""
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_longlong * 4)()
        _capi.check(self.lib.gx_kernel_times(self._h, ms, n), "gx_kernel_times")
        names = ("stage1", "stage2", "shock_flags", "boundary_fluence_tally")
        return {k: (ms[i], n[i]) for i, k in enumerate(names)}

    @property
    def stream(self) -> int:
        return self.lib.gx_stream(self._h) or 0
