"""ctypes view of the CPU restatement (libgx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker and the timed CPU baseline, never
by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgx_oracle.so")
_lib = None

FIELD = {
    "FEP_1": 0, "FEP_2": 1, "FEP_3": 2, "FEM_1": 3, "FEM_2": 4, "FEM_3": 5,
    "N": 6, "D": 7, "M": 8, "V_1": 9, "V_2": 10, "V_3": 11,
    "S_1": 12, "S_2": 13, "S_3": 14, "E": 15, "G": 16, "P": 17, "T": 18,
    "SB": 19, "DS": 20, "CS": 21, "MN": 22,
}
CONSERVED = ("D", "S_1", "S_2", "S_3", "G")
TALLY_N = 12  # Tally_F_P % Value: BaryonNumber, Momentum_1..3, KineticEnergy,
# AngularMomentum_1..3, GravitationalEnergy, TotalEnergy, FluidEnergy,
# InternalEnergy
PRIMITIVE = ("N", "V_1", "V_2", "V_3", "E")

_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)


def build() -> None:
    """Compile libgx_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.gxo_create.restype = ctypes.c_void_p
        lib.gxo_create.argtypes = [_PI, _PD, _PD, _PI, _D, _D, ctypes.c_int,
                                   ctypes.c_int, _D, _D, _D, ctypes.c_int]
        lib.gxo_destroy.argtypes = [ctypes.c_void_p]
        lib.gxo_set_primitive.argtypes = [ctypes.c_void_p] + [_PD] * 5
        lib.gxo_set_conserved.argtypes = [ctypes.c_void_p] + [_PD] * 5
        lib.gxo_time_step_candidate.restype = _D
        lib.gxo_time_step_candidate.argtypes = [ctypes.c_void_p, _PD]
        lib.gxo_step.argtypes = [ctypes.c_void_p, _D]
        lib.gxo_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _PD]
        lib.gxo_get_sources.argtypes = [ctypes.c_void_p, ctypes.c_int, _PD]
        lib.gxo_get_flags.argtypes = [ctypes.c_void_p, ctypes.c_int, _PD]
        lib.gxo_fallback_counts.argtypes = [ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_long), ctypes.c_int]
        lib.gxo_set_flags.argtypes = [ctypes.c_void_p, ctypes.c_int, _PD]
        lib.gxo_geometry.argtypes = [ctypes.c_void_p, ctypes.c_int, _PD, _PD, _PD]
        lib.gxo_limited_slope.restype = _D
        lib.gxo_limited_slope.argtypes = [_D, _D, _D, _D, _D]
        lib.gxo_face_flux.argtypes = [_PD, _PD, ctypes.c_int, _D, ctypes.c_int,
                                      _D, _PD]
        lib.gxo_boundary_fluence.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_int, _PD]
        lib.gxo_compute_tally.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_PD] * 4
        lib.gxo_set_boundary_fluence.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_int, _PD]
        lib.gxo_surface_integral.argtypes = [ctypes.c_void_p, _PD]
        lib.gxo_set_reconstruction.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.gxo_parabolic_face_values.argtypes = [_D] * 8 + [_PD]
        _lib = lib
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_PD)


def limited_slope(dv_l, dv_r, dx_l, dx_r, theta=1.4) -> float:
    return _load().gxo_limited_slope(dv_l, dv_r, dx_l, dx_r, theta)


def parabolic_face_values(vm, vc, vp, xm, xc, xp, dxl, dxr):
    """(vI, vO) of the parabolic reconstruction of one cell."""
    out = np.zeros(2)
    _load().gxo_parabolic_face_values(vm, vc, vp, xm, xc, xp, dxl, dxr, _ptr(out))
    return float(out[0]), float(out[1])


def face_flux(prim_l, prim_r, d, gamma=1.4, riemann="HLLC", df=0.0):
    out = np.zeros(5)
    pl = np.ascontiguousarray(prim_l, dtype=np.float64)
    pr = np.ascontiguousarray(prim_r, dtype=np.float64)
    _load().gxo_face_flux(_ptr(pl), _ptr(pr), d, gamma,
                          1 if riemann == "HLLC" else 0, df, _ptr(out))
    return out


class OracleFluid:
    """The reference step on the CPU, on the reference's padded layout."""

    def __init__(self, chart, gamma=1.4, theta=1.4, use_limiter=True,
                 riemann="HLLC", shock_threshold=0.1, courant=0.7,
                 n_min=0.0, n_threads=0, reconstruction="LINEAR"):
        lib = _load()
        self.chart = chart
        self.shape = chart.m[::-1]
        n = (ctypes.c_int * 3)(*chart.n)
        lo = (ctypes.c_double * 3)(*chart.lo)
        hi = (ctypes.c_double * 3)(*chart.hi)
        bc = (ctypes.c_int * 6)(*chart.bc_codes)
        self._h = lib.gxo_create(n, lo, hi, bc, gamma, theta, int(use_limiter),
                                 1 if riemann == "HLLC" else 0,
                                 shock_threshold, courant, n_min, n_threads)
        if not self._h:
            raise MemoryError("gxo_create failed")
        self._lib = lib
        if reconstruction not in ("LINEAR", "PARABOLIC"):
            raise ValueError(f"ReconstructionType not recognized: {reconstruction}")
        lib.gxo_set_reconstruction(self._h, int(reconstruction == "PARABOLIC"))

    def close(self):
        if self._h:
            self._lib.gxo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_primitive(self, N, V1, V2, V3, E):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (N, V1, V2, V3, E)]
        for a in arrs:
            assert a.shape == self.shape, (a.shape, self.shape)
        self._lib.gxo_set_primitive(self._h, *[_ptr(a) for a in arrs])

    def set_conserved(self, D, S1, S2, S3, G):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (D, S1, S2, S3, G)]
        self._lib.gxo_set_conserved(self._h, *[_ptr(a) for a in arrs])

    def time_step_candidate(self):
        inv = ctypes.c_double(0.0)
        ts = self._lib.gxo_time_step_candidate(self._h, ctypes.byref(inv))
        return ts, inv.value

    def step(self, dt: float):
        self._lib.gxo_step(self._h, float(dt))

    def field(self, name) -> np.ndarray:
        out = np.empty(self.shape)
        self._lib.gxo_get_field(self._h, FIELD[name], _ptr(out))
        return out

    def sources(self) -> np.ndarray:
        """Current % Sources of the divergence after the last step (padded,
        D, S_1..3, G, DS)."""
        out = np.empty((6,) + self.shape)
        for f in range(6):
            self._lib.gxo_get_sources(self._h, f, _ptr(out[f]))
        return out

    def fallback_counts(self, reset=False):
        """Proper cells ComputeFromConserved repaired since the last reset:
        {"stage": (vacuum, energy), "store": (vacuum, energy)} for the stage-2
        input Y and for StoreSolution; vacuum = D <= N_Min, energy = E < 0."""
        out = (ctypes.c_long * 4)()
        self._lib.gxo_fallback_counts(self._h, out, int(bool(reset)))
        return {"stage": (int(out[0]), int(out[1])), "store": (int(out[2]), int(out[3]))}

    def flags(self, d) -> np.ndarray:
        out = np.empty(self.shape)
        self._lib.gxo_get_flags(self._h, d, _ptr(out))
        return out

    def set_flags(self, d, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._lib.gxo_set_flags(self._h, d, _ptr(a))

    def geometry(self, d):
        m = self.chart.m[d]
        c, wl, wr = np.empty(m), np.empty(m), np.empty(m)
        self._lib.gxo_geometry(self._h, d, _ptr(c), _ptr(wl), _ptr(wr))
        return c, wl, wr

    def conserved(self):
        return [self.field(f) for f in CONSERVED]

    def primitive(self):
        return [self.field(f) for f in PRIMITIVE]

    def boundary_fluence(self, b) -> np.ndarray:
        """Fluence of the last step through face b = 2*d (inner) or 2*d + 1
        (outer): D, S_1, S_2, S_3, G over the transverse proper cells, shape
        (5,) + the chart's extents without dimension d (numpy order)."""
        d = b // 2
        shape = tuple(self.chart.n[e] for e in (2, 1, 0) if e != d)
        out = np.empty((5,) + shape)
        for f in range(5):
            self._lib.gxo_boundary_fluence(self._h, b, f, _ptr(out[f]))
        return out

    def compute_tally(self, compute_change=True):
        """Interior, boundary, total and change tallies (12 values each)."""
        outs = [np.zeros(TALLY_N) for _ in range(4)]
        self._lib.gxo_compute_tally(self._h, int(compute_change), *[_ptr(o) for o in outs])
        return dict(zip(("interior", "boundary", "total", "change"), outs))

    def set_boundary_fluence(self, b, fluence):
        for f in range(5):
            a = np.ascontiguousarray(fluence[f], dtype=np.float64)
            self._lib.gxo_set_boundary_fluence(self._h, b, f, _ptr(a))

    def surface_integral(self) -> np.ndarray:
        out = np.zeros(TALLY_N)
        self._lib.gxo_surface_integral(self._h, _ptr(out))
        return out
