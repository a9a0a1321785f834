"""ctypes binding of libgx_fluid.so (the C-ABI declared in include/gx_fluid.h).

The library is loaded from genasis_amd/lib; a missing library raises -- there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgx_fluid.so")

GX_OK = 0
GX_BC_PERIODIC, GX_BC_OUTFLOW, GX_BC_REFLECTING = 0, 1, 2
GX_HLL, GX_HLLC = 0, 1
GX_FP64, GX_FP32 = 0, 1
GX_N_FIELDS = 23

FIELD_NAMES = (
    "FastEigenspeedPlus_1", "FastEigenspeedPlus_2", "FastEigenspeedPlus_3",
    "FastEigenspeedMinus_1", "FastEigenspeedMinus_2", "FastEigenspeedMinus_3",
    "ComovingBaryonDensity", "ConservedBaryonDensity", "BaryonMass",
    "Velocity_1", "Velocity_2", "Velocity_3",
    "MomentumDensity_1", "MomentumDensity_2", "MomentumDensity_3",
    "InternalEnergy", "ConservedEnergy", "Pressure", "Temperature",
    "EntropyPerBaryon", "ConservedEntropy", "SoundSpeed", "MachNumber",
)

GX_LINEAR, GX_PARABOLIC = 0, 1
GX_RESTART_PRIMITIVE = 0
GX_RESTART_EXACT = 1

# Conservation tallies (Tally_F_P % Value order, 0-based)
TALLY_NAMES = (
    "BaryonNumber", "Momentum_1", "Momentum_2", "Momentum_3", "KineticEnergy",
    "AngularMomentum_1", "AngularMomentum_2", "AngularMomentum_3",
    "GravitationalEnergy", "TotalEnergy", "FluidEnergy", "InternalEnergy",
)
GX_TALLY_N = len(TALLY_NAMES)

# Every symbol include/gx_fluid.h declares.
EXPORTS = (
    "gx_default_params", "gx_create", "gx_destroy", "gx_create_bricks", "gx_get_bricks", "gx_upload", "gx_download",
    "gx_upload_features", "gx_download_features", "gx_compute_time_step",
    "gx_compute_time_step_noalloc",
    "gx_step_rk2", "gx_sync", "gx_totals", "gx_set_state", "gx_set_primitive",
    "gx_get_conserved", "gx_get_primitive", "gx_get_flags", "gx_set_flags",
    "gx_set_tuning", "gx_get_tuning", "gx_set_tile", "gx_get_tile", "gx_get_geometry_pow2",
    "gx_set_record_sources", "gx_get_sources",
    "gx_stream", "gx_set_profiling", "gx_kernel_times",
    "gx_self_test_division", "gx_last_error",
    "gx_get_boundary_fluence", "gx_compute_tally",
    "gx_write_checkpoint", "gx_read_checkpoint",
    "gx_batch_create", "gx_batch_destroy", "gx_batch_set_primitive",
    "gx_batch_set_integrator", "gx_batch_step", "gx_batch_sync", "gx_batch_cfl",
    "gx_batch_times", "gx_batch_get_conserved", "gx_batch_get_flags",
    "gx_batch_set_tuning", "gx_batch_get_tuning", "gx_batch_set_tiles",
    "gx_batch_get_tiles", "gx_batch_autotune",
    "gx_batch_cells", "gx_batch_stream",
    "gx_adjoint_forward", "gx_adjoint_backward", "gx_adjoint_kinetic_energy",
    "gx_adjoint_get_bar", "gx_adjoint_face_vjp", "gx_adjoint_slope_vjp",
)


class GxGrid(ctypes.Structure):
    _fields_ = [
        ("nCells", ctypes.c_int * 3),
        ("MinCoordinate", ctypes.c_double * 3),
        ("MaxCoordinate", ctypes.c_double * 3),
        ("BoundaryCondition", ctypes.c_int * 6),
    ]


class GxParams(ctypes.Structure):
    _fields_ = [
        ("AdiabaticIndex", ctypes.c_double),
        ("LimiterParameter", ctypes.c_double),
        ("ShockThreshold", ctypes.c_double),
        ("CourantFactor", ctypes.c_double),
        ("BaryonDensityMin", ctypes.c_double),
        ("BaryonMassReference", ctypes.c_double),
        ("UseLimiter", ctypes.c_int),
        ("RiemannSolverType", ctypes.c_int),
        ("Precision", ctypes.c_int),
        ("ReconstructionType", ctypes.c_int),
    ]


class GxError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load the library (once) and declare the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GxError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp, d, i, ll = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_longlong
    pd = P(ctypes.c_double)
    sig = {
        "gx_default_params": (None, [P(GxParams)]),
        "gx_create": (i, [P(GxGrid), P(GxParams), P(vp)]),
        "gx_destroy": (i, [vp]),
        "gx_create_bricks": (i, [P(GxGrid), P(GxParams), P(ctypes.c_int), P(vp)]),
        "gx_get_bricks": (i, [vp, P(ctypes.c_int)]),
        "gx_upload": (i, [vp, pd, ll, P(ctypes.c_int)]),
        "gx_download": (i, [vp, pd, ll, P(ctypes.c_int)]),
        "gx_upload_features": (i, [vp, pd, ll]),
        "gx_download_features": (i, [vp, pd, ll]),
        "gx_compute_time_step": (i, [vp, pd]),
        "gx_compute_time_step_noalloc": (i, [P(GxGrid), P(GxParams), vp, vp, vp]),
        "gx_step_rk2": (i, [vp, d, d]),
        "gx_sync": (i, [vp]),
        "gx_totals": (i, [vp, pd]),
        "gx_get_boundary_fluence": (i, [vp, pd]),
        "gx_compute_tally": (i, [vp, i, pd, pd, pd, pd]),
        "gx_write_checkpoint": (i, [vp, ctypes.c_char_p, d, ll, i]),
        "gx_read_checkpoint": (i, [vp, ctypes.c_char_p, i, pd, P(ctypes.c_longlong),
                                   P(ctypes.c_int)]),
        "gx_set_state": (i, [vp, pd, pd]),
        "gx_set_primitive": (i, [vp, pd]),
        "gx_get_conserved": (i, [vp, pd]),
        "gx_get_primitive": (i, [vp, pd]),
        "gx_get_flags": (i, [vp, P(ctypes.c_ubyte)]),
        "gx_set_flags": (i, [vp, P(ctypes.c_ubyte)]),
        "gx_set_tuning": (i, [vp, i]),
        "gx_get_tuning": (i, [vp, P(ctypes.c_int)]),
        "gx_set_tile": (i, [vp, i, i]),
        "gx_get_tile": (i, [vp, P(ctypes.c_int), P(ctypes.c_int)]),
        "gx_get_geometry_pow2": (i, [vp, P(ctypes.c_int)]),
        "gx_set_record_sources": (i, [vp, i]),
        "gx_get_sources": (i, [vp, pd]),
        "gx_stream": (vp, [vp]),
        "gx_set_profiling": (i, [vp, i]),
        "gx_kernel_times": (i, [vp, pd, P(ctypes.c_longlong)]),
        "gx_self_test_division": (i, [ll, ctypes.c_ulonglong, P(ctypes.c_longlong)]),
        "gx_last_error": (ctypes.c_char_p, []),
        "gx_batch_create": (i, [i, P(GxGrid), P(GxParams), P(vp)]),
        "gx_batch_destroy": (i, [vp]),
        "gx_batch_set_primitive": (i, [vp, i, pd]),
        "gx_batch_set_integrator": (i, [vp, d, i, i]),
        "gx_batch_step": (i, [vp, i]),
        "gx_batch_sync": (i, [vp]),
        "gx_batch_cfl": (i, [vp, pd]),
        "gx_batch_times": (i, [vp, pd, pd, P(ctypes.c_int)]),
        "gx_batch_get_conserved": (i, [vp, i, pd]),
        "gx_batch_get_flags": (i, [vp, i, P(ctypes.c_ubyte)]),
        "gx_batch_set_tuning": (i, [vp, i, P(ctypes.c_int), P(ctypes.c_int), i]),
        "gx_batch_get_tuning": (i, [vp, i, P(ctypes.c_int), P(ctypes.c_int),
                                    P(ctypes.c_int)]),
        "gx_batch_set_tiles": (i, [vp, i, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
                                   P(ctypes.c_int), i]),
        "gx_batch_get_tiles": (i, [vp, i, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
                                   P(ctypes.c_int), P(ctypes.c_int)]),
        "gx_batch_autotune": (i, [vp, i]),
        "gx_batch_cells": (ll, [vp]),
        "gx_batch_stream": (vp, [vp]),
        "gx_adjoint_forward": (i, [vp, i, i, i]),
        "gx_adjoint_backward": (i, [vp, pd, pd, pd]),
        "gx_adjoint_kinetic_energy": (i, [vp, i, d, d, pd, pd]),
        "gx_adjoint_get_bar": (i, [vp, pd]),
        "gx_adjoint_face_vjp": (None, [pd, pd, i, i, d, d, i, pd, pd, pd, pd, pd]),
        "gx_adjoint_slope_vjp": (None, [d, d, d, d, d, d, i, d, pd]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != GX_OK:
        msg = _lib.gx_last_error().decode() if _lib is not None else ""
        raise GxError(f"{what} failed ({rc}): {msg}")
