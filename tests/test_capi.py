"""The C-ABI library loads and exports every entry point include/gx_fluid.h
declares; host-side argument checks run without a GPU."""
import ctypes
import os
import re

import pytest

from genasis_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "gx_fluid.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gx_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert _declared() == sorted(_capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_default_params_and_argument_errors():
    lib = _capi.load()
    p = _capi.GxParams()
    lib.gx_default_params(ctypes.byref(p))
    assert (p.AdiabaticIndex, p.LimiterParameter, p.ShockThreshold, p.CourantFactor,
            p.BaryonDensityMin, p.BaryonMassReference, p.UseLimiter,
            p.RiemannSolverType, p.Precision) == (1.4, 1.4, 0.1, 0.7, 0.0, 1.0, 1,
                                                  _capi.GX_HLLC, _capi.GX_FP64)
    g = _capi.GxGrid()
    for d in range(3):
        g.nCells[d] = 1  # too small: rejected before any device call
        g.MaxCoordinate[d] = 1.0
    h = ctypes.c_void_p()
    rc = lib.gx_create(ctypes.byref(g), ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and b"nCells" in lib.gx_last_error()
    for d in range(3):
        g.nCells[d] = 8
    g.BoundaryCondition[0] = _capi.GX_BC_PERIODIC
    g.BoundaryCondition[1] = _capi.GX_BC_OUTFLOW
    rc = lib.gx_create(ctypes.byref(g), ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and b"PERIODIC" in lib.gx_last_error()
    g.BoundaryCondition[1] = _capi.GX_BC_PERIODIC
    p.Precision = 2
    rc = lib.gx_create(ctypes.byref(g), ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and b"Precision" in lib.gx_last_error()
    p.Precision = _capi.GX_FP64
    # spacings outside the range of the exactly rounded reciprocal division
    for big in (1.0e40, 1.0e-160):
        for d in range(3):
            g.MaxCoordinate[d] = big
        rc = lib.gx_create(ctypes.byref(g), ctypes.byref(p), ctypes.byref(h))
        assert rc == -1 and b"spacings" in lib.gx_last_error(), big
    assert lib.gx_step_rk2(None, 0.0, 1.0) == -1
    assert lib.gx_set_record_sources(None, 1) == -1
    assert lib.gx_get_sources(None, None) == -1 and b"gx_get_sources" in lib.gx_last_error()


def test_brick_arguments_are_checked():
    """nBricks must divide nCells and leave two cells per brick (host-side
    checks, before any device work)."""
    lib = _capi.load()
    g = _capi.GxGrid()
    for d in range(3):
        g.nCells[d] = 12
        g.MinCoordinate[d] = 0.0
        g.MaxCoordinate[d] = 1.0
    for i in range(6):
        g.BoundaryCondition[i] = _capi.GX_BC_OUTFLOW
    h = ctypes.c_void_p()
    for nb in ((5, 1, 1), (1, 12, 1), (0, 1, 1)):
        arr = (ctypes.c_int * 3)(*nb)
        assert lib.gx_create_bricks(ctypes.byref(g), None, arr, ctypes.byref(h)) == -1, nb
        assert b"nBricks" in lib.gx_last_error() or b"2 cells" in lib.gx_last_error()
