"""Per-field error norms between two states on one chart (equal spacing).

L1 is the volume-weighted mean of |a - b| (the integral over the box divided
by its volume), Linf the maximum; both also relative to the mean and maximum
of |b|.
"""
from __future__ import annotations

import numpy as np

CONSERVED = ("D", "S_1", "S_2", "S_3", "G")
PRIMITIVE = ("N", "V_1", "V_2", "V_3", "E")


def field_errors(a: np.ndarray, b: np.ndarray, names=CONSERVED) -> dict:
    """{name: {"L1", "Linf", "L1_rel", "Linf_rel"}} for five-field states
    a (tested) and b (reference) of shape (5, n2, n1, n0)."""
    out = {}
    for f, name in enumerate(names):
        d = np.abs(np.asarray(a[f], dtype=np.float64) - np.asarray(b[f], dtype=np.float64))
        ref = np.abs(np.asarray(b[f], dtype=np.float64))
        l1, linf = float(d.mean()), float(d.max())
        mref, xref = float(ref.mean()), float(ref.max())
        out[name] = {
            "L1": l1,
            "Linf": linf,
            "L1_rel": l1 / mref if mref > 0 else float("nan"),
            "Linf_rel": linf / xref if xref > 0 else float("nan"),
        }
    return out
