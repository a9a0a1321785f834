"""Single-level rectangular chart: cell layout and 1-D geometry.

Restates the geometry that the GenASiS chart builds for a GALILEAN,
RECTANGULAR, EQUAL-spacing chart:

* ``SetEdgeEqual`` (Modules/Mathematics/Manifolds/Atlases/Charts/ChartBasics/
  Chart_Template.f90:699-723): edges by repeated addition of one width;
* ghost edges (Chart_Template.f90:541-549), whose widths are held in variables
  declared ``real(KDL)`` (Chart_Template.f90:477-478) -- KDL is the default
  logical kind, 4 -- so each ghost width is rounded to single precision;
* ``SetCenterRectangular`` and ``SetWidthLeftRight`` (Chart_Template.f90:
  827-841, 985-987).

Boundary conditions follow the reference names (``'PERIODIC'``,
``'OUTFLOW'``, ``'REFLECTING'``, Modules/Mathematics/Manifolds/Atlases/
Intercharts/Atlas_SC__Template.f90:244-324).
"""
from __future__ import annotations

import numpy as np

N_GHOST = 2  # nGhostLayers (ChartHeader_SL__Form.f90:362)

BC_CODES = {"PERIODIC": 0, "OUTFLOW": 1, "REFLECTING": 2}


def _edges(n: int, lo: float, hi: float) -> np.ndarray:
    """Edge(1-g : n+g+1) as a numpy array of n + 2g + 1 values."""
    g = N_GHOST
    e = np.zeros(n + 2 * g + 1)
    off = g - 1  # Edge(i) -> e[i + off]
    e[1 + off] = lo
    width = (hi - lo) / n
    for ic in range(2, n + 2):
        e[ic + off] = e[ic - 1 + off] + width
    for ic in range(1, g + 1):
        w_ig = float(np.float32(e[ic + 1 + off] - e[ic + off]))
        w_og = float(np.float32(e[n - ic + 2 + off] - e[n - ic + 1 + off]))
        e[1 - ic + off] = e[2 - ic + off] - w_ig
        e[n + 1 + ic + off] = e[n + ic + off] + w_og
    return e


class Chart:
    """A 3-D single-brick chart of ``n`` proper cells per dimension.

    ``bc`` is a sequence of three (inner, outer) pairs of boundary names; a
    periodic dimension must be periodic on both faces.
    """

    def __init__(self, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                 bc=(("PERIODIC", "PERIODIC"),) * 3):
        self.n = tuple(int(v) for v in n)
        if len(self.n) != 3 or min(self.n) < 1:
            raise ValueError("nCells must hold three positive sizes")
        self.lo = tuple(float(v) for v in lo)
        self.hi = tuple(float(v) for v in hi)
        self.bc = tuple((str(a).upper(), str(b).upper()) for a, b in bc)
        for d, (a, b) in enumerate(self.bc):
            if a not in BC_CODES or b not in BC_CODES:
                raise ValueError(f"boundary condition not recognized: {a}, {b}")
            if (a == "PERIODIC") != (b == "PERIODIC"):
                raise ValueError(f"dimension {d}: PERIODIC must hold on both faces")
        self.g = N_GHOST
        self.m = tuple(v + 2 * self.g for v in self.n)
        self.edge, self.center, self.width_left, self.width_right = [], [], [], []
        for d in range(3):
            e = _edges(self.n[d], self.lo[d], self.hi[d])
            c = 0.5 * (e[:-1] + e[1:])
            self.edge.append(e)
            self.center.append(c)
            self.width_left.append(c - e[:-1])
            self.width_right.append(e[1:] - c)

    @property
    def bc_codes(self):
        return [BC_CODES[self.bc[d][s]] for d in range(3) for s in range(2)]

    @property
    def n_proper(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def n_values(self) -> int:
        return self.m[0] * self.m[1] * self.m[2]

    def padded_centers(self):
        """Cell centres as three (m2, m1, m0) arrays over the padded layout."""
        z, y, x = np.meshgrid(self.center[2], self.center[1], self.center[0],
                              indexing="ij")
        return x, y, z

    def proper(self, a: np.ndarray) -> np.ndarray:
        """The proper-cell block of a padded (m2, m1, m0) array."""
        g = self.g
        return a[g:-g, g:-g, g:-g]
