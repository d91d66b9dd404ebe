"""Oracle of the state diagnostics (conserved totals, extrema, positivity
monitor, errors against a second state).

Test infrastructure only (see ``oracle/__init__.py``).

The reference has no run-time monitor, so this module defines what the GPU
reductions compute, over the interior cells of a reference-format state
(a dict of ghosted (nx1+2, nx2+2, nx3+2) float64 arrays):

* totals: sum of U * V, with V the cell volumes of ``build_grid``
  (geometry.py:14-194); summed exactly (``math.fsum``) of the rounded
  products, so the oracle value is the correctly rounded total;
* min / max per field;
* L1: sum |U - ref| * V / sum V, Linf: max |U - ref|;
* the ideal-gas pressure of ``compute_closure`` (navier_stokes.py:94-100),
  its minimum, and the number of cells with rho <= 0 or P <= 0.

Pinned (``tests/test_oracle_diagnostics.py``) by the closed-form volumes of
Cartesian, cylindrical and spherical domains, the closed-form total of a
constant and of a linear field (the midpoint rule is exact for linear
functions), the pressure against ``oracle.navier_stokes.compute_closure``
(itself pinned by the reference's golden vectors), and planted single-cell
differences and violations whose L1, Linf and counts are known by hand.
"""

import math

import numpy as np

_IN = (slice(1, -1), slice(1, -1), slice(1, -1))


def pressure(U, gamma):
    """navier_stokes.py:94-100 -- P over the interior cells."""
    rho = U["mass_density"][_IN]
    vx1 = U["momentum_x1"][_IN] / rho
    vx2 = U["momentum_x2"][_IN] / rho
    vx3 = U["momentum_x3"][_IN] / rho
    E = U["energy_density"][_IN] / rho
    return (gamma - 1.0) * rho * (E - (vx1 ** 2 + vx2 ** 2 + vx3 ** 2) / 2.0)


def diagnostics(U, names, grid, ref=None, gamma=None):
    """Reductions of the interior of ``U`` (fields ``names``) on ``grid``."""
    V = np.asarray(grid["cell"]["geometry"]["volume"], dtype=np.float64)
    vol = math.fsum(V.ravel())
    out = {"total": {}, "min": {}, "max": {}, "l1": None, "linf": None, "volume": vol,
           "pressure_min": None, "nonpositive_cells": None}
    if ref is not None:
        out["l1"], out["linf"] = {}, {}
    for name in names:
        u = np.asarray(U[name], dtype=np.float64)[_IN]
        out["total"][name] = math.fsum((u * V).ravel())
        out["min"][name] = float(np.min(u))
        out["max"][name] = float(np.max(u))
        if ref is not None:
            e = np.abs(u - np.asarray(ref[name], dtype=np.float64)[_IN])
            out["l1"][name] = math.fsum((e * V).ravel()) / vol
            out["linf"][name] = float(np.max(e))
    if gamma is not None:
        P = pressure(U, gamma)
        rho = np.asarray(U["mass_density"], dtype=np.float64)[_IN]
        out["pressure_min"] = float(np.min(P))
        out["nonpositive_cells"] = int(np.count_nonzero(~((rho > 0) & (P > 0))))
    return out
