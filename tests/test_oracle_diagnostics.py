"""Pins of the diagnostics oracle (conserved totals, extrema, errors against a
second state, the pressure positivity monitor) by closed forms, by the
pressure of the golden-pinned NS closure, and by planted single-cell cases
whose answers are known by hand."""

import math

import numpy as np
import pytest

from oracle import diagnostics as od
from oracle import grid as og
from oracle import navier_stokes as ons

NS = ("mass_density", "momentum_x1", "momentum_x2", "momentum_x3", "energy_density")


def _cfg(n, lo, hi):
    return {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0], "x2_min": lo[1], "x2_max": hi[1],
            "nx2": n[1], "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}


def _state(n, rng, names=NS):
    shp = tuple(x + 2 for x in n)
    U = {k: rng.uniform(-1.0, 1.0, shp) for k in names}
    if "mass_density" in U:
        U["mass_density"] = rng.uniform(0.5, 2.0, shp)
        U["energy_density"] = rng.uniform(4.0, 6.0, shp)
    return U


@pytest.mark.parametrize("coord,lo,hi,n,closed", [
    ("cartesian", (0.0, -0.3, 0.1), (0.7, 1.1, 0.6), (6, 7, 5), 0.7 * 1.4 * 0.5),
    ("cylindrical", (0.1, 0.0, -1.0), (2.0, 2 * np.pi, 1.0), (16, 12, 8),
     (2.0 ** 2 - 0.1 ** 2) / 2 * 2 * np.pi * 2.0),
    ("spherical", (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi), (16, 12, 10),
     (1.0 - 0.1 ** 3) / 3 * (np.cos(0.05) - np.cos(np.pi - 0.05)) * 2 * np.pi),
])
def test_volume_and_constant_field_totals_match_closed_forms(coord, lo, hi, n, closed):
    g = og.build_grid(coord, _cfg(n, lo, hi))
    U = {"T": np.full(tuple(x + 2 for x in n), 2.5)}
    d = od.diagnostics(U, ("T",), g)
    assert d["volume"] == pytest.approx(closed, rel=1e-13)
    assert d["total"]["T"] == pytest.approx(2.5 * closed, rel=1e-13)
    assert d["min"]["T"] == d["max"]["T"] == 2.5
    assert d["l1"] is None and d["pressure_min"] is None


def test_linear_field_total_is_its_integral():
    """Midpoint rule is exact for a linear field: sum u V = Vol * u(domain centre)."""
    n, lo, hi = (8, 6, 5), (0.0, -1.0, 0.5), (2.0, 1.0, 1.5)
    g = og.build_grid("cartesian", _cfg(n, lo, hi))
    x1, x2, x3 = (g["cell"]["coordinates"][k] for k in ("x1", "x2", "x3"))
    u = np.zeros(tuple(x + 2 for x in n))
    u[1:-1, 1:-1, 1:-1] = 3.0 + 2.0 * x1 - 1.5 * x2 + 4.0 * x3
    d = od.diagnostics({"u": u}, ("u",), g)
    vol = 2.0 * 2.0 * 1.0
    centre = 3.0 + 2.0 * 1.0 - 1.5 * 0.0 + 4.0 * 1.0
    assert d["total"]["u"] == pytest.approx(vol * centre, rel=1e-13)
    assert d["min"]["u"] == float(np.min(u[1:-1, 1:-1, 1:-1]))


def test_pressure_equals_the_ns_closure_pressure():
    rng = np.random.default_rng(4)
    n = (6, 5, 4)
    g = og.build_grid("cartesian", _cfg(n, (0, 0, 0), (1, 1, 1)))
    U = _state(n, rng)
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 0.0,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    C = ons.compute_closure(U, params, g)
    P_ref = np.asarray(C["P"])
    if P_ref.shape != tuple(n):
        P_ref = P_ref[1:-1, 1:-1, 1:-1]
    P = od.pressure(U, 1.4)
    assert np.array_equal(P, P_ref)
    d = od.diagnostics(U, NS, g, gamma=1.4)
    assert d["pressure_min"] == float(np.min(P_ref))
    assert d["nonpositive_cells"] == 0


def test_planted_violations_are_counted():
    rng = np.random.default_rng(5)
    n = (5, 4, 3)
    g = og.build_grid("cartesian", _cfg(n, (0, 0, 0), (1, 1, 1)))
    U = _state(n, rng)
    U["energy_density"][2, 2, 2] = 0.0        # P < 0 there
    U["mass_density"][3, 1, 1] = -1.0         # rho < 0 (P > 0 as well: both count once)
    U["mass_density"][0, 2, 2] = -5.0         # a ghost cell: not counted
    U["momentum_x1"][1, 3, 2] = np.nan        # NaN pressure
    d = od.diagnostics(U, NS, g, gamma=1.4)
    assert d["nonpositive_cells"] == 3
    assert math.isnan(d["pressure_min"])
    assert math.isnan(d["min"]["momentum_x1"]) and math.isnan(d["total"]["momentum_x1"])


def test_single_cell_difference_gives_known_l1_and_linf():
    n = (4, 3, 2)
    g = og.build_grid("cylindrical", _cfg(n, (0.5, 0.0, 0.0), (1.5, np.pi, 1.0)))
    rng = np.random.default_rng(6)
    U = {"a": rng.uniform(-1, 1, (6, 5, 4)), "b": rng.uniform(-1, 1, (6, 5, 4))}
    ref = {k: v.copy() for k, v in U.items()}
    ref["b"][2, 3, 1] += 0.25        # interior cell (i, j, k) = (1, 2, 0)
    ref["a"][5, 1, 1] += 10.0        # a ghost cell: ignored
    d = od.diagnostics(U, ("a", "b"), g, ref=ref)
    V = g["cell"]["geometry"]["volume"]
    assert d["linf"]["a"] == 0.0 and d["l1"]["a"] == 0.0
    e = abs(U["b"][2, 3, 1] - ref["b"][2, 3, 1])
    assert d["linf"]["b"] == e
    assert d["l1"]["b"] == pytest.approx(e * V[1, 2, 0] / d["volume"], rel=1e-15)
